"""Pins the oracle (oracle/numpy_oracle.py, oracle/brute_wmc.c) against what the mathematics
fixes: closed-form counts of small graphs, brute-force enumeration, exact big-integer and
rational arithmetic, scaling identities and slicing invariance.  Planted mistakes in the
oracle must be caught by these checks."""

import itertools
import math
import random
from fractions import Fraction

import numpy as np
import pytest

import graphs
from oracle import brute, numpy_oracle
from tensororder_amd import network, plan_io, planners
from tensororder_amd.formula import parse_dimacs_text
from tensororder_amd.instances import random_3cnf, vertex_cover_cnf
from tensororder_amd.plan import ALL_SLICERS


def _plan_dict(formula, planner, nslice=0, seed=0, max_rank=22):
    net = network.cnf_count(formula)
    plan = planners.best_plan(net, planner, solver=None, seed=seed)
    assert plan.maxrank <= max_rank, "test plan too large for the CPU oracle"
    if nslice:
        ALL_SLICERS["greedy_mem"].slice_until(plan, slices=nslice)
    return plan_io.plan_to_dict(plan)


def _weights(formula):
    return {v: formula._variables[v] for v in formula.variables}


# -- leaf values -----------------------------------------------------------------------

@pytest.mark.parametrize("rank", range(0, 6))
def test_or_leaf_truth_table(rank):
    for pos in itertools.product((True, False), repeat=rank):
        for o in [None] + list(range(rank)):
            t = numpy_oracle.build_leaf({"kind": "or", "literals_positive": list(pos), "output_index": o})
            for idx in itertools.product((0, 1), repeat=rank):
                lits = [v if p else 1 - v for v, p in zip(idx, pos)]
                if o is None:
                    want = 1 if any(lits) else 0
                else:
                    want = 1 if lits[o] == int(any(lits[i] for i in range(rank) if i != o)) else 0
                assert t[idx] == want


@pytest.mark.parametrize("rank", range(0, 6))
def test_var_leaf(rank):
    t = numpy_oracle.build_leaf({"kind": "var", "rank": rank, "w_neg": 0.3, "w_pos": 0.7})
    if rank == 0:
        assert t[()] == pytest.approx(1.0)
        return
    for idx in itertools.product((0, 1), repeat=rank):
        want = 0.3 if all(i == 0 for i in idx) else (0.7 if all(i == 1 for i in idx) else 0.0)
        assert t[idx] == want


# -- tensordot semantics -------------------------------------------------------------

def test_tensordot_matches_einsum():
    rng = np.random.default_rng(0)
    letters = "abcdefghijklmnop"
    for _ in range(60):
        ra, rb = rng.integers(0, 6), rng.integers(0, 6)
        s = int(rng.integers(0, min(ra, rb) + 1))
        la = list(rng.permutation(ra)[:s])
        lb = list(rng.permutation(rb)[:s])
        a = rng.standard_normal((2,) * ra)
        b = rng.standard_normal((2,) * rb)
        sa = list(letters[:ra])
        sb = list(letters[8:8 + rb])
        for i, j in zip(la, lb):
            sb[j] = sa[i]
        out = [c for i, c in enumerate(sa) if i not in la] + [c for j, c in enumerate(sb) if j not in lb]
        want = np.einsum("%s,%s->%s" % ("".join(sa), "".join(sb), "".join(out)), a, b)
        got = numpy_oracle.tensordot(a, b, (la, lb))
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_join_maps_order():
    # left-unshared edges in order, then right-unshared; pairs in left-axis order
    lmap, rmap, free = numpy_oracle.join_maps([5, 1, 7, 3], [3, 9, 5, 2])
    assert lmap == [0, 3] and rmap == [2, 0] and free == [1, 7, 9, 2]


# -- brute force ---------------------------------------------------------------------

def test_brute_c_matches_python():
    for seed in range(5):
        f = parse_dimacs_text(random_3cnf(10, 25, seed), "cachet")
        w = _weights(f)
        assert brute.brute_wmc(10, f.clauses, w) == pytest.approx(brute.brute_wmc_py(10, f.clauses, w), rel=1e-13)


# -- closed forms (every planner, exact integers) ----------------------------------------

CLOSED = [
    ("K4", graphs.complete(4), 5),
    ("K33", graphs.complete_bipartite(3, 3), 15),
    ("prism", graphs.prism(), 13),
    ("Q3", graphs.cube(), 35),
    ("Petersen", graphs.petersen(), 76),
    ("C9", graphs.cycle(9), graphs.lucas(9)),
    ("C12", graphs.cycle(12), graphs.lucas(12)),
    ("P10", graphs.path(10), graphs.fibonacci(12)),
]


@pytest.mark.parametrize("name,graph,count", CLOSED, ids=[c[0] for c in CLOSED])
@pytest.mark.parametrize("planner", ["line", "factor"])
def test_closed_form_counts(name, graph, count, planner):
    n, edges = graph
    assert graphs.count_independent_sets(n, edges) == count
    f = graphs.vc_formula(n, edges)
    d = _plan_dict(f, planner)
    exact = numpy_oracle.contract_sliced(d, "bigint")
    assert isinstance(exact, int) and exact == count
    assert numpy_oracle.contract_sliced(d) == count
    d2 = _plan_dict(f, planner, nslice=2)
    assert numpy_oracle.contract_sliced(d2, "bigint") == count


# -- weighted counts vs brute force, sliced and unsliced --------------------------------

@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("planner", ["line", "factor"])
def test_weighted_vs_brute(seed, planner):
    # line graphs of dense 3-CNFs are wide, so the line planner gets vertex cover
    if planner == "line":
        f = parse_dimacs_text(vertex_cover_cnf(18, seed, True), "cachet")
    else:
        f = parse_dimacs_text(random_3cnf(14, 40, seed), "cachet")
    want = brute.brute_wmc(len(f.variables), f.clauses, _weights(f))
    for nslice in (0, 3):
        d = _plan_dict(f, planner, nslice=nslice, seed=seed)
        assert numpy_oracle.contract_sliced(d) == pytest.approx(want, rel=1e-12)


def test_vertex_cover_vs_brute():
    f = parse_dimacs_text(vertex_cover_cnf(20, 3, True), "cachet")
    want = brute.brute_wmc(20, f.clauses, _weights(f))
    for planner in ("line", "factor"):
        d = _plan_dict(f, planner, nslice=2)
        assert numpy_oracle.contract_sliced(d) == pytest.approx(want, rel=1e-12)


def test_exact_rational_count():
    f = parse_dimacs_text(random_3cnf(12, 30, 7), "cachet")
    d = _plan_dict(f, "factor", nslice=2)
    exact = numpy_oracle.contract_sliced(d, "bigint", exact=True)
    assert isinstance(exact, Fraction)
    # enumerate exactly with the same binary weights
    total = Fraction(0)
    for bits in itertools.product((0, 1), repeat=12):
        if all(any((bits[abs(l) - 1] == 1) == (l > 0) for l in c) for c in f.clauses):
            w = Fraction(1)
            for v in range(12):
                w *= Fraction(f._variables[v + 1][bits[v]])
            total += w
    assert exact == total
    assert float(numpy_oracle.contract_sliced(d)) == pytest.approx(float(total), rel=1e-13)


def test_weight_scaling_identity():
    # doubling both literal weights of every variable multiplies the count by 2^n
    f = parse_dimacs_text(random_3cnf(12, 30, 3), "cachet")
    base = numpy_oracle.contract_sliced(_plan_dict(f, "factor"))
    for v in f.variables:
        w0, w1 = f._variables[v]
        f.set_variable_weight(v, 2 * w0, 2 * w1)
    assert numpy_oracle.contract_sliced(_plan_dict(f, "factor")) == pytest.approx(base * 2 ** 12, rel=1e-13)


def test_slice_sum_equals_unsliced_and_prefix():
    f = parse_dimacs_text(vertex_cover_cnf(16, 5, True), "cachet")
    d0 = _plan_dict(f, "factor")
    d4 = _plan_dict(f, "factor", nslice=4)
    full = numpy_oracle.contract_sliced(d0)
    assert len(d4["groups_to_slice"]) == 4
    assert numpy_oracle.contract_sliced(d4) == pytest.approx(full, rel=1e-13)
    vals = numpy_oracle.slice_values(d4, list(itertools.product((0, 1), repeat=4)))
    assert sum(vals) == pytest.approx(full, rel=1e-13)
    assert numpy_oracle.contract_sliced(d4, num_slice_limit=5) == pytest.approx(sum(vals[:5]), rel=1e-13)


# -- planted mistakes are caught -------------------------------------------------------

def test_planted_mistake_in_edge_maps_is_caught(monkeypatch):
    f = parse_dimacs_text(random_3cnf(12, 34, 1), "cachet")
    want = brute.brute_wmc(12, f.clauses, _weights(f))
    d = _plan_dict(f, "factor")
    real = numpy_oracle.join_maps

    def reversed_right(lf, rf):
        lmap, rmap, free = real(lf, rf)
        return lmap, rmap[::-1], free

    monkeypatch.setattr(numpy_oracle, "join_maps", reversed_right)
    bad = numpy_oracle.contract_sliced(d)
    assert bad != pytest.approx(want, rel=1e-9)


def test_planted_mistake_in_or_leaf_is_caught(monkeypatch):
    f = parse_dimacs_text(random_3cnf(12, 34, 2), "cachet")
    want = brute.brute_wmc(12, f.clauses, _weights(f))
    d = _plan_dict(f, "factor")
    real = numpy_oracle.build_leaf

    def sloppy(desc, dtype=np.float64, exact=False):
        out = real(desc, dtype, exact)
        if desc["kind"] == "or" and desc["output_index"] is None:
            out[...] = 1  # forgets the all-false zero
        return out

    monkeypatch.setattr(numpy_oracle, "build_leaf", sloppy)
    assert numpy_oracle.contract_sliced(d) != pytest.approx(want, rel=1e-9)


def test_planted_mistake_in_slice_value_is_caught(monkeypatch):
    f = parse_dimacs_text(vertex_cover_cnf(16, 2, True), "cachet")
    d0 = _plan_dict(f, "factor")
    d3 = _plan_dict(f, "factor", nslice=3)
    full = numpy_oracle.contract_sliced(d0)
    real = numpy_oracle.slice_leaves

    def always_zero(plan, leaves, assignment, groups=None):
        return real(plan, leaves, tuple(0 for _ in assignment), groups)

    monkeypatch.setattr(numpy_oracle, "slice_leaves", always_zero)
    assert numpy_oracle.contract_sliced(d3) != pytest.approx(full, rel=1e-9)

"""ORACLE -- test infrastructure only.  Nothing in the product imports this module.

A CPU restatement of the reference's numpy execution path, working directly from a plan
document (``tensororder-plan/1`` JSON dict), written independently of the package's
host structures so that both can be checked against each other:

* leaf values: ``tensor_network/tensor_network_constructions.py:69-99`` (OR gate,
  optional output index) and ``:144-152`` (diagonal variable tensor);
  dense leaves as ``tensor.py:43-49``;
* per-slice leaf values keep each sliced axis with size 1
  (``tensor.py:20-25`` and ``:70-73``), groups enumerated in ``itertools.product``
  order (``tensor_network.pyx:363-401``);
* join edge maps and free-edge order: ``contraction_methods/contraction_tree.pyx:248-288``;
* post-order walk calling ``numpy.tensordot``: ``tensor_network.pyx:142-156`` with
  ``tensor_apis/numpy_apis.py:42-43``;
* slice sum from an integer 0 in product order, optionally truncated:
  ``tensor_apis/base_api.py:17-28``.

Entry types follow ``numpy_apis.py:13-27``: ``float64`` (default), ``float32``,
``int64``, ``uint64`` and ``bigint`` (``object`` arrays of Python ints, or of
``fractions.Fraction`` for exact weighted counts when ``exact=True``).
"""

import itertools
from fractions import Fraction

import numpy as np

DTYPES = {
    "float64": np.float64,
    "float32": np.float32,
    "float16": np.float16,
    "int": np.int64,
    "uint": np.uint64,
    "bigint": object,
}


def _weight(x, dtype, exact):
    if dtype is object:
        return Fraction(x) if exact else (int(x) if float(x).is_integer() else x)
    return x


def build_leaf(desc, dtype=np.float64, exact=False):
    """Dense C-ordered value of one leaf descriptor."""
    kind = desc["kind"]
    if kind == "var":
        r = desc["rank"]
        w0 = _weight(desc["w_neg"], dtype, exact)
        w1 = _weight(desc["w_pos"], dtype, exact)
        out = np.full((2,) * r, 0, dtype=dtype)
        if r == 0:
            out[()] = w0 + w1
        else:
            out[(0,) * r] = w0
            out[(1,) * r] = w1
        return out
    if kind == "or":
        pos = desc["literals_positive"]
        r = len(pos)
        o = desc["output_index"]
        out = np.full((2,) * r, 1, dtype=dtype)
        false_pt = tuple(0 if p else 1 for p in pos)
        if o is None:
            out[false_pt] = 0
        else:
            # entry is 1 iff literal o == OR(other literals); literal value of index
            # value v is v for a positive literal and 1-v for a negative one
            for idx in itertools.product((0, 1), repeat=r):
                lits = [v if p else 1 - v for v, p in zip(idx, pos)]
                others = any(lits[i] for i in range(r) if i != o)
                out[idx] = 1 if lits[o] == int(others) else 0
        return out
    if kind == "dense":
        vals = [_weight(v, dtype, exact) for v in desc["values"]]
        return np.array(vals, dtype=dtype).reshape(desc["shape"])
    raise ValueError(kind)


def join_maps(left_free, right_free):
    """(left_edge_map, right_edge_map, free_edges) of one join."""
    lmap, rmap, free = [], [], []
    for i, e in enumerate(left_free):
        if e in right_free:
            lmap.append(i)
            rmap.append(right_free.index(e))
        else:
            free.append(e)
    free += [e for e in right_free if e not in left_free]
    return lmap, rmap, free


def walk(plan, leaf_values, tensordot=np.tensordot):
    """Contract ``leaf_values`` (one array per tensor id) along the plan's tree."""
    stack = []
    for tok in plan["tree"]:
        if tok >= 0:
            stack.append((leaf_values[tok], list(plan["index_lists"][tok])))
        else:
            b, bf = stack.pop()
            a, af = stack.pop()
            lmap, rmap, free = join_maps(af, bf)
            stack.append((tensordot(a, b, (lmap, rmap)), free))
    assert len(stack) == 1
    return stack[0][0]


def _slice_groups(plan):
    """Per group: list of (tensor, axis) positions of its edges."""
    locs = {}
    for t, lst in enumerate(plan["index_lists"]):
        for ax, e in enumerate(lst):
            locs.setdefault(e, []).append((t, ax))
    return [[loc for e in group for loc in locs[e]] for group in plan["groups_to_slice"]]


def slice_leaves(plan, leaves, assignment, groups=None):
    """Leaf values of one slice: every axis in group g fixed to assignment[g] (size 1)."""
    groups = _slice_groups(plan) if groups is None else groups
    index = [[slice(None)] * len(plan["index_lists"][t]) for t in range(len(leaves))]
    touched = set()
    for locs, v in zip(groups, assignment):
        for t, ax in locs:
            index[t][ax] = slice(v, v + 1)
            touched.add(t)
    return [leaves[t][tuple(index[t])] if t in touched else leaves[t] for t in range(len(leaves))]


def contract_sliced(plan, entry_type="float64", num_slice_limit=None, exact=False):
    """Sum over slices (product order) of the contracted value of each slice."""
    dtype = DTYPES[entry_type]
    leaves = [build_leaf(d, dtype, exact) for d in plan["tensors"]]
    groups = _slice_groups(plan)
    assignments = itertools.product((0, 1), repeat=len(groups))
    if num_slice_limit is not None:
        assignments = itertools.islice(assignments, num_slice_limit)
    result = 0
    for a in assignments:
        result += walk(plan, slice_leaves(plan, leaves, a, groups))[()]
    return result


def slice_values(plan, assignments, entry_type="float64"):
    """The contracted value of each listed slice (assignment tuples)."""
    dtype = DTYPES[entry_type]
    leaves = [build_leaf(d, dtype) for d in plan["tensors"]]
    groups = _slice_groups(plan)
    return [walk(plan, slice_leaves(plan, leaves, a, groups))[()] for a in assignments]


def tensordot(a, b, axes):
    """The reference's join primitive (``numpy_apis.py:42-43``)."""
    return np.tensordot(a, b, axes)

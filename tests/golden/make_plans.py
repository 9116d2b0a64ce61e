"""Generate the committed plan fixtures (run in the build container, not on the GPU box).

Each fixture is a ``tensororder-plan/1`` document (see tensororder_amd/plan_io.py) whose
``meta`` also holds the DIMACS instance, the decomposition the plan was extracted from
(PACE .td text), and expected counts from the oracle (float64, exact where feasible,
brute-force enumeration for small instances).  Decompositions come from FlowCutter built
from the reference's own sources by ``make -C oracle ref`` (oracle/_ref/flow_cutter_pace17);
the anytime solver's stream is timing dependent, so the decomposition actually used is
stored in the fixture and the plan is rebuilt from it deterministically.

    python tests/golden/make_plans.py [name ...]
"""

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import brute, numpy_oracle  # noqa: E402
from tensororder_amd import decomposition, network, plan_io, planners  # noqa: E402
from tensororder_amd.formula import parse_dimacs_text  # noqa: E402
from tensororder_amd.instances import random_3cnf, vertex_cover_cnf  # noqa: E402
from tensororder_amd.plan import ALL_SLICERS, SlicedExecutionPlan  # noqa: E402

FLOWCUTTER = os.path.join(ROOT, "oracle", "_ref", "flow_cutter_pace17")
BUNDLED = "/root/reference/benchmarks/cubic_vertex_cover/cubic_vc_50_0.cnf"


def _bundled():
    with open(BUNDLED) as f:
        return f.read()


# name: (dimacs thunk, weight format, planner, FlowCutter seconds (0 = first decomposition),
#        slicing (slice_until kwargs), expectations)
SPECS = {
    "vc50_unweighted_line": (lambda: vertex_cover_cnf(50, 0, False), "unweighted", "line", 0, {},
                             ["float64", "bigint"]),
    "cubic_vc_50_0_line": (_bundled, "unweighted", "line", 0, {}, ["float64", "bigint"]),
    "vc22_weighted_line_s3": (lambda: vertex_cover_cnf(22, 4, True), "cachet", "line", 0, {"slices": 3},
                              ["float64", "brute"]),
    "vc24_weighted_factor_s5": (lambda: vertex_cover_cnf(24, 5, True), "cachet", "factor", 0, {"slices": 5},
                                ["float64", "brute"]),
    "cnf20_factor_s4": (lambda: random_3cnf(20, 60, 11, True), "cachet", "factor", 0, {"slices": 4},
                        ["float64", "brute"]),
    "cnf40_factor_s6": (lambda: random_3cnf(40, 80, 12, True), "cachet", "factor", 0, {"slices": 6},
                        ["float64"]),
    "vc90_weighted_line": (lambda: vertex_cover_cnf(90, 6, True), "cachet", "line", 5, {}, ["float64"]),
    "vc200_weighted_line": (lambda: vertex_cover_cnf(200, 0, True), "cachet", "line", 240, {}, ["float64"]),
    # the 120-variable / 240-clause instance of the original workload description has a
    # factor-Flow plan of max rank 60 (treewidth 53): 12 slices leave rank 48, far beyond
    # any memory.  70 variables at the same clause ratio gives rank 38 -> 25 with 12 slices.
    "cnf70_factor_s12": (lambda: random_3cnf(70, 140, 0, True), "cachet", "factor", 120, {"slices": 12},
                         ["prefix:32"]),
}


def build(name, td_text=None):
    thunk, wfmt, planner, seconds, slicing, expect = SPECS[name]
    dimacs = thunk()
    formula = parse_dimacs_text(dimacs, wfmt)
    net = network.cnf_count(formula)
    graph = net.line_structure_text() if planner == "line" else net.structure_text()
    nverts = net.num_edges() if planner == "line" else len(net)
    if td_text is None:
        cmd = [FLOWCUTTER, "-s", "0", "-p", "100"]
        tds = decomposition.run_solver(cmd, graph, max_decompositions=1 if seconds == 0 else 100000,
                                       time_limit=max(seconds, 10))
    else:
        tds = decomposition.parse_all(td_text)
    best = None
    for td in tds:
        text = td.to_text(nverts)
        if planner == "line":
            tree, nn = planners.plan_line(net, td)
        else:
            tree, nn = planners.plan_factor(net, td, seed=0)
        plan = SlicedExecutionPlan(tree, nn)
        if best is None or best[0].maxrank > plan.maxrank:
            best = (plan, text, td.width())
    plan, td_text, width = best
    unsliced_rank = plan.maxrank
    if slicing:
        ALL_SLICERS["greedy_mem"].slice_until(plan, **slicing)
    meta = {
        "name": name,
        "dimacs": dimacs,
        "weights": wfmt,
        "planner": planner + "-Flow",
        "solver": "FlowCutter PACE17 (-s 0 -p 100), best of its first %s"
                  % ("decomposition" if seconds == 0 else "%d s" % seconds),
        "decomposition": td_text,
        "width": width,
        "maxrank_unsliced": unsliced_rank,
        "maxrank": plan.maxrank,
        "memory_entries": plan.memory,
        "slicer": "greedy_mem" if slicing else None,
        "slicing": slicing,
        "expected": {},
    }
    doc = plan_io.plan_to_dict(plan, meta)
    if "float64" in expect:
        doc["meta"]["expected"]["float64"] = float(numpy_oracle.contract_sliced(doc))
    if "bigint" in expect:
        doc["meta"]["expected"]["bigint"] = str(numpy_oracle.contract_sliced(doc, "bigint"))
    for e in expect:
        if e.startswith("prefix:"):
            k = int(e.split(":")[1])
            doc["meta"]["expected"]["prefix"] = {
                "slices": k, "float64": float(numpy_oracle.contract_sliced(doc, num_slice_limit=k))}
    if "brute" in expect:
        w = {v: formula._variables[v] for v in formula.variables}
        doc["meta"]["expected"]["brute"] = brute.brute_wmc(len(formula.variables), formula.clauses, w)
    return doc


def main(names):
    for name in names or list(SPECS):
        path = os.path.join(HERE, name + ".json")
        td_text = None
        if os.path.exists(path) and "--replan" not in sys.argv:
            with open(path) as f:
                td_text = json.load(f)["meta"]["decomposition"]
        doc = build(name, td_text)
        with open(path, "w") as f:
            json.dump(doc, f, separators=(",", ":"))
        m = doc["meta"]
        print(name, "maxrank", m["maxrank"], "unsliced", m["maxrank_unsliced"], "groups",
              len(doc["groups_to_slice"]), "expected", m["expected"], flush=True)


if __name__ == "__main__":
    main([a for a in sys.argv[1:] if not a.startswith("--")])

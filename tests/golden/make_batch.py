"""Generate the mixed-batch fixture: 512 weighted vertex-cover instances on seeded random
3-regular graphs with even n ~ U[30, 150] (run in the build container, not on the GPU box).

Each instance is planned with the line-graph planner over a FlowCutter decomposition of its
line graph (FlowCutter built from the reference's sources into oracle/_ref; 1 s per
instance, the decomposition of smallest plan max rank in its stream is kept, the
reference's rule).  The fixture stores, per instance, n, the graph seed, the decomposition
(zlib-compressed PACE .td text, base64) and the oracle's float64 count of the plan.

    python tests/golden/make_batch.py [count]
"""

import base64
import json
import multiprocessing
import os
import sys
import zlib

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import numpy_oracle  # noqa: E402
from tensororder_amd import decomposition, network, plan_io, planners  # noqa: E402
from tensororder_amd.formula import parse_dimacs_text  # noqa: E402
from tensororder_amd.instances import mixed_batch_sizes, vertex_cover_cnf  # noqa: E402
from tensororder_amd.plan import SlicedExecutionPlan  # noqa: E402

FLOWCUTTER = os.path.join(ROOT, "oracle", "_ref", "flow_cutter_pace17")


def one(args):
    seed, n = args
    net = network.cnf_count(parse_dimacs_text(vertex_cover_cnf(n, seed, True), "cachet"))
    tds = decomposition.run_solver([FLOWCUTTER, "-s", "0", "-p", "100"], net.line_structure_text(),
                                   max_decompositions=100000, time_limit=1.0)
    best = None
    for td in tds:
        tree, nn = planners.plan_line(net, td)
        if best is None or tree.maxrank < best[0].maxrank:
            best = (tree, td)
    tree, td = best
    plan = SlicedExecutionPlan(tree, net)
    count = float(numpy_oracle.contract_sliced(plan_io.plan_to_dict(plan)))
    text = td.to_text(net.num_edges())
    return {"n": n, "seed": seed, "maxrank": plan.maxrank, "expected": count,
            "decomposition": base64.b64encode(zlib.compress(text.encode(), 9)).decode()}


def main():
    count = int(sys.argv[1]) if len(sys.argv) > 1 else 512
    sizes = mixed_batch_sizes(count, 0, low=30, high=150)
    with multiprocessing.Pool(6) as pool:
        rows = pool.map(one, list(enumerate(sizes)), chunksize=4)
    doc = {"format": "tensororder-batch/1", "name": "vc_batch%d" % count, "weights": "cachet",
           "planner": "line-Flow", "solver": "FlowCutter PACE17 (-s 0 -p 100), best plan of its first 1 s",
           "graphs": "vertex_cover_cnf(n, seed, weighted=True), n from mixed_batch_sizes(%d, 0, 30, 150)" % count,
           "instances": rows}
    with open(os.path.join(HERE, "vc_batch%d.json" % count), "w") as f:
        json.dump(doc, f, separators=(",", ":"))
    ranks = [r["maxrank"] for r in rows]
    print("instances", len(rows), "max rank", min(ranks), "-", max(ranks))


if __name__ == "__main__":
    main()

"""Generate the held-out candidate sets of the learned plan choice (run in the build
container, not on the GPU box): eight seeded random 3-CNFs of 60-67 variables (clause
ratio 2, cachet weights U(0.1, 0.9), seeds 10000-10007), none of them seen in training;
at this size every candidate's full contraction takes well under a second on one MI355X,
so that all of them can be measured.

Per formula, FlowCutter (built from the reference's sources into oracle/_ref by
``make -C oracle ref``) runs under four seeds for a fixed time each and the last
decomposition of each run is kept.  Plans are rebuilt from them deterministically by
``planners.plan_candidates`` (decompositions x slice specs), every candidate sliced to the
same number of groups: the reference rule's plan sliced to max rank <= 24.

    python tests/golden/make_heldout.py [seconds] [--relabel]
"""

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tensororder_amd import decomposition, network, planners  # noqa: E402
from tensororder_amd.formula import parse_dimacs_text  # noqa: E402
from tensororder_amd.instances import random_3cnf  # noqa: E402

FLOWCUTTER = os.path.join(ROOT, "oracle", "_ref", "flow_cutter_pace17")
SEEDS = (0, 2, 3, 7)
RANK_TARGET = 24


def formula(i):
    n = 60 + i
    return n, random_3cnf(n, 2 * n, 10000 + i, True)


def main():
    relabel = "--relabel" in sys.argv  # keep the stored decompositions, rebuild the labels
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    seconds = float(args[0]) if args else 4.0
    old = None
    if relabel:
        with open(os.path.join(HERE, "heldout_candidates.json")) as f:
            old = json.load(f)["sets"]
    docs = []
    for i in range(8):
        n, dimacs = formula(i)
        net = network.cnf_count(parse_dimacs_text(dimacs, "cachet"))
        text = net.structure_text()
        tds = old[i]["decompositions"] if relabel else []
        for seed in SEEDS if not relabel else ():
            got = decomposition.run_solver([FLOWCUTTER, "-s", str(seed), "-p", "100"], text,
                                           max_decompositions=100000, time_limit=seconds)
            td = got[-1]
            tds.append({"solver_seed": seed, "width": td.width(), "text": td.to_text(len(net))})
        parsed = [decomposition.parse_all(t["text"])[0] for t in tds]
        # slices: the reference rule's plan (smallest unsliced max rank) down to RANK_TARGET
        unsliced = []
        for t in tds:  # on fresh parses: building a plan extends the decomposition it is given
            tree, _ = planners.plan_factor(net, decomposition.parse_all(t["text"])[0], seed=0)
            unsliced.append(tree.maxrank)
        slices = max(4, min(unsliced) - RANK_TARGET)
        plans, labels = planners.plan_candidates(net, "factor", parsed, slices)
        ref = planners.reference_candidate(labels)
        docs.append({"format": "tensororder-candidates/1", "name": "heldout%d_v%d" % (i, n), "dimacs": dimacs,
                     "weights": "cachet", "planner": "factor-Flow", "slices": slices,
                     "solver": "FlowCutter PACE17 (-s SEED -p 100), last decomposition after %g s" % seconds,
                     "decompositions": tds, "candidates": labels, "reference": ref})
        print("formula", i, "vars", n, "widths", [t["width"] for t in tds], "unsliced", unsliced, "slices", slices,
              "candidates", len(plans), "reference", ref, "maxranks", sorted({l["maxrank"] for l in labels}),
              flush=True)
    with open(os.path.join(HERE, "heldout_candidates.json"), "w") as f:
        json.dump({"format": "tensororder-candidate-sets/1", "sets": docs}, f, separators=(",", ":"))


if __name__ == "__main__":
    main()

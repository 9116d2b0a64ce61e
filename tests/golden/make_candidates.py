"""Generate the candidate-plan fixture of the 4096-slice random 3-CNF (run in the build
container, not on the GPU box).

FlowCutter (built from the reference's sources into oracle/_ref by ``make -C oracle ref``)
runs under four seeds for a fixed time each; the last decomposition of each run is kept.
The fixture stores the DIMACS instance and those decompositions (PACE .td text); plans are
rebuilt from them deterministically by ``planners.plan_candidates`` (decompositions x
slice specs, 12 slice groups each).

    python tests/golden/make_candidates.py [seconds] [--relabel]
"""

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tensororder_amd import decomposition, network, planners  # noqa: E402
from tensororder_amd.formula import parse_dimacs_text  # noqa: E402

FLOWCUTTER = os.path.join(ROOT, "oracle", "_ref", "flow_cutter_pace17")
SEEDS = (0, 2, 3, 7)


def main():
    relabel = "--relabel" in sys.argv  # keep the stored decompositions, rebuild the labels
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    seconds = float(args[0]) if args else 15.0
    with open(os.path.join(HERE, "cnf70_factor_s12.json")) as f:
        dimacs = json.load(f)["meta"]["dimacs"]
    net = network.cnf_count(parse_dimacs_text(dimacs, "cachet"))
    text = net.structure_text()
    tds = []
    if relabel:
        with open(os.path.join(HERE, "cnf70_candidates.json")) as f:
            old = json.load(f)
        tds, seconds = old["decompositions"], float(old["solver"].split("after ")[1].split(" s")[0])
    for seed in SEEDS if not relabel else ():
        got = decomposition.run_solver([FLOWCUTTER, "-s", str(seed), "-p", "100"], text,
                                       max_decompositions=100000, time_limit=seconds)
        td = got[-1]
        tds.append({"solver_seed": seed, "width": td.width(), "text": td.to_text(len(net))})
        print("seed", seed, "decompositions", len(got), "last width", td.width(), flush=True)
    plans, labels = planners.plan_candidates(net, "factor", [decomposition.parse_all(t["text"])[0] for t in tds], 12)
    ref = planners.reference_candidate(labels)
    doc = {"format": "tensororder-candidates/1", "name": "cnf70_factor_s12_candidates", "dimacs": dimacs,
           "weights": "cachet", "planner": "factor-Flow", "slices": 12,
           "solver": "FlowCutter PACE17 (-s SEED -p 100), last decomposition after %g s" % seconds,
           "decompositions": tds, "candidates": labels, "reference": ref}
    with open(os.path.join(HERE, "cnf70_candidates.json"), "w") as f:
        json.dump(doc, f, separators=(",", ":"))
    for i, lab in enumerate(labels):
        print(i, "*" if i == ref else " ", lab, flush=True)


if __name__ == "__main__":
    main()

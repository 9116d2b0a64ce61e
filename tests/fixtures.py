"""Paths of the committed plan fixtures (tensororder-plan/1 documents) under tests/golden;
the candidate-set and batch fixtures there have their own formats."""

import glob
import json
import os

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def plan_fixture_paths():
    out = []
    for path in sorted(glob.glob(os.path.join(GOLDEN, "*.json"))):
        with open(path) as f:
            head = f.read(64)
        if '"format":"tensororder-plan/1"' in head.replace(" ", ""):
            out.append(path)
    return out

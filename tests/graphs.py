"""Small graphs with known independent-set counts (= vertex-cover counts) and helpers
turning a graph into a vertex-cover CNF."""

import itertools

from tensororder_amd.formula import parse_dimacs_text


def complete(n):
    return n, [(i, j) for i in range(n) for j in range(i + 1, n)]


def complete_bipartite(a, b):
    return a + b, [(i, a + j) for i in range(a) for j in range(b)]


def cycle(n):
    return n, [(i, (i + 1) % n) for i in range(n)]


def path(n):
    return n, [(i, i + 1) for i in range(n - 1)]


def prism():
    return 6, [(0, 1), (1, 2), (2, 0), (3, 4), (4, 5), (5, 3), (0, 3), (1, 4), (2, 5)]


def cube():
    n = 8
    return n, [(u, u ^ (1 << b)) for u in range(n) for b in range(3) if u < u ^ (1 << b)]


def petersen():
    outer = [(i, (i + 1) % 5) for i in range(5)]
    spokes = [(i, i + 5) for i in range(5)]
    inner = [(5 + i, 5 + (i + 2) % 5) for i in range(5)]
    return 10, outer + spokes + inner


def lucas(n):
    a, b = 2, 1
    for _ in range(n):
        a, b = b, a + b
    return a


def fibonacci(n):
    a, b = 0, 1
    for _ in range(n):
        a, b = b, a + b
    return a


def count_independent_sets(n, edges):
    total = 0
    for bits in itertools.product((0, 1), repeat=n):
        if all(not (bits[u] and bits[v]) for u, v in edges):
            total += 1
    return total


def vc_dimacs(n, edges, weights=None):
    lines = ["p cnf %d %d" % (n, len(edges))]
    if weights is not None:
        lines += ["w %d %.17g" % (v + 1, p) for v, p in enumerate(weights)]
    lines += ["%d %d 0" % (u + 1, v + 1) for u, v in edges]
    return "\n".join(lines) + "\n"


def vc_formula(n, edges, weights=None):
    return parse_dimacs_text(vc_dimacs(n, edges, weights), "cachet" if weights is not None else "unweighted")

"""Seeded synthetic instances of the measured workloads, written as DIMACS text.

* vertex cover of a random 3-regular graph as monotone 2-CNF, one clause ``u v 0`` per
  edge, unweighted or with cachet weights ``w v p``, p ~ U(0.1, 0.9);
* random 3-CNF: each clause has 3 distinct uniform variables with uniform random signs.

Each instance draws from its own ``numpy.random.Generator(PCG64(seed))``.  3-regular
graphs come from the configuration model, rejecting pairings with loops or multi-edges.
Weights are printed with 17 significant digits so they parse back exactly.
"""

import numpy as np


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def random_regular_graph(n, d, rng, max_tries=10000):
    """Edges (0-based, u < v) of a uniform simple d-regular graph on n vertices."""
    if (n * d) % 2:
        raise ValueError("n*d must be even")
    for _ in range(max_tries):
        stubs = np.repeat(np.arange(n), d)
        rng.shuffle(stubs)
        pairs = stubs.reshape(-1, 2)
        u = np.minimum(pairs[:, 0], pairs[:, 1])
        v = np.maximum(pairs[:, 0], pairs[:, 1])
        if np.any(u == v):
            continue
        keys = u.astype(np.int64) * n + v
        if len(np.unique(keys)) != len(keys):
            continue
        order = np.lexsort((v, u))
        return [(int(u[i]), int(v[i])) for i in order]
    raise RuntimeError("no simple pairing found")


def _weights(n, rng):
    return rng.uniform(0.1, 0.9, size=n)


def vertex_cover_cnf(n, seed, weighted):
    """DIMACS text of (weighted) vertex cover on a seeded random 3-regular graph."""
    rng = _rng(seed)
    edges = random_regular_graph(n, 3, rng)
    lines = ["p cnf %d %d" % (n, len(edges))]
    if weighted:
        for v, p in enumerate(_weights(n, rng), start=1):
            lines.append("w %d %.17g" % (v, p))
    lines += ["%d %d 0" % (u + 1, v + 1) for u, v in edges]
    return "\n".join(lines) + "\n"


def random_3cnf(num_vars, num_clauses, seed, weighted=True):
    """DIMACS text of a seeded random 3-CNF with optional cachet weights."""
    rng = _rng(seed)
    clauses = []
    for _ in range(num_clauses):
        vs = rng.choice(num_vars, size=3, replace=False) + 1
        signs = rng.integers(0, 2, size=3) * 2 - 1
        clauses.append([int(v * s) for v, s in zip(vs, signs)])
    lines = ["p cnf %d %d" % (num_vars, num_clauses)]
    if weighted:
        for v, p in enumerate(_weights(num_vars, rng), start=1):
            lines.append("w %d %.17g" % (v, p))
    lines += [" ".join(map(str, c)) + " 0" for c in clauses]
    return "\n".join(lines) + "\n"


def mixed_batch_sizes(count, seed, low=30, high=150):
    """Even vertex counts n ~ U[low, high] for a batch of vertex-cover instances."""
    rng = _rng(seed)
    return [int(x) for x in (rng.integers(low // 2, high // 2 + 1, size=count) * 2)]

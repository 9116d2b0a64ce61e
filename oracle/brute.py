"""ORACLE -- test infrastructure only.

ctypes wrapper of ``brute_wmc.c`` (built by ``oracle/Makefile`` into ``oracle/_build``)
plus a pure-Python enumerator for tiny formulas that cross-checks it.
"""

import ctypes
import itertools
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "libbrute_wmc.so")


def _formula_arrays(num_vars, clauses, weights):
    pos = np.zeros(len(clauses), dtype=np.uint64)
    neg = np.zeros(len(clauses), dtype=np.uint64)
    for i, c in enumerate(clauses):
        for lit in c:
            bit = np.uint64(1) << np.uint64(abs(lit) - 1)
            if lit > 0:
                pos[i] |= bit
            else:
                neg[i] |= bit
    w_neg = np.array([weights[v][0] for v in range(1, num_vars + 1)], dtype=np.float64)
    w_pos = np.array([weights[v][1] for v in range(1, num_vars + 1)], dtype=np.float64)
    return pos, neg, w_neg, w_pos


def brute_wmc(num_vars, clauses, weights):
    """Weighted model count by enumeration in C.  ``weights[v] = (w(not v), w(v))``."""
    lib = ctypes.CDLL(_LIB)
    lib.brute_wmc.restype = ctypes.c_double
    pos, neg, w_neg, w_pos = _formula_arrays(num_vars, clauses, weights)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    r = lib.brute_wmc(num_vars, len(clauses), p(pos), p(neg), p(w_neg), p(w_pos))
    if r < 0:
        raise ValueError("too many variables for enumeration")
    return r


def brute_wmc_py(num_vars, clauses, weights):
    """Pure-Python enumeration (small formulas only)."""
    total = 0.0
    for bits in itertools.product((0, 1), repeat=num_vars):
        if all(any((bits[abs(l) - 1] == 1) == (l > 0) for l in c) for c in clauses):
            w = 1.0
            for v in range(num_vars):
                w *= weights[v + 1][bits[v]]
            total += w
    return total

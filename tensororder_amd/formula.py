"""DIMACS CNF parsing and literal-weight formats.

Restates the reference reduction input: ``util/boolean_formula.py:5-9`` (weight
formats) and ``util/boolean_formula.py:199-262`` (``Formula.parse_DIMACS``).  The
variable order matters downstream: it fixes the tensor ids of the network built by
``network.cnf_count``.  Variables named on ``w`` lines come first (in the order those
lines appear), then every remaining variable ``1..n`` in increasing order
(``boolean_formula.py:254-261``).
"""

from collections import OrderedDict
from enum import Enum


class WeightFormat(Enum):
    unweighted = 0
    cachet = 1
    minic2d = 2
    mcc = 3


class Formula:
    """A CNF formula with per-literal multiplicative weights."""

    def __init__(self):
        # var id -> [weight of the negative literal, weight of the positive literal]
        self._variables = OrderedDict()
        self._clauses = []

    def add_clause(self, literals):
        self._clauses.append(list(literals))

    def fresh_variable(self, neg_weight, pos_weight):
        var = len(self._variables) + 1
        self._variables[var] = [neg_weight, pos_weight]
        return var

    @property
    def clauses(self):
        return list(self._clauses)

    @property
    def variables(self):
        return list(self._variables.keys())

    def literal_weight(self, lit):
        return self._variables[abs(lit)][1 if lit > 0 else 0]

    def set_literal_weight(self, lit, weight):
        if abs(lit) not in self._variables:
            self._variables[abs(lit)] = [1, 1]
        self._variables[abs(lit)][1 if lit > 0 else 0] = weight

    def set_variable_weight(self, var, neg_weight, pos_weight):
        self._variables[var] = [neg_weight, pos_weight]

    @staticmethod
    def parse_DIMACS(lines, weight_format):
        """Parse DIMACS text (an iterable of lines) under ``weight_format``.

        Semantics follow ``boolean_formula.py:199-262``: ``c weights`` lines carry
        miniC2D weights (positive weight first in each pair); ``w v p`` lines carry cachet
        weights (p = -1 means 1/1) or, under mcc, a single literal weight; a ``w`` line in
        any other format is an error; every variable up to the ``p`` line's count that
        has no weight yet gets 0.5/0.5 under cachet and 1/1 otherwise.
        """
        if isinstance(weight_format, str):
            weight_format = WeightFormat[weight_format]
        result = Formula()
        num_vars = 0
        for line in lines:
            if weight_format == WeightFormat.minic2d and line.startswith("c weights"):
                weights = line.split(" ")[2:]
                for i in range(len(weights) // 2):
                    result.set_variable_weight(
                        i + 1, float(weights[2 * i + 1]), float(weights[2 * i])
                    )
            elif len(line) == 0 or line[0] == "c":
                continue
            elif line[0] == "p":
                num_vars = int(line.split()[2])
            elif line[0] == "w":
                args = line.split()
                var = int(args[1])
                weight = float(args[2])
                if weight_format == WeightFormat.cachet:
                    if weight == -1:
                        result.set_variable_weight(var, 1, 1)
                    else:
                        result.set_variable_weight(var, 1 - weight, weight)
                elif weight_format == WeightFormat.mcc:
                    result.set_literal_weight(var, weight)
                else:
                    raise RuntimeError("w lines cannot be used in " + str(weight_format))
            else:
                literals = [lit for lit in map(int, line.split()) if lit != 0]
                if len(literals) == 0:
                    continue
                result.add_clause(literals)

        default = 0.5 if weight_format == WeightFormat.cachet else 1
        for var in range(1, num_vars + 1):
            if var not in result._variables:
                result.set_variable_weight(var, default, default)
        return result


def parse_dimacs_text(text, weight_format):
    """Parse a DIMACS string; a convenience over ``Formula.parse_DIMACS``."""
    return Formula.parse_DIMACS(text.splitlines(keepends=True), weight_format)

"""Sliced execution plans and slicers.

Mirror of ``tensor_network/sliced_execution_plan.py:4-108`` and
``tensor_network/slicers.py:6-124``.  A plan holds the contraction tree, the (possibly
factored) network and ``groups_to_slice``: each group is one variable's full class of
equivalent edges (``find_equivalent_edges``), so fixing a group to 0/1 halves every
tensor containing it and doubles the number of independent slices.
"""

import collections
import math
import random


class SlicedExecutionPlan:
    def __init__(self, tree, network):
        self.tree = tree
        self.network = network
        self.edges_to_slice = set()
        self.groups_to_slice = []
        self.FLOPs, self.memory, self.next_edge_to_slice, self.maxrank = self.tree.estimate_cost(
            self.edges_to_slice
        )

    def contract_small(self, below_size, tensor_api):
        """Early contraction (``sliced_execution_plan.py:29-56``): reroot so a large tensor
        is on top, regroup small neighbours, contract every join below ``below_size`` now,
        and renumber the slice groups' edges."""
        self.tree = self.tree.reroot(self.network, self.maxrank)
        self.tree.regroup(below_size)
        self.tree, new_eid = self.network.identify_partial(self.tree, tensor_api, below_size)
        self.edges_to_slice = set(new_eid[e] for e in self.edges_to_slice if new_eid[e] >= 0)
        self.groups_to_slice = [set(new_eid[e] for e in g if new_eid[e] >= 0) for g in self.groups_to_slice]
        self.next_edge_to_slice = None

    def slice_at(self, edge):
        equivalent = self.network.find_equivalent_edges(edge)
        self.groups_to_slice.append(equivalent)
        self.edges_to_slice |= equivalent
        self.FLOPs, self.memory, self.next_edge_to_slice, self.maxrank = self.tree.estimate_cost(
            self.edges_to_slice
        )

    @property
    def total_FLOPs(self):
        return self.FLOPs * (2 ** len(self.groups_to_slice))

    @property
    def num_slices(self):
        return 2 ** len(self.groups_to_slice)

    @property
    def widths(self):
        widths = {"Max Rank": self.maxrank}
        if hasattr(self.tree, "treewidth"):
            widths["Treewidth"] = self.tree.treewidth
        if hasattr(self.tree, "branchwidth"):
            widths["Branchwidth"] = self.tree.branchwidth
        return widths

    def statistics(self):
        """The plan statistics the reference reports (``sliced_execution_plan.py:98-108``)."""
        return {
            "# Sliced": len(self.groups_to_slice),
            "# Network Slices": 2 ** len(self.groups_to_slice),
            "Estimated Memory": float(self.memory),
            "Estimated FLOPs": float(self.total_FLOPs),
        }


class BaseSlicer:
    def slice_once(self, plan):
        raise NotImplementedError

    def slice_until(self, plan, memory=None, rank=None, slices=None):
        """Slice until memory <= limit (entries), then rank <= limit, then at least
        ``slices`` groups (``slicers.py:10-24``)."""
        while memory is not None and memory < plan.memory:
            self.slice_once(plan)
        while rank is not None and rank < plan.maxrank:
            self.slice_once(plan)
        while slices is not None and len(plan.groups_to_slice) < slices:
            self.slice_once(plan)


class GreedyMemSlicer(BaseSlicer):
    """Slice the cost model's best edge (``slicers.py:27-33``)."""

    def slice_once(self, plan):
        plan.slice_at(plan.next_edge_to_slice)


class GreedyLargestSlicer(BaseSlicer):
    """Slice a random edge of the largest intermediate (``slicers.py:36-49``)."""

    def slice_once(self, plan):
        max_size, edges = 0, None
        for node in plan.tree.iterate_postorder():
            size = sum(0 if e in plan.edges_to_slice else 1 for e in node.free_edges)
            if size > max_size:
                max_size, edges = size, node.free_edges
        plan.slice_at(random.choice(edges))


class GreedyMostSlicer(GreedyMemSlicer):
    """Slice the edge incident to the most over-limit intermediates, then fall back to
    ``greedy_mem`` (``slicers.py:52-99``)."""

    def slice_until(self, plan, memory=None, rank=None, slices=None):
        limit = min(
            float("inf") if rank is None else rank,
            float("inf") if memory is None else math.floor(math.log2(memory)),
        )
        if not math.isinf(limit):
            rep = plan.network.equivalent_edge_sets()
            nodes = list(plan.tree.iterate_postorder())
            sizes = [sum(0 if e in plan.edges_to_slice else 1 for e in n.free_edges) for n in nodes]
            large = collections.defaultdict(set)
            for i in range(len(nodes)):
                if sizes[i] > limit:
                    for e in nodes[i].free_edges:
                        if e not in plan.edges_to_slice:
                            large[rep[e]].add(i)
            if len(large) > 0:
                while True:
                    max_count = max(len(n) for n in large.values())
                    if max_count == 0:
                        break
                    chosen = random.choice([e for e in large if len(large[e]) == max_count])
                    for i in list(large[chosen]):
                        for e in nodes[i].free_edges:
                            if e not in plan.edges_to_slice:
                                large[rep[e]].discard(i)
                    plan.slice_at(chosen)
        super().slice_until(plan, memory, rank, slices)


class DisabledSlicer(BaseSlicer):
    def slice_once(self, plan):
        raise RuntimeError("Slicing is disabled")

    def slice_until(self, plan, memory=None, rank=None, slices=None):
        if memory is not None and memory < plan.memory:
            raise RuntimeError("Slicing is disabled but plan memory is too large ")
        if rank is not None and rank < plan.maxrank:
            raise RuntimeError("Slicing is disabled but plan maxrank is too large")
        if slices is not None and len(plan.groups_to_slice) < slices:
            raise RuntimeError("Slicing is disabled; --minimum_slice is disallowed")


ALL_SLICERS = {
    "greedy_mem": GreedyMemSlicer(),
    "disable": DisabledSlicer(),
    "greedy_largest": GreedyLargestSlicer(),
    "greedy_most": GreedyMostSlicer(),
}

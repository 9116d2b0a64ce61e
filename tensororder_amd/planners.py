"""Contraction-tree extraction from tree decompositions (the line-* and factor-* planners).

* ``line_graph_tree`` restates ``contraction_methods/line_graph_method.py:8-35``: walk a
  tree decomposition of the network's line graph (vertices = network edges) in
  post-order from bag 1, joining the children's trees and then every remaining tensor
  whose edges all lie in the current bag.
* ``factor_tree_plan`` restates ``contraction_methods/factor_tree_method.pyx:18-253``
  (Factor-Tree, Dudek et al. 2019, Thm 4): attach one leaf per network edge to a bag
  holding both of its tensors, split bags to degree <= 3, then factor high-rank tensors
  while propagating edges up the decomposition.  The network is modified (factored) in
  place, so callers pass a copy.

Two intentional differences, neither affecting the value of the contraction: the
recursion of ``factor_tree`` is an explicit post-order loop, and per-node maps iterate in
insertion order (the reference iterates C++ ``unordered_map`` order).
"""

import random

from .decomposition import (
    TreeDecomposition,
    min_degree_decomposition,
    parse_graph,
    run_solver,
)
from .tree import TreeContext


def line_graph_tree(network, td, root=1):
    ctx = TreeContext()
    trees = [ctx.leaf(network, i) for i in range(len(network))]
    edges = [frozenset(ctx.nodes[t].free_edges) for t in trees]

    def combine(node, children_trees):
        bag = set(td.bags[node])
        matches = {(tree, fe) for tree, fe in zip(trees, edges) if bag.issuperset(fe)}
        for tree, fe in matches:
            trees.remove(tree)
            edges.remove(fe)
        complete = ctx.empty()
        for tree in children_trees:
            complete = ctx.join(complete, tree)
        for tree, _ in matches:
            complete = ctx.join(complete, tree)
        return complete

    result = td.traverse_postorder(root, combine)
    tree = ctx.get_tree(ctx.include_rank_zero_tensors(network, result))
    tree.treewidth = td.width()
    return tree


def _edge_leaves(network, td, rng):
    """For every network edge, extend a bag containing both of its tensors with a new leaf
    and map that leaf to the edge (``factor_tree_method.pyx:18-69``)."""
    n = len(network)
    pairs = {}
    for node in range(len(td.bags)):
        for a in td.bags[node]:
            for b in td.bags[node]:
                if a != b:
                    pairs.setdefault(a * n + b, []).append(node)
    cache, node_to_edge = {}, {}
    for edge_id in range(network.num_edges()):
        e = network._edges[edge_id]
        key = min(e[1], e[2]) * n + max(e[1], e[2])
        if key not in cache:
            cands = pairs.get(key, [])
            idx = 0 if len(cands) == 1 else rng.randint(0, len(cands) - 1)
            cache[key] = td.extend_leaf(cands[idx])
        else:
            cache[key] = td.extend_leaf(cache[key])
        node_to_edge[cache[key]] = edge_id
    return node_to_edge


def _factor_tree(neighbors, root, network, ctx, node_to_edge):
    """Factor-Tree over a rooted degree-<=3 tree (``factor_tree_method.pyx:77-229``)."""
    # Explicit post-order: a node is finished after children[0]'s subtree, then
    # children[1]'s subtree -- the call order of the recursive formulation.
    results = {}
    stack = [(root, -1, False)]
    while stack:
        node, parent, expanded = stack.pop()
        children = [c for c in neighbors[node] if c != parent]
        if not expanded and children:
            stack.append((node, parent, True))
            for c in reversed(children):
                stack.append((c, node, False))
            continue
        if len(children) == 1:
            results[node] = results.pop(children[0])
            continue
        if len(children) == 0:
            tree, inter = ctx.empty(), {}
            if node in node_to_edge:
                e = network._edges[node_to_edge[node]]
                for t in (e[1], e[2]):
                    if len(network._index_lists[t]) <= 1:
                        tree = ctx.join(tree, ctx.leaf(network, t))
                    else:
                        inter[t] = e[0]
            results[node] = (tree, inter)
            continue
        left_tree, left_ie = results.pop(children[0])
        right_tree, right_ie = results.pop(children[1])
        res_ie = {}
        to_contract = []
        for t in list(left_ie):
            if t not in right_ie:
                res_ie[t] = left_ie[t]
            elif left_ie[t] == right_ie[t]:
                continue
            elif len(network._index_lists[t]) == 2:
                if len(left_ie) < len(right_ie):
                    left_tree = ctx.join(left_tree, ctx.leaf(network, t))
                else:
                    right_tree = ctx.join(right_tree, ctx.leaf(network, t))
            else:
                idx = network._index_lists[t]
                li = idx.index(left_ie[t])
                ri = idx.index(right_ie[t])
                new_tensor, new_edge = network.factor_out(t, li, ri)
                res_ie[t] = new_edge
                to_contract.append(new_tensor)
        for t in right_ie:
            if t not in left_ie:
                res_ie[t] = right_ie[t]
        rank_l, rank_r, rank_res = len(left_ie), len(right_ie), len(res_ie)
        g_l, g_r, g_res = [], [], []
        for t in to_contract:
            if rank_l <= rank_r and rank_l <= rank_res:
                g_l.append(ctx.leaf(network, t))
                rank_l += 1
            elif rank_r <= rank_res:
                g_r.append(ctx.leaf(network, t))
                rank_r += 1
            else:
                g_res.append(ctx.leaf(network, t))
                rank_res += 1
        for leaf in g_l:
            left_tree = ctx.join(left_tree, leaf)
        for leaf in g_r:
            right_tree = ctx.join(right_tree, leaf)
        tree = ctx.join(left_tree, right_tree)
        for leaf in g_res:
            tree = ctx.join(tree, leaf)
        results[node] = (tree, res_ie)
    return results[root]


def factor_tree_plan(network, td, seed=None):
    """Factor-Tree contraction tree of ``network`` (modified in place) from a tree
    decomposition of its tensor graph.  Returns the tree."""
    rng = random.Random(seed)
    node_to_edge = _edge_leaves(network, td, rng)
    td.split_high_degree_nodes()
    root = td.extend_leaf(1)
    ctx = TreeContext()
    tree_id, _ = _factor_tree(td.neighbors, root, network, ctx, node_to_edge)
    tree = ctx.get_tree(ctx.include_rank_zero_tensors(network, tree_id))
    return tree


def decompose(graph_text, solver=None, max_decompositions=1, tie_seeds=None):
    """Tree decompositions of a PACE graph: from an external anytime solver (argv list,
    e.g. a FlowCutter binary with ``-s seed -p 100``) or, with ``solver=None``, the
    built-in greedy min-degree decomposition (one per entry of ``tie_seeds``; None =
    deterministic ties)."""
    if solver is None:
        n, edges = parse_graph(graph_text)
        return [min_degree_decomposition(n, edges, seed=s) for s in (tie_seeds or [None])]
    return run_solver(solver, graph_text, max_decompositions=max_decompositions)


def plan_line(network, td):
    """(tree, network) for the line-graph planner given a decomposition of the line graph."""
    return line_graph_tree(network, td), network


def plan_factor(network, td, seed=None):
    """(tree, factored copy of network) for the factor-tree planner."""
    net = network.copy()
    tree = factor_tree_plan(net, td, seed=seed)
    tree.treewidth = None
    return tree, net


def candidate_plans(network, planner, tie_seeds, seed=0):
    """One unsliced plan per built-in decomposition (random tie orders ``tie_seeds``)."""
    from .plan import SlicedExecutionPlan

    text = network.line_structure_text() if planner == "line" else network.structure_text()
    out = []
    for td in decompose(text, None, tie_seeds=tie_seeds):
        width = td.width()
        tree, net = plan_line(network, td) if planner == "line" else plan_factor(network, td, seed=seed)
        tree.treewidth = width
        out.append(SlicedExecutionPlan(tree, net))
    return out


def best_plan(network, planner, solver=None, max_decompositions=1, seed=None):
    """Smallest-max-rank tree over the decompositions, the reference's rule
    (``planning.py:159-162``).  ``planner`` is ``"line"`` or ``"factor"``."""
    from .plan import SlicedExecutionPlan

    text = network.line_structure_text() if planner == "line" else network.structure_text()
    best = None
    for td in decompose(text, solver, max_decompositions):
        width = td.width()
        if planner == "line":
            tree, net = plan_line(network, td)
        else:
            tree, net = plan_factor(network, td, seed=seed)
        tree.treewidth = width
        if best is None or best.tree.maxrank > tree.maxrank:
            best = SlicedExecutionPlan(tree, net)
    return best


# Slice sets tried per decomposition when choosing a plan by predicted cost: the reference's
# greedy_mem (``slicers.py:27-33``), greedy_most at rank caps (``slicers.py:52-99``) and
# greedy_largest under fixed random seeds (``slicers.py:36-49``); (slicer, slice_until
# keyword arguments, random seed).
SLICE_SPECS = (
    ("greedy_mem", {}, 0),
    ("greedy_most", {"rank": 25}, 0),
    ("greedy_most", {"rank": 26}, 0),
    ("greedy_most", {"rank": 27}, 0),
    ("greedy_most", {"rank": 28}, 0),
    ("greedy_largest", {}, 0),
    ("greedy_largest", {}, 1),
    ("greedy_largest", {}, 2),
)


def plan_candidates(network, planner, decompositions, slices, specs=SLICE_SPECS, seed=0):
    """Candidate plans of one network: every decomposition (TreeDecomposition) x every
    slice spec, each sliced to exactly ``slices`` distinct groups (a spec that needs more
    groups to meet its own limits, or that slices one group twice, is dropped).  Returns (plans, labels); a label names the
    decomposition index, its unsliced max rank and the slice spec."""
    import random as _random

    from .plan import ALL_SLICERS, SlicedExecutionPlan

    plans, labels = [], []
    for di, td in enumerate(decompositions):
        width = td.width()
        for slicer, kw, rseed in specs:
            if planner == "line":
                tree, net = plan_line(network, td)
            else:
                tree, net = plan_factor(network, td, seed=seed)
            tree.treewidth = width
            plan = SlicedExecutionPlan(tree, net)
            unsliced = plan.maxrank
            state = _random.getstate()
            _random.seed(rseed)
            try:
                ALL_SLICERS[slicer].slice_until(plan, slices=slices, **kw)
            finally:
                _random.setstate(state)
            if len(plan.groups_to_slice) != slices:
                continue
            # greedy_largest may pick an edge that is already sliced (slicers.py:36-49
            # chooses among all free edges of the largest node) and slice its group again:
            # the first copy then fixes no leaf and the sum over it doubles the count (the
            # reference's own execution does the same).  Such a plan is no candidate.
            if len({frozenset(g) for g in plan.groups_to_slice}) != len(plan.groups_to_slice):
                continue
            plans.append(plan)
            labels.append({"decomposition": di, "width": width, "unsliced_maxrank": unsliced,
                           "slicer": slicer, "args": dict(kw), "seed": rseed, "maxrank": plan.maxrank})
    return plans, labels


def reference_candidate(labels):
    """The reference's choice among candidates: the decomposition whose plan has the
    smallest unsliced max rank (first on ties, ``planning.py:159-162``), sliced by
    greedy_mem."""
    best = None
    for i, lab in enumerate(labels):
        if lab["slicer"] != "greedy_mem":
            continue
        if best is None or lab["unsliced_maxrank"] < labels[best]["unsliced_maxrank"]:
            best = i
    return best

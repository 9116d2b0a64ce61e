"""Binary contraction trees and the reference cost model.

Host-side mirror of ``contraction_methods/contraction_tree.pyx``:

* node fields and accessors (:4-47, :200-209) and ``iterate_postorder`` (:174-183);
* ``compute_join_properties`` (:248-288): a join's shared axes are paired in ascending
  left-axis order (``left_edge_map[k]`` <-> ``right_edge_map[k]``) and its free edges are
  the left operand's unshared edges in order followed by the right operand's -- exactly
  ``numpy.tensordot``'s output axis order;
* ``include_rank_zero_tensors`` (:302-310), ``sort_small`` (:137-162);
* ``estimate_cost`` (:382-496): FLOPs, memory cap (in entries), max rank and the best
  edge to slice, which drives the ``greedy_mem`` slicer.
"""


class _Node:
    __slots__ = (
        "left", "right", "tensor_index", "free_edges",
        "left_edge_map", "right_edge_map", "maxrank", "is_leaf",
    )

    def __init__(self):
        self.left = -1
        self.right = -1
        self.tensor_index = -1
        self.free_edges = []
        self.left_edge_map = []
        self.right_edge_map = []
        self.maxrank = 0
        self.is_leaf = False


class _Cost:
    __slots__ = ("FLOPs", "total_memory", "local_memory", "largest_tensor",
                 "best_edge", "best_edge_memory", "open_edge_total_memory")

    def __init__(self):
        self.FLOPs = 0.0
        self.total_memory = 0.0
        self.local_memory = 0.0
        self.largest_tensor = 0
        self.best_edge = -1
        self.best_edge_memory = 0.0
        self.open_edge_total_memory = {}


def _at(d, key):
    """``unordered_map::operator[]``: read, inserting a zero entry when absent."""
    if key not in d:
        d[key] = 0.0
    return d[key]


class TreeContext:
    """Storage for tree nodes; ``leaf`` / ``join`` return node ids (-1 is empty)."""

    def __init__(self):
        self.nodes = []

    def get_data(self, index):
        return self.nodes[index]

    def get_tree(self, index):
        if index < 0:
            raise RuntimeError("Cannot get interface to empty tree")
        return ContractionTree(self, index)

    def empty(self):
        return -1

    def leaf(self, network, tensor_index):
        return self.leaf_manual(tensor_index, network.index_list(tensor_index))

    def leaf_manual(self, tensor_index, free_edges):
        n = _Node()
        n.is_leaf = True
        n.tensor_index = tensor_index
        n.free_edges = list(free_edges)
        n.maxrank = len(n.free_edges)
        self.nodes.append(n)
        return len(self.nodes) - 1

    def compute_join_properties(self, node):
        j = self.nodes[node]
        lt = self.nodes[j.left]
        rt = self.nodes[j.right]
        j.free_edges = []
        j.left_edge_map = []
        j.right_edge_map = []
        rpos = {}
        for idx, e in enumerate(rt.free_edges):
            rpos.setdefault(e, idx)
        lset = set(lt.free_edges)
        for i, e in enumerate(lt.free_edges):
            if e in rpos:
                j.left_edge_map.append(i)
                j.right_edge_map.append(rpos[e])
            else:
                j.free_edges.append(e)
        for e in rt.free_edges:
            if e not in lset:
                j.free_edges.append(e)
        j.maxrank = max(len(j.free_edges), lt.maxrank, rt.maxrank)
        return node

    def join(self, left, right):
        if left == -1:
            return right
        if right == -1:
            return left
        n = _Node()
        n.left = left
        n.right = right
        self.nodes.append(n)
        return self.compute_join_properties(len(self.nodes) - 1)

    def include_rank_zero_tensors(self, network, tree):
        combined = self.empty()
        for tensor_id in range(len(network)):
            if len(network.index_list(tensor_id)) == 0:
                combined = self.join(combined, self.leaf(network, tensor_id))
        return self.join(combined, tree)

    def group_if_below(self, upper, upper_left, lower_left, if_below):
        """At (A (B C)) with A = upper's ``upper_left`` child and B = the lower join's
        ``lower_left`` child, regroup to (C (A B)) when rank(A B) < if_below
        (``contraction_tree.pyx:324-380``).  Returns 1 if regrouped."""
        un = self.nodes[upper]
        if upper_left:
            chosen, lower = un.left, un.right
        else:
            chosen, lower = un.right, un.left
        ln = self.nodes[lower]
        if lower_left:
            lchosen, lother = ln.left, ln.right
        else:
            lchosen, lother = ln.right, ln.left
        a, b = self.nodes[chosen].free_edges, self.nodes[lchosen].free_edges
        rank = len(a) + len(b)
        for e in a:
            if e in b:
                rank -= 2
        if rank < if_below:
            if upper_left:
                un.left = lother
            else:
                un.right = lother
            if lower_left:
                ln.right = chosen
            else:
                ln.left = chosen
            self.compute_join_properties(lower)
            self.compute_join_properties(upper)
            return 1
        return 0

    def postorder_ids(self, root):
        out = []
        stack = [(root, False)]
        while stack:
            nid, expanded = stack.pop()
            n = self.nodes[nid]
            if expanded or n.is_leaf:
                out.append(nid)
            else:
                stack.append((nid, True))
                stack.append((n.right, False))
                stack.append((n.left, False))
        return out

    def estimate_cost(self, root, sliced_edges):
        costs = {}
        for nid in self.postorder_ids(root):
            costs[nid] = self._cost_at(nid, costs, sliced_edges)
            n = self.nodes[nid]
            if not n.is_leaf:
                del costs[n.left]
                del costs[n.right]
        return costs[root]

    def _cost_at(self, nid, costs, sliced):
        node = self.nodes[nid]
        r = _Cost()
        r.local_memory = 1.0
        for e in node.free_edges:
            if e not in sliced:
                r.local_memory *= 2
        if node.is_leaf:
            r.FLOPs = 0.0
            r.total_memory = r.local_memory
            r.largest_tensor = 0
            for e in node.free_edges:
                if e not in sliced:
                    r.open_edge_total_memory[e] = r.total_memory / 2
                    r.largest_tensor += 1
            r.best_edge = -1
            r.best_edge_memory = float(2 ** 64)
            return r

        left = costs[node.left]
        right = costs[node.right]
        lnode = self.nodes[node.left]
        rnode = self.nodes[node.right]

        left_cap = left.total_memory
        right_cap = left.local_memory + right.total_memory
        local_cap = r.local_memory + 2 * left.local_memory + 2 * right.local_memory
        if left_cap >= right_cap and left_cap >= local_cap:
            r.total_memory = left_cap
            r.best_edge = left.best_edge
            r.best_edge_memory = left.best_edge_memory
        elif right_cap >= local_cap:
            r.total_memory = right_cap
            r.best_edge = right.best_edge
            r.best_edge_memory = left.local_memory + right.best_edge_memory
        else:
            r.total_memory = local_cap
            r.best_edge = left.best_edge
            r.best_edge_memory = local_cap

        log_terms = int((len(lnode.free_edges) + len(rnode.free_edges) - len(node.free_edges)) / 2)
        log_sums = len(node.free_edges)
        if log_terms + log_sums > 100:
            r.FLOPs = float(2 ** 100)
        else:
            r.FLOPs = left.FLOPs + right.FLOPs + float(2 ** (log_terms + log_sums))

        r.largest_tensor = 0
        for e in node.free_edges:
            if e not in sliced:
                r.largest_tensor += 1
        r.largest_tensor = max(r.largest_tensor, left.largest_tensor, right.largest_tensor)

        lo, ro, res = left.open_edge_total_memory, right.open_edge_total_memory, r.open_edge_total_memory
        for e in lnode.free_edges:
            if e not in ro:
                res[e] = max(
                    max(_at(lo, e), (left.local_memory / 2) + right.total_memory),
                    (r.local_memory / 2) + 2 * (left.local_memory / 2) + 2 * right.local_memory,
                )
        for e in rnode.free_edges:
            if e not in lo:
                res[e] = max(
                    max(left.total_memory, left.local_memory + _at(ro, e)),
                    (r.local_memory / 2) + 2 * left.local_memory + 2 * right.local_memory,
                )
            elif e not in sliced:
                new_cap = max(
                    max(_at(lo, e), (left.local_memory / 2) + _at(ro, e)),
                    r.local_memory + 2 * (left.local_memory / 2) + 2 * (right.local_memory / 2),
                )
                if new_cap < r.best_edge_memory:
                    r.best_edge = e
                    r.best_edge_memory = new_cap
        return r


class ContractionTree:
    """A view of one node of a ``TreeContext`` (``contraction_tree.pyx:4-47``)."""

    def __init__(self, context, node):
        self._context = context
        self._node = node
        self._info = context.get_data(node)

    @property
    def context(self):
        return self._context

    @property
    def node_id(self):
        return self._node

    @property
    def left(self):
        return ContractionTree(self._context, self._info.left)

    @property
    def right(self):
        return ContractionTree(self._context, self._info.right)

    @property
    def tensor_index(self):
        return self._info.tensor_index

    @property
    def is_leaf(self):
        return self._info.is_leaf

    @property
    def maxrank(self):
        return self._info.maxrank

    @property
    def left_edge_map(self):
        return self._info.left_edge_map

    @property
    def right_edge_map(self):
        return self._info.right_edge_map

    @property
    def rank(self):
        return len(self._info.free_edges)

    @property
    def free_edges(self):
        return self._info.free_edges

    def iterate_postorder(self):
        processed = [(self, False)]
        while processed:
            node, expanded = processed.pop()
            if expanded or node.is_leaf:
                yield node
            else:
                processed.append((node, True))
                processed.append((node.right, False))
                processed.append((node.left, False))

    def reroot(self, network, large_rank):
        """A new tree whose top joins the first subtree of max rank >= large_rank with the
        remaining pieces along its path (``contraction_tree.pyx:49-89``)."""
        ctx = TreeContext()
        stack, path_up = [], []
        for node in self.iterate_postorder():
            if node.is_leaf:
                stack.append(ctx.leaf(network, node.tensor_index))
                continue
            right = stack.pop()
            left = stack.pop()
            if not path_up and node.maxrank >= large_rank:
                path_up.append(ctx.join(left, right))
                stack.append(-1)
            elif right == -1:
                path_up.append(left)
                stack.append(-1)
            elif left == -1:
                path_up.append(right)
                stack.append(-1)
            else:
                stack.append(ctx.join(left, right))
        if not path_up:
            return ctx.get_tree(stack[0])
        top = path_up[-1]
        for i in range(len(path_up) - 1, 0, -1):
            top = ctx.join(path_up[i - 1], top)
        return ctx.get_tree(top)

    def regroup(self, rank_limit):
        """Regroup (A (B C)) with A, B and (A B) small and C large into (C (A B)), in place
        (``contraction_tree.pyx:91-135``)."""
        ctx = self._context
        stack = []
        for nid in ctx.postorder_ids(self._node):
            n = ctx.nodes[nid]
            if n.is_leaf:
                stack.append((None, n.maxrank < rank_limit))
                continue
            r_sol, r_small = stack.pop()
            l_sol, l_small = stack.pop()
            if r_sol is not None and l_small:
                if ctx.group_if_below(nid, True, r_sol, rank_limit) == 1:
                    l_small, r_small = False, True
            elif l_sol is not None and r_small:
                if ctx.group_if_below(nid, False, l_sol, rank_limit) == 1:
                    l_small, r_small = True, False
            if l_small and not r_small:
                sol = True
            elif r_small and not l_small:
                sol = False
            else:
                sol = None
            stack.append((sol, n.maxrank < rank_limit))

    def sort_small(self):
        """Swap operands so the lower-rank one is on the right; recompute join
        properties wherever something below changed (``contraction_tree.pyx:137-162``)."""
        ctx = self._context
        stack = []
        for nid in ctx.postorder_ids(self._node):
            n = ctx.nodes[nid]
            if n.is_leaf:
                stack.append(0)
                continue
            changed = stack.pop()
            changed += stack.pop()
            if len(ctx.nodes[n.left].free_edges) < len(ctx.nodes[n.right].free_edges):
                n.left, n.right = n.right, n.left
                changed += 1
            if changed > 0:
                ctx.compute_join_properties(nid)
                stack.append(1)
            else:
                stack.append(0)

    def estimate_cost(self, slices=frozenset()):
        """(FLOPs, memory cap in entries, best edge to slice, max rank)."""
        c = self._context.estimate_cost(self._node, slices)
        return c.FLOPs, c.total_memory, c.best_edge, c.largest_tensor


def is_tree_complete(contraction_tree, tensor_network):
    """Every tensor appears exactly once as a leaf."""
    seen = []
    for node in contraction_tree.iterate_postorder():
        if node.is_leaf:
            seen.append(node.tensor_index)
    return sorted(seen) == list(range(len(tensor_network)))

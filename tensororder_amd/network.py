"""Tensor networks for weighted model counting: tensors, wiring and slicing.

Host-side mirror of the reference's tensor-network layer so that plans, slices and
leaf values are produced exactly as the reference produces them:

* lazy tensors and slicing -- ``tensor_network/tensor.py:1-119``;
* OR / variable (COPY) leaf tensors and the CNF reduction --
  ``tensor_network/tensor_network_constructions.py:7-159``;
* the network store and its slice enumeration -- ``tensor_network/tensor_network.pyx``
  (``connect`` :54-66, ``add_node`` :68-76, ``identify`` :142-156, ``factor_out``
  :281-325, ``slice_groups`` :363-401, ``get_tensor_slices`` :404-442,
  ``remove_sliced_indices_from`` :444-468, ``find_equivalent_edges`` :470-489).

All indices have dimension 2.  A tensor of rank r is a dense C-ordered array of shape
(2,)*r: axis 0 is the most significant bit of the flat index.
"""

from collections import Counter
from itertools import product, islice

from .formula import Formula


class Tensor:
    """A lazily built tensor of a fixed shape (``tensor.py:4-33``)."""

    def __init__(self, shape, label=None):
        self.shape = tuple(shape)
        self.label = label

    @property
    def rank(self):
        return len(self.shape)

    @property
    def display(self):
        return {} if self.label is None else {"label": self.label}

    def get_slice(self, which_edge, value):
        # The sliced axis is kept with size 1: a one-element fancy index on the first
        # slice of a tensor (``tensor.py:20-25``), a size-1 range afterwards (:70-73).
        lookup = [
            (value,) if i == which_edge else slice(0, self.shape[i])
            for i in range(len(self.shape))
        ]
        return SlicedTensor(self, lookup)

    @property
    def diagonal(self):
        return False

    def build(self, tensor_factory):
        raise NotImplementedError()

    def get_factor_components(self, left_indices, right_indices):
        raise NotImplementedError()


class BuiltTensor(Tensor):
    """A tensor wrapping an already materialised array (``tensor.py:36-52``)."""

    def __init__(self, base, label=None):
        Tensor.__init__(self, base.shape, label)
        self.base = base

    def build(self, tensor_factory):
        result = tensor_factory(self.base.shape)
        if len(self.base.shape) == 0:
            result[()] = self.base[()]
        else:
            result[:] = self.base[:]
        return result


class SlicedTensor(Tensor):
    """A parent tensor with some axes fixed to one value (``tensor.py:55-76``)."""

    def __init__(self, parent, slice_lookup):
        Tensor.__init__(self, parent.shape, parent.label)
        self._parent = parent
        self._slice_lookup = slice_lookup

    @property
    def diagonal(self):
        return self._parent.diagonal

    def build(self, tensor_factory):
        result = self._parent.build(tensor_factory)
        return result[tuple(self._slice_lookup)]

    def get_slice(self, which_edge, value):
        new_lookup = list(self._slice_lookup)
        new_lookup[which_edge] = slice(value, value + 1)
        return SlicedTensor(self._parent, new_lookup)


class SliceSequence:
    """All slices of one leaf, as used by the batched (slice-parallel) executors.

    ``reordered_tensor`` moves the sliced axes of the leaf to the front and flattens them
    into one row index, so slice ``assignment`` of the whole network selects row
    ``sum(slice_lookup[g] * assignment[g])`` of the leaf (``tensor.py:79-119``).
    """

    def __init__(self, base_tensor, full_sliced_indices, local_to_global):
        self._base_tensor = base_tensor
        self._full_sliced_indices = full_sliced_indices
        self._local_to_global = local_to_global

    def is_constant(self):
        return len(self._local_to_global) == 0

    @property
    def shape(self):
        original = list(self._base_tensor.shape)
        for local_index in self._local_to_global:
            original[local_index] = 1
        return tuple(original)

    @property
    def local_to_global(self):
        return dict(self._local_to_global)

    def reordered_tensor(self, tensor_api):
        sliced, unsliced, resulting_shape = [], [], []
        for i, size in enumerate(self._base_tensor.shape):
            if i in self._local_to_global:
                sliced.append(i)
            else:
                unsliced.append(i)
                resulting_shape.append(size)
        slice_lookup = [0] * len(self._full_sliced_indices)
        offset = 1
        for i in reversed(sliced):
            slice_lookup[self._local_to_global[i]] += offset
            offset *= self._base_tensor.shape[i]
        reordered = tensor_api.reshape(
            tensor_api.transpose(self._base_tensor, perm=sliced + unsliced),
            tuple([offset] + resulting_shape),
        )
        return reordered, slice_lookup, tuple(resulting_shape)


class OrTensor(Tensor):
    """OR gate over signed literals (``tensor_network_constructions.py:44-119``).

    Without an output index the entry is 1 unless every literal is false.  With an
    output index o, the entry is 1 iff literal o equals the OR of the other literals.
    A literal is false at index value 0 when positive and 1 when negative.
    """

    def __init__(self, literals_positive, output_index=None):
        super().__init__([2] * len(literals_positive), label="or")
        self._literals_positive = list(literals_positive)
        self._output_index = output_index

    @property
    def output_index(self):
        return self._output_index

    @property
    def literals_positive(self):
        return list(self._literals_positive)

    def build(self, tensor_factory):
        result = tensor_factory(self.shape, 1)
        all_false = [0 if pos else 1 for pos in self._literals_positive]
        if self._output_index is None:
            result[tuple(all_false)] = 0
        else:
            o = self._output_index
            out_false = all_false[o]
            result[
                tuple(out_false if i == o else slice(0, 2) for i in range(len(self.shape)))
            ] = 0
            result[tuple(all_false)] = 1
            flipped = list(all_false)
            flipped[o] = 1 - flipped[o]
            result[tuple(flipped)] = 0
        return result

    def get_factor_components(self, left_indices, right_indices):
        left_literals = [self._literals_positive[i] for i in left_indices] + [True]
        right_literals = [self._literals_positive[i] for i in right_indices] + [True]
        o = self._output_index
        if o is None:
            return OrTensor(left_literals, len(left_literals) - 1), OrTensor(right_literals)
        if o in left_indices:
            return (
                OrTensor(left_literals, left_indices.index(o)),
                OrTensor(right_literals, len(right_literals) - 1),
            )
        if o in right_indices:
            return (
                OrTensor(left_literals, len(left_literals) - 1),
                OrTensor(right_literals, right_indices.index(o)),
            )
        raise RuntimeError("Provided indices must partition the indices of this tensor")


class VariableTensor(Tensor):
    """Diagonal COPY tensor of a variable (``tensor_network_constructions.py:122-159``).

    Entry w(not x) at 0...0, w(x) at 1...1, 0 elsewhere; rank 0 gives w(not x) + w(x).
    """

    def __init__(self, rank, positive_weight, negative_weight, label=None):
        super().__init__((2,) * rank, label=label)
        self._positive_weight = positive_weight
        self._negative_weight = negative_weight

    @property
    def diagonal(self):
        return True

    @property
    def weights(self):
        """(w(not x), w(x))."""
        return self._negative_weight, self._positive_weight

    def build(self, tensor_factory):
        result = tensor_factory(self.shape, 0)
        if len(self.shape) == 0:
            result[()] = self._negative_weight + self._positive_weight
        else:
            result[(0,) * self.rank] = self._negative_weight
            result[(1,) * self.rank] = self._positive_weight
        return result

    def get_factor_components(self, left_indices, right_indices):
        left = VariableTensor(
            len(left_indices) + 1, self._positive_weight, self._negative_weight
        )
        right = VariableTensor(len(right_indices) + 1, 1, 1)
        return left, right


class TensorNetwork:
    """Tensors plus the edges (dimension-2 indices) connecting pairs of them.

    ``index_list(t)[i]`` is the edge id on axis i of tensor t; an unconnected axis holds
    a negative id.  ``copy`` shares tensor objects but copies the wiring, matching the
    value semantics of the reference's C++ vectors (``tensor_network.pyx:12-23``).
    """

    def __init__(self, base=None):
        if base is not None:
            self._nodes = list(base._nodes)
            self._edges = [list(e) for e in base._edges]
            self._index_lists = [list(x) for x in base._index_lists]
            self._disconnected_edge_id = base._disconnected_edge_id
        else:
            self._nodes = []
            self._edges = []  # [edge id, tensor1 id, tensor2 id]
            self._index_lists = []
            self._disconnected_edge_id = -1

    def copy(self):
        return TensorNetwork(self)

    @property
    def tensors(self):
        return iter(self._nodes)

    def tensor(self, tensor_id):
        return self._nodes[tensor_id]

    def __getitem__(self, tensor_id):
        return self._nodes[tensor_id]

    def __len__(self):
        return len(self._nodes)

    def index_list(self, tensor_id):
        return list(self._index_lists[tensor_id])

    @property
    def edges(self):
        return iter(
            {"id": e[0], "tensor1_id": e[1], "tensor2_id": e[2]} for e in self._edges
        )

    def edge(self, edge_index):
        e = self._edges[edge_index]
        return {"id": e[0], "tensor1_id": e[1], "tensor2_id": e[2]}

    def num_edges(self):
        return len(self._edges)

    def connected(self, tensor_id, edge_id):
        return self._index_lists[tensor_id][edge_id] >= 0

    def connect(self, tensor1, edge1, tensor2, edge2):
        if tensor1 == tensor2:
            raise ValueError("Self loops are not allowed")
        if self._index_lists[tensor1][edge1] >= 0:
            raise ValueError("Index of first tensor has already been assigned")
        if self._index_lists[tensor2][edge2] >= 0:
            raise ValueError("Index of second tensor has already been assigned")
        edge_id = len(self._edges)
        self._index_lists[tensor1][edge1] = edge_id
        self._index_lists[tensor2][edge2] = edge_id
        self._edges.append([edge_id, tensor1, tensor2])
        return edge_id

    def add_node(self, tensor):
        self._nodes.append(tensor)
        rank = tensor.rank
        self._index_lists.append([self._disconnected_edge_id - i for i in range(rank)])
        self._disconnected_edge_id -= rank
        return [(len(self._nodes) - 1, i) for i in range(rank)]

    # -- graph exports consumed by decomposition solvers -------------------------

    def structure_text(self):
        """The tensor graph in PACE ``p tw`` form, 1-based (``tensor_network.pyx:78-83``)."""
        lines = ["p tw %d %d" % (len(self._index_lists), len(self._edges))]
        lines += ["%d %d" % (e[1] + 1, e[2] + 1) for e in self._edges]
        return "\n".join(lines) + "\n"

    def line_structure_text(self):
        """The line graph (edges adjacent when they share a tensor), PACE ``p tw`` form
        (``tensor_network.pyx:98-108``)."""
        pairs = []
        for conn in self._index_lists:
            for i in range(len(conn)):
                for j in range(i + 1, len(conn)):
                    pairs.append("%d %d" % (conn[i] + 1, conn[j] + 1))
        return "\n".join(["p tw %d %d" % (len(self._edges), len(pairs))] + pairs) + "\n"

    # -- contraction --------------------------------------------------------------

    def identify(self, contraction_tree, tensor_api):
        """Contract along the tree in post-order with ``tensor_api.tensordot``
        (``tensor_network.pyx:142-156``)."""
        stack = []
        for node in contraction_tree.iterate_postorder():
            if node.is_leaf:
                stack.append(self._nodes[node.tensor_index].build(tensor_api.create_tensor))
            else:
                right = stack.pop()
                left = stack.pop()
                stack.append(
                    tensor_api.tensordot(left, right, (node.left_edge_map, node.right_edge_map))
                )
        return stack[0]

    # -- early contraction -----------------------------------------------------------

    def contract_pair(self, left_index, left_edge_map, right_index, right_edge_map, free_edges, tensor_api):
        """Contract two tensors now; the result becomes a new built tensor wired to
        ``free_edges`` (``tensor_network.pyx:158-178``)."""
        left = self._nodes[left_index].build(tensor_api.create_tensor)
        right = self._nodes[right_index].build(tensor_api.create_tensor)
        result = tensor_api.tensordot(left, right, (left_edge_map, right_edge_map))
        if hasattr(result, "numpy"):
            result = result.numpy()
        self.add_node(BuiltTensor(result))
        result_index = len(self._nodes) - 1
        for i, e in enumerate(free_edges):
            self._index_lists[result_index][i] = e
            if self._edges[e][1] == left_index or self._edges[e][1] == right_index:
                self._edges[e][1] = result_index
            elif self._edges[e][2] == left_index or self._edges[e][2] == right_index:
                self._edges[e][2] = result_index
        self._nodes[left_index] = None
        self._nodes[right_index] = None
        return result_index

    def identify_partial(self, contraction_tree, tensor_api, contract_below):
        """Contract every join whose operands are already built and whose result has rank
        < contract_below; return the remaining tree over the compacted network and the map
        old edge id -> new edge id (or a negative value) (``tensor_network.pyx:180-244``)."""
        from .tree import TreeContext

        ctx = TreeContext()
        stack = []
        for node in contraction_tree.iterate_postorder():
            if node.is_leaf:
                stack.append((True, node.tensor_index))
                continue
            right_created, right = stack.pop()
            left_created, left = stack.pop()
            if left_created and right_created and len(node.free_edges) < contract_below:
                stack.append((True, self.contract_pair(left, node.left_edge_map, right, node.right_edge_map,
                                                       node.free_edges, tensor_api)))
            else:
                if left_created:
                    left = ctx.leaf(self, left)
                if right_created:
                    right = ctx.leaf(self, right)
                stack.append((False, ctx.join(left, right)))
        top_created, top = stack.pop()
        if top_created:
            top = ctx.leaf(self, top)
        new_tid, fixed = [], 0
        for cur in range(len(self._nodes)):
            if self._nodes[cur] is not None:
                self._nodes[fixed] = self._nodes[cur]
                self._index_lists[fixed] = self._index_lists[cur]
                new_tid.append(fixed)
                fixed += 1
            else:
                new_tid.append(-1)
        self._nodes = self._nodes[:fixed]
        self._index_lists = self._index_lists[:fixed]
        new_eid, fixed_e = [], 0
        for cur in range(len(self._edges)):
            # the reference tests the first endpoint twice (tensor_network.pyx:228); edges
            # are retargeted by contract_pair, so both endpoints agree on survival
            if new_tid[self._edges[cur][1]] >= 0 and new_tid[self._edges[cur][1]] >= 0:
                self._edges[fixed_e] = [fixed_e, new_tid[self._edges[cur][1]], new_tid[self._edges[cur][2]]]
                new_eid.append(fixed_e)
                fixed_e += 1
            else:
                new_eid.append(-2)
        self._edges = self._edges[:fixed_e]
        for t in range(fixed):
            self._index_lists[t] = [new_eid[e] for e in self._index_lists[t]]
        for n in ctx.nodes:
            if n.is_leaf:
                n.tensor_index = new_tid[n.tensor_index]
            n.free_edges = [new_eid[e] for e in n.free_edges]
        return ctx.get_tree(top), new_eid

    # -- factoring (Factor-Tree planner) -----------------------------------------

    def factor_out(self, tensor_index, dim1, dim2):
        """Split axes (dim1, dim2) of a tensor into a new tensor joined to the rest by a
        fresh edge; returns (new tensor id, fresh edge id) (``tensor_network.pyx:281-325``)."""
        idx = self._index_lists[tensor_index]
        rank = len(idx)
        if rank <= 2:
            raise ValueError("Unable to factor rank 2 or smaller tensor")
        if rank == 3:
            return tensor_index, idx[3 - dim1 - dim2]
        right_indices = [j for j in range(rank) if j != dim1 and j != dim2]
        left, right = self._nodes[tensor_index].get_factor_components([dim1, dim2], right_indices)
        new_tensor_index = len(self._index_lists)
        self._nodes.append(left)
        self._nodes[tensor_index] = right
        self._index_lists.append([idx[dim1], idx[dim2]])
        self._index_lists[tensor_index] = [idx[j] for j in right_indices]
        for edge_id in self._index_lists[new_tensor_index]:
            if edge_id >= 0:
                if self._edges[edge_id][1] == tensor_index:
                    self._edges[edge_id][1] = new_tensor_index
                else:
                    self._edges[edge_id][2] = new_tensor_index
        edge_id = len(self._edges)
        self._index_lists[tensor_index].append(edge_id)
        self._index_lists[new_tensor_index].append(edge_id)
        self._edges.append([edge_id, tensor_index, new_tensor_index])
        return new_tensor_index, edge_id

    # -- slicing ------------------------------------------------------------------

    def _locate(self, e):
        """(tensor, axis) of both endpoints of edge e."""
        t1, t2 = self._edges[e][1], self._edges[e][2]
        infos = []
        found = False
        for i, x in enumerate(self._index_lists[t1]):
            if x == e:
                infos.append((t1, i))
                found = True
                break
        for i, x in enumerate(self._index_lists[t2]):
            if x == e:
                infos.append((t2, i))
                break
        if not found:
            raise RuntimeError("Unable to determine size of edge " + str(e))
        return infos

    def slice_groups(self, edge_groups):
        """Yield one network per assignment of the edge groups, in ``itertools.product``
        order (last group fastest); every axis of a group takes the group's value
        (``tensor_network.pyx:363-401``)."""
        if len(edge_groups) == 0:
            yield self
            return
        index_values, tensor_infos = [], []
        for group in edge_groups:
            if len(group) == 0:
                continue
            infos = []
            for e in group:
                infos.extend(self._locate(e))
            tensor_infos.append(infos)
            t, ax = infos[-1]
            index_values.append(list(range(self._nodes[t].shape[ax])))
        for assignment in product(*index_values):
            tn_slice = self.copy()
            for infos, value in zip(tensor_infos, assignment):
                for t, ax in infos:
                    tn_slice._nodes[t] = tn_slice._nodes[t].get_slice(ax, value)
            yield tn_slice

    def get_tensor_slices(self, edge_groups, tensor_factory):
        """Per-leaf slice sequences for batched execution (``tensor_network.pyx:404-442``).

        Returns (number of slices, one ``SliceSequence`` per tensor, iterator over
        assignments in product order)."""
        if len(edge_groups) == 0:
            return 1, [SliceSequence(t.build(tensor_factory), ([],), {}) for t in self._nodes], [[]]
        index_values = []
        lookups = [{} for _ in range(len(self._nodes))]
        total = 1
        which = 0
        for group in edge_groups:
            if len(group) == 0:
                continue
            size = 0
            for e in group:
                infos = self._locate(e)
                for k, (t, ax) in enumerate(infos):
                    lookups[t][ax] = which
                    if k == 0:
                        size = self._nodes[t].shape[ax]
            index_values.append(list(range(size)))
            total *= size
            which += 1
        return total, [
            SliceSequence(t.build(tensor_factory), index_values, lookup)
            for t, lookup in zip(self._nodes, lookups)
        ], product(*index_values)

    def slice_lookups(self, edge_groups):
        """(number of slices, per tensor {axis: slice group}) as ``get_tensor_slices``
        finds them, without building any tensor."""
        lookups = [{} for _ in range(len(self._nodes))]
        total, which = 1, 0
        for group in edge_groups:
            if len(group) == 0:
                continue
            size = 0
            for e in group:
                for k, (t, ax) in enumerate(self._locate(e)):
                    lookups[t][ax] = which
                    if k == 0:
                        size = self._nodes[t].shape[ax]
            total *= size
            which += 1
        return total, lookups

    def remove_sliced_indices_from(self, contraction_tree, edge_groups):
        """A copy of the tree whose leaves omit sliced edges (``tensor_network.pyx:444-468``)."""
        from .tree import TreeContext

        sliced = set()
        for group in edge_groups:
            sliced.update(group)
        ctx = TreeContext()
        stack = []
        for node in contraction_tree.iterate_postorder():
            if node.is_leaf:
                free = [e for e in self._index_lists[node.tensor_index] if e >= 0 and e not in sliced]
                stack.append(ctx.leaf_manual(node.tensor_index, free))
            else:
                right = stack.pop()
                left = stack.pop()
                stack.append(ctx.join(left, right))
        return ctx.get_tree(stack.pop())

    def find_equivalent_edges(self, edge):
        """All edges forced equal to ``edge`` through diagonal tensors
        (``tensor_network.pyx:470-489``)."""
        to_process = [edge]
        seen = {edge}
        while to_process:
            e = to_process.pop()
            for t in (self._edges[e][1], self._edges[e][2]):
                if self._nodes[t].diagonal:
                    for other in self._index_lists[t]:
                        if other not in seen:
                            to_process.append(other)
                            seen.add(other)
        return seen

    def equivalent_edge_sets(self):
        rep = {}
        for e in range(self.num_edges()):
            if e not in rep:
                for f in self.find_equivalent_edges(e):
                    rep[f] = e
        return rep


def cnf_count(formula_or_lines, weight_format=None):
    """Reduce a CNF to a tensor network whose full contraction is its weighted model
    count (``tensor_network_constructions.py:7-41``).

    One diagonal ``VariableTensor`` per variable (rank = number of occurrences) in
    ``Formula.variables`` order, then one ``OrTensor`` per clause.  A clause's k-th
    literal is wired to the last still-unconnected axis of its variable, so a
    variable's last axis meets the first clause it occurs in.
    """
    if isinstance(formula_or_lines, Formula):
        formula = formula_or_lines
    else:
        formula = Formula.parse_DIMACS(formula_or_lines, weight_format)
    network = TensorNetwork()
    count = Counter()
    for clause in formula.clauses:
        count.update(map(abs, clause))
    variable_edges = {}
    for var in formula.variables:
        variable_edges[var] = network.add_node(
            VariableTensor(count[var], formula.literal_weight(var), formula.literal_weight(-var))
        )
    for clause in formula.clauses:
        clause_edges = network.add_node(OrTensor([lit > 0 for lit in clause]))
        for clause_edge, lit in zip(clause_edges, clause):
            network.connect(*clause_edge, *variable_edges[abs(lit)].pop())
    return network


ALL_CONSTRUCTIONS = {"wmc": cnf_count}

"""A neutral on-disk plan format (JSON) replacing the reference's pickled ``.con`` files.

The reference stores plans as pickles of Cython objects (``planning.py:108-114``, loaded
at ``execution.py:95-96`` and ``slicing.py:71-72``), readable only with the reference
importable.  This format carries everything the executor needs, as plain data:

``tensors``        leaf descriptors: ``{"kind": "var", "rank", "w_neg", "w_pos"}``,
                   ``{"kind": "or", "literals_positive", "output_index"}`` or
                   ``{"kind": "dense", "shape", "values"}`` (C-order, float64)
``index_lists``    edge id on each axis of each tensor
``edges``          ``[tensor1, tensor2]`` per edge id
``tree``           post-order token list: a tensor id pushes a leaf, ``-1`` joins the top
                   two entries (left = deeper of the two)
``groups_to_slice`` list of edge-id groups, in slicing order
Floats are written with ``repr`` and so round-trip exactly.
"""

import json

import numpy as np

from .network import BuiltTensor, OrTensor, TensorNetwork, VariableTensor
from .plan import SlicedExecutionPlan
from .tree import TreeContext

FORMAT = "tensororder-plan/1"


def _tensor_desc(t):
    if isinstance(t, VariableTensor):
        w_neg, w_pos = t.weights
        return {"kind": "var", "rank": t.rank, "w_neg": float(w_neg), "w_pos": float(w_pos)}
    if isinstance(t, OrTensor):
        return {
            "kind": "or",
            "literals_positive": [bool(x) for x in t.literals_positive],
            "output_index": t.output_index,
        }
    arr = np.asarray(t.build(lambda shape, d=None: np.full(shape, 0 if d is None else d, dtype=np.float64)))
    return {"kind": "dense", "shape": list(arr.shape), "values": [float(x) for x in arr.ravel()]}


def _tensor_from(d):
    if d["kind"] == "var":
        return VariableTensor(d["rank"], d["w_pos"], d["w_neg"])
    if d["kind"] == "or":
        return OrTensor(d["literals_positive"], d["output_index"])
    if d["kind"] == "dense":
        return BuiltTensor(np.asarray(d["values"], dtype=np.float64).reshape(d["shape"]))
    raise ValueError("unknown tensor kind " + str(d["kind"]))


def tree_tokens(tree):
    return [n.tensor_index if n.is_leaf else -1 for n in tree.iterate_postorder()]


def tree_from_tokens(tokens, network):
    ctx = TreeContext()
    stack = []
    for tok in tokens:
        if tok >= 0:
            stack.append(ctx.leaf(network, tok))
        else:
            right = stack.pop()
            left = stack.pop()
            stack.append(ctx.join(left, right))
    if len(stack) != 1:
        raise ValueError("malformed tree token list")
    return ctx.get_tree(stack[0])


def plan_to_dict(plan, meta=None):
    net = plan.network
    return {
        "format": FORMAT,
        "tensors": [_tensor_desc(net.tensor(i)) for i in range(len(net))],
        "index_lists": [net.index_list(i) for i in range(len(net))],
        "edges": [[e[1], e[2]] for e in net._edges],
        "tree": tree_tokens(plan.tree),
        "groups_to_slice": [sorted(int(e) for e in g) for g in plan.groups_to_slice],
        "meta": dict(meta or {}),
    }


def network_from_dict(d):
    net = TensorNetwork()
    for t in d["tensors"]:
        net._nodes.append(_tensor_from(t))
    net._index_lists = [list(x) for x in d["index_lists"]]
    net._edges = [[i, e[0], e[1]] for i, e in enumerate(d["edges"])]
    neg = [x for lst in net._index_lists for x in lst if x < 0]
    net._disconnected_edge_id = min(neg) - 1 if neg else -1
    return net


def plan_from_dict(d):
    if d.get("format") != FORMAT:
        raise ValueError("not a %s document" % FORMAT)
    net = network_from_dict(d)
    tree = tree_from_tokens(d["tree"], net)
    plan = SlicedExecutionPlan(tree, net)
    for group in d["groups_to_slice"]:
        g = set(group)
        plan.groups_to_slice.append(g)
        plan.edges_to_slice |= g
    if d["groups_to_slice"]:
        plan.FLOPs, plan.memory, plan.next_edge_to_slice, plan.maxrank = plan.tree.estimate_cost(
            plan.edges_to_slice
        )
    plan.meta = dict(d.get("meta", {}))
    return plan


def save_plan(plan, path, meta=None):
    with open(path, "w") as f:
        json.dump(plan_to_dict(plan, meta), f, separators=(",", ":"))


def load_plan(path):
    with open(path) as f:
        return plan_from_dict(json.load(f))


BATCH_FORMAT = "tensororder-batch/1"


def batch_plans(doc):
    """Plans of a ``tensororder-batch/1`` document: weighted vertex-cover instances, each
    rebuilt from (n, graph seed) and its stored line-graph decomposition.  Returns
    (plans, expected float64 counts)."""
    import base64
    import zlib

    from . import decomposition, network, planners
    from .formula import parse_dimacs_text
    from .instances import vertex_cover_cnf
    from .plan import SlicedExecutionPlan

    if doc.get("format") != BATCH_FORMAT:
        raise ValueError("not a %s document" % BATCH_FORMAT)
    plans, expected = [], []
    for row in doc["instances"]:
        net = network.cnf_count(parse_dimacs_text(vertex_cover_cnf(row["n"], row["seed"], True), doc["weights"]))
        text = zlib.decompress(base64.b64decode(row["decomposition"])).decode()
        td = decomposition.parse_all(text)[0]
        tree, nn = planners.plan_line(net, td)
        tree.treewidth = td.width()
        plans.append(SlicedExecutionPlan(tree, nn))
        expected.append(row["expected"])
    return plans, expected

"""Tree decompositions: PACE ``.td`` parsing, the tree surgery the planners need, a
solver subprocess driver and a built-in greedy min-degree decomposer.

Mirror of ``decompositions/tree_decomposition.pyx`` (``extend_leaf`` :30-60,
``split_high_degree_nodes`` :62-94, ``traverse_postorder`` :96-127, ``width`` :129-138,
``parse`` :140-193) and of the solver driver in
``decompositions/decomposition_solver.py:12-64``.  Bags hold 0-based vertex ids.
"""

import subprocess


class TreeDecomposition:
    def __init__(self):
        self.bags = []
        self.neighbors = []

    def add_node(self, bag):
        self.bags.append(list(bag))
        self.neighbors.append([])
        return len(self.neighbors) - 1

    def add_edge(self, src, dst):
        self.neighbors[src].append(dst)
        self.neighbors[dst].append(src)

    def extend_leaf(self, node):
        """Attach a new leaf (copy of ``node``'s bag) next to ``node`` without changing the
        degree of ``node`` or its neighbours, unless the decomposition has one node."""
        new_leaf = self.add_node(self.bags[node])
        if new_leaf == 1:
            self.neighbors[node].append(new_leaf)
            self.neighbors[new_leaf].append(node)
            return new_leaf
        neighbor = self.neighbors[node][0]
        new_internal = self.add_node(self.bags[node])
        self.neighbors[node][0] = new_internal
        self.neighbors[new_internal].append(node)
        self.neighbors[new_internal].append(neighbor)
        nb = self.neighbors[neighbor]
        for i in range(len(nb)):
            if nb[i] == node:
                nb[i] = new_internal
                break
        self.add_edge(new_internal, new_leaf)
        return new_leaf

    def split_high_degree_nodes(self):
        """Rewire so that no node has degree above 3."""
        i = 0
        while i < len(self.neighbors):
            while len(self.neighbors[i]) > 3:
                n1 = self.neighbors[i].pop()
                n2 = self.neighbors[i].pop()
                new_internal = self.add_node(self.bags[i])
                self.neighbors[i].append(new_internal)
                self.neighbors[new_internal] += [i, n1, n2]
                for nb in (n1, n2):
                    lst = self.neighbors[nb]
                    for j in range(len(lst)):
                        if lst[j] == i:
                            lst[j] = new_internal
                            break
            i += 1

    def traverse_postorder(self, root, func, *args):
        """result(node) = func(node, [result(child) ...], *args), children in
        neighbour order, computed without recursion."""
        results = {}
        processed = [(root, None, False)]
        while processed:
            node, parent, expanded = processed.pop()
            base = []
            if node != root and len(self.neighbors[node]) == 1:
                results[node] = func(node, base, *args)
            elif expanded:
                for n in self.neighbors[node]:
                    if n != parent:
                        base.append(results[n])
                results[node] = func(node, base, *args)
                for n in self.neighbors[node]:
                    if n != parent:
                        del results[n]
            else:
                processed.append((node, parent, True))
                for child in self.neighbors[node]:
                    if child != parent:
                        processed.append((child, node, False))
        return results[root]

    def width(self):
        return max((len(b) for b in self.bags), default=0) - 1

    def to_text(self, num_vertices):
        lines = ["s td %d %d %d" % (len(self.bags), self.width() + 1, num_vertices)]
        for i, bag in enumerate(self.bags):
            lines.append(("b %d %s" % (i + 1, " ".join(str(v + 1) for v in bag))).rstrip())
        seen = set()
        for u, nbs in enumerate(self.neighbors):
            for v in nbs:
                if (v, u) not in seen:
                    seen.add((u, v))
                    lines.append("%d %d" % (u + 1, v + 1))
        return "\n".join(lines) + "\n"

    @staticmethod
    def parse(lines):
        """Parse one decomposition from an iterator of ``.td`` lines; ``c`` comments are
        skipped and a ``=`` line or EOF ends it.  Returns None at EOF before a header."""
        it = iter(lines)
        header = None
        for line in it:
            if not line.strip() or line[0] in "cO=":
                continue
            if line[0] != "s":
                raise RuntimeError("Unexpected line prefix in: " + line)
            header = line.split()
            break
        if header is None:
            return None
        if header[1] != "td":
            raise RuntimeError("Decomposition header unknown: " + str(header))
        numbags, bag_size = int(header[2]), int(header[3])
        td = TreeDecomposition()
        for _ in range(numbags):
            td.add_node([])
        read = 0
        for line in it:
            if not line.strip() or line[0] in "cO":
                continue
            if line[0] != "b":
                raise RuntimeError("Unexpected line prefix in: " + line)
            bag = line.split()
            node = int(bag[1]) - 1
            if len(bag) - 2 > bag_size:
                raise RuntimeError("Bag larger than identified width")
            if 0 <= node < numbags:
                td.bags[node] = [int(b) - 1 for b in bag[2:]]
            read += 1
            if read == numbags:
                break
        if read < numbags:
            raise RuntimeError("EOF reached while reading bags")
        for line in it:
            if not line.strip() or line[0] in "cO":
                continue
            if line[0] == "=":
                break
            parts = line.split()
            td.add_edge(int(parts[0]) - 1, int(parts[1]) - 1)
        return td


def parse_all(text):
    """Every decomposition in a solver's ``=``-separated output stream."""
    it = iter(text.splitlines(keepends=True))
    out = []
    while True:
        td = TreeDecomposition.parse(it)
        if td is None:
            return out
        out.append(td)


def run_solver(command, graph_text, max_decompositions=1, time_limit=10.0):
    """Run an anytime PACE solver on ``graph_text`` (stdin) and collect up to
    ``max_decompositions`` ``=``-terminated decompositions from its stdout
    (``decomposition_solver.py:12-46``).  ``command`` is an argv list.

    Anytime solvers print only improved decompositions, so the wait is bounded by
    ``time_limit`` seconds; the solver is then stopped with SIGTERM, on which FlowCutter
    prints its best decomposition once more.  The first decomposition (FlowCutter's
    greedy min-degree order on small graphs) arrives without waiting.
    """
    import os
    import select
    import time

    proc = subprocess.Popen(command, stdin=subprocess.PIPE, stdout=subprocess.PIPE)
    proc.stdin.write(graph_text.encode())
    proc.stdin.close()
    fd = proc.stdout.fileno()
    buf = b""
    deadline = time.monotonic() + time_limit
    try:
        while buf.count(b"\n=") < max_decompositions:
            left = deadline - time.monotonic()
            if left <= 0:
                break
            ready, _, _ = select.select([fd], [], [], left)
            if not ready:
                break
            data = os.read(fd, 1 << 16)
            if not data:
                break
            buf += data
        if buf.count(b"\n=") < max_decompositions and proc.poll() is None:
            proc.terminate()
            end = time.monotonic() + 5.0
            while time.monotonic() < end:
                ready, _, _ = select.select([fd], [], [], max(0.0, end - time.monotonic()))
                if not ready:
                    break
                data = os.read(fd, 1 << 16)
                if not data:
                    break
                buf += data
    finally:
        if proc.poll() is None:
            proc.kill()
        proc.wait()
    text = buf.decode()
    # keep only complete '='-terminated decompositions, plus a final one printed on SIGTERM
    return parse_all(text)[:max_decompositions] if max_decompositions else parse_all(text)


def parse_graph(text):
    """(number of vertices, list of 0-based edges) from PACE ``p tw`` text."""
    n, edges = 0, []
    for line in text.splitlines():
        if not line or line[0] == "c":
            continue
        parts = line.split()
        if parts[0] == "p":
            n = int(parts[2])
        else:
            edges.append((int(parts[0]) - 1, int(parts[1]) - 1))
    return n, edges


def min_degree_decomposition(num_vertices, edges, seed=None):
    """Tree decomposition from greedy minimum-degree elimination (ties: smallest id, or a
    seeded random order of the vertices when ``seed`` is given).

    Eliminating v creates bag {v} + N(v) and makes N(v) a clique; v's bag is attached to
    the bag of whichever neighbour is eliminated next.  Components are chained so the
    result is one tree.  Every vertex of the graph appears in some bag.
    """
    adj = [set() for _ in range(num_vertices)]
    for u, v in edges:
        if u != v:
            adj[u].add(v)
            adj[v].add(u)
    import heapq

    if seed is None:
        tie = list(range(num_vertices))
    else:
        import random

        tie = list(range(num_vertices))
        random.Random(seed).shuffle(tie)
    heap = [(len(adj[v]), tie[v], v) for v in range(num_vertices)]
    heapq.heapify(heap)
    eliminated = [False] * num_vertices
    order_pos = [0] * num_vertices
    bags, bag_nbrs = [], []
    pos = 0
    while heap:
        d, _, v = heapq.heappop(heap)
        if eliminated[v] or d != len(adj[v]):
            continue
        eliminated[v] = True
        order_pos[v] = pos
        pos += 1
        nb = sorted(adj[v])
        bags.append([v] + nb)
        bag_nbrs.append(nb)
        for a in nb:
            adj[a].discard(v)
        for i in range(len(nb)):
            for j in range(i + 1, len(nb)):
                adj[nb[i]].add(nb[j])
                adj[nb[j]].add(nb[i])
        for a in nb:
            heapq.heappush(heap, (len(adj[a]), tie[a], a))
    td = TreeDecomposition()
    for b in bags:
        td.add_node(b)
    roots = []
    for i, nb in enumerate(bag_nbrs):
        if nb:
            parent = min(order_pos[a] for a in nb)
            td.add_edge(i, parent)
        else:
            roots.append(i)
    for a, b in zip(roots, roots[1:]):
        td.add_edge(a, b)
    if len(td.bags) == 1:
        td.add_node(td.bags[0])
        td.add_edge(0, 1)
    return td

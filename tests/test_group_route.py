"""CPU check of the lane routing rule of the grouped sparse kernels (`group_route` in
thea_amd/csrc/query.hip): for every G = 2^lg, the 2^2 routes of each of the G lanes of a group reach
every element range of a binary tree exactly once, on random unbalanced trees. Route pth of lane gl
takes child slot bit j of (pth | gl << 2) at depth j for depths 0 .. 2 + lg - 1; a range met on the
way belongs to the route whose bits above j are zero; a route that ends at an internal node owns
that node's subtree.
"""
import random


def _tree(n_ranges, rng):
    """Random binary tree with n_ranges leaves: nodes[i] = (child0, child1), ~r for range r."""
    nodes, next_range = [], [0]

    def build(n):
        if n == 1:
            r = next_range[0]
            next_range[0] += 1
            return ~r
        i = len(nodes)
        nodes.append(None)
        k = rng.randint(1, n - 1)
        nodes[i] = (build(k), build(n - k))
        return i

    build(n_ranges)
    return nodes


def _ranges_under(nodes, ref, out):
    if ref < 0:
        out.append(~ref)
        return
    for c in nodes[ref]:
        _ranges_under(nodes, c, out)


def _lane_ranges(nodes, lg, gl, free=2):
    got = []
    depth = free + lg
    for pth in range(1 << free):
        route = pth | (gl << free)
        node = 0
        for j in range(depth):
            ref = nodes[node][(route >> j) & 1]
            if ref < 0:
                if (route >> (j + 1)) == 0:
                    got.append(~ref)
                break
            if j + 1 == depth:
                _ranges_under(nodes, ref, got)
                break
            node = ref
    return got


def test_routes_partition_the_tree():
    rng = random.Random(11)
    for _ in range(1500):
        n = rng.randint(2, 200)
        nodes = _tree(n, rng)
        for lg in range(7):
            seen = []
            for gl in range(1 << lg):
                seen += _lane_ranges(nodes, lg, gl)
            assert sorted(seen) == list(range(n)), (n, lg)

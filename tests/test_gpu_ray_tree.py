"""GPU: the 4-wide quantized ray tree against the binary BVH it is folded from.

Every binary node at even depth must appear as one 64-byte node whose children are its
grandchildren through inner children and its leaf children (empty leaves dropped), in order, with
the same references (inner ones renumbered) and leaf counts; each child's box, decoded as
fmaf(q, 2^e, origin), must contain the binary tree's exact box for it and be at most two grid steps
looser per side; the node's grid must span it and be no coarser than twice its extent needs. The
ray kernels' parity
with brute force is in test_gpu_knn_ray_sample.py; this pins the structure they walk.
"""
import numpy as np
import pytest

from conftest import random_mesh

pytestmark = pytest.mark.gpu


def _binary(b):
    F, I, _ = b.export()
    n = len(F)
    parent = -np.ones(n, np.int64)
    for c in (0, 1):
        r = I[:, 12 + c]
        inner = r >= 0
        parent[r[inner]] = np.nonzero(inner)[0]
    depth = np.zeros(n, np.int64)
    for i in range(n):
        p = parent[i]
        while p >= 0:
            depth[i] += 1
            p = parent[p]
    return F, I, depth


def _child(F, I, i, c):
    return F[i, 6 * c:6 * c + 3], F[i, 6 * c + 3:6 * c + 6], int(I[i, 12 + c]), int(I[i, 14 + c])


def _decode(q, e, origin):
    """fmaf(q, 2^e, origin) in float32: q 2^e is exact, and the float64 sum rounds once more."""
    return (np.float64(q) * np.ldexp(1.0, e) + np.float64(origin)).astype(np.float32)


def _check(b):
    F, I, depth = _binary(b)
    Q = b.exportRayTree()
    even = np.nonzero(depth % 2 == 0)[0]
    assert len(Q) == len(even), f"{len(Q)} ray-tree nodes for {len(even)} even-depth binary nodes"
    qidx = -np.ones(len(F), np.int64)
    qidx[even] = np.arange(len(even))
    Qf = Q.view(np.float32)
    Qu = Q.view(np.uint32)
    leaves_q, leaves_b = [], []
    for c in (0, 1):
        r, cnt = I[:, 12 + c], I[:, 14 + c]
        leaves_b += [(int(x), int(y)) for x, y in zip(r[(r < 0) & (cnt > 0)], cnt[(r < 0) & (cnt > 0)])]
    for i in even:
        want = []
        for c in (0, 1):
            lo, hi, r, cnt = _child(F, I, i, c)
            if r >= 0:
                for g in (0, 1):
                    glo, ghi, gr, gc = _child(F, I, r, g)
                    if gr >= 0:
                        want.append((glo, ghi, int(qidx[gr]), 0))
                    elif gc > 0:
                        want.append((glo, ghi, gr, gc))
            elif cnt > 0:
                want.append((lo, hi, r, cnt))
        row = Q[qidx[i]]
        meta = int(Qu[qidx[i], 3])
        nc = meta >> 24
        e = np.array([(meta >> (8 * a) & 0xFF) - 127 for a in range(3)])
        origin = Qf[qidx[i], 0:3]
        assert nc == len(want), f"node {i}: {nc} children, expected {len(want)}"
        refs = row[10:14]
        cnts = [int(Qu[qidx[i], 14]) & 0xFFFF, int(Qu[qidx[i], 14]) >> 16,
                int(Qu[qidx[i], 15]) & 0xFFFF, int(Qu[qidx[i], 15]) >> 16]
        span_lo = np.min([w[0] for w in want], axis=0) if want else np.zeros(3, np.float32)
        span_hi = np.max([w[1] for w in want], axis=0) if want else np.zeros(3, np.float32)
        assert np.array_equal(origin, span_lo), f"node {i}: origin {origin} vs {span_lo}"
        for a in range(3):
            # 255 steps reach the node's far side, and the grid is not coarser than needed (the
            # step at most doubles the extent / 255, plus one step for the decode's rounding)
            assert _decode(255, e[a], origin[a]) >= span_hi[a]
            ext = float(span_hi[a]) - float(origin[a])
            if e[a] > -126:
                assert 255 * np.ldexp(1.0, int(e[a])) <= 2 * ext + np.ldexp(1.0, int(e[a])), \
                    f"node {i} axis {a}: exponent {e[a]} for extent {ext}"
        for k, (lo, hi, r, cnt) in enumerate(want):
            assert refs[k] == r, f"node {i} child {k}: ref {refs[k]} expected {r}"
            assert cnts[k] == cnt, f"node {i} child {k}: count {cnts[k]} expected {cnt}"
            if r < 0:
                leaves_q.append((r, cnt))
            for a in range(3):
                ql = (int(Qu[qidx[i], 4 + a]) >> (8 * k)) & 0xFF
                qh = (int(Qu[qidx[i], 7 + a]) >> (8 * k)) & 0xFF
                dlo, dhi = _decode(ql, e[a], origin[a]), _decode(qh, e[a], origin[a])
                assert dlo <= lo[a] and dhi >= hi[a], f"node {i} child {k} axis {a}: [{dlo}, {dhi}] misses [{lo[a]}, {hi[a]}]"
                step = np.ldexp(1.0, int(e[a]))
                assert float(lo[a]) - float(dlo) <= 2 * step and float(dhi) - float(hi[a]) <= 2 * step
        for k in range(len(want), 4):
            assert refs[k] == np.iinfo(np.int32).min
    assert sorted(leaves_q) == sorted(leaves_b), "leaf sets differ"


@pytest.mark.parametrize("leaf", [1, 2, 4, 8])
def test_ray_tree_random_mesh(leaf):
    import thea_amd
    g = np.random.default_rng(40 + leaf)
    v, t = random_mesh(g, 5000, size=0.05)
    print(f"case ray tree random leaf={leaf}", flush=True)
    _check(thea_amd.MeshBvh(thea_amd.Mesh(v, t), leaf_size=leaf))


def test_ray_tree_flat_and_far_meshes():
    """A mesh in the plane z = 0 (zero extent: the smallest exponent) and one far from the origin
    (large coordinates, small extents)."""
    import thea_amd
    from thea_amd import workloads as W
    g = np.random.default_rng(7)
    v, t = random_mesh(g, 3000, size=0.05)
    v[:, 2] = 0.0
    print("case ray tree flat", flush=True)
    _check(thea_amd.MeshBvh(thea_amd.Mesh(v, t), leaf_size=2))
    sv, st = W.uv_sphere(20, 40)
    sv = (sv * 0.01 + np.array([1.0e5, -2.5e4, 3.0e3])).astype(np.float32)
    print("case ray tree far", flush=True)
    _check(thea_amd.MeshBvh(thea_amd.Mesh(sv, st.astype(np.int32)), leaf_size=1))


def test_ray_tree_single_leaf_and_points():
    import thea_amd
    g = np.random.default_rng(8)
    v, t = random_mesh(g, 3, size=0.2)
    print("case ray tree single leaf", flush=True)
    b = thea_amd.MeshBvh(thea_amd.Mesh(v, t), leaf_size=4)
    Q = b.exportRayTree()
    assert len(Q) == 1
    _check(b)
    pb = thea_amd.PointBvh(g.uniform(-1, 1, size=(100, 3)).astype(np.float32))
    assert len(pb.exportRayTree()) == 0

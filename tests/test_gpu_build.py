"""GPU tests of the BVH build: the exported tree is a partition of the elements into leaves, and
every child box is exactly the union of what it holds (tight, so it both contains its elements and
adds nothing). Covers the span-local and straddling parts of the refit (trees larger than one
1024-element span, many spans, duplicate Morton codes) and the serial atomic refit it replaces."""
import os

import numpy as np
import pytest

from conftest import random_mesh

pytestmark = pytest.mark.gpu


def _check_tree(bvh, elo, ehi):
    f, i, ids = bvh.export()
    n = len(elo)
    assert np.array_equal(np.sort(ids), np.arange(n)), "leaf order is not a permutation of the elements"
    elo, ehi = elo[ids], ehi[ids]
    box = [f[:, 0:6], f[:, 6:12]]
    ref = [i[:, 12], i[:, 13]]
    cnt = [i[:, 14], i[:, 15]]
    leaf_start, leaf_cnt, leaf_box = [], [], []
    for c in range(2):
        leaf = ref[c] < 0
        internal = ~leaf
        # internal child: its box is the union of that node's two child boxes
        ch = ref[c][internal]
        assert np.all((ch > 0) & (ch < len(f)))
        u_lo = np.minimum(f[ch, 0:3], f[ch, 6:9])
        u_hi = np.maximum(f[ch, 3:6], f[ch, 9:12])
        assert np.array_equal(box[c][internal, 0:3], u_lo)
        assert np.array_equal(box[c][internal, 3:6], u_hi)
        leaf_start.append(~ref[c][leaf])
        leaf_cnt.append(cnt[c][leaf])
        leaf_box.append(box[c][leaf])
    st = np.concatenate(leaf_start)
    ct = np.concatenate(leaf_cnt)
    bx = np.concatenate(leaf_box)
    o = np.argsort(st, kind="stable")
    st, ct, bx = st[o], ct[o], bx[o]
    assert st[0] == 0 and np.array_equal(st[1:], (st + ct)[:-1]) and st[-1] + ct[-1] == n, "leaves do not tile"
    assert np.all(ct >= 1) and np.all(ct <= bvh.info["leaf_size"])
    assert bvh.info["leaves"] == len(st), (bvh.info["leaves"], len(st))
    assert np.array_equal(bx[:, 0:3], np.minimum.reduceat(elo, st, axis=0))
    assert np.array_equal(bx[:, 3:6], np.maximum.reduceat(ehi, st, axis=0))
    # every internal node except the root is referenced exactly once
    refs = np.concatenate([ref[0][ref[0] >= 0], ref[1][ref[1] >= 0]])
    assert np.array_equal(np.sort(refs), np.arange(1, len(f)))
    # max_depth: nodes on the path from the root to the deepest leaf (the root counts 1), level by level
    depth, level, frontier = 0, 1, np.array([0])
    while len(frontier):
        if np.any(ref[0][frontier] < 0) or np.any(ref[1][frontier] < 0):
            depth = level
        kids = np.concatenate([ref[0][frontier], ref[1][frontier]])
        frontier = kids[kids >= 0]
        level += 1
    assert bvh.info["max_depth"] == depth, (bvh.info["max_depth"], depth)


def _tri_boxes(v, t):
    p = v[t]
    return p.min(axis=1), p.max(axis=1)


@pytest.mark.parametrize("nt,leaf", [(5, 4), (1000, 4), (1025, 2), (70_000, 4), (300_000, 8), (300_000, 1)])
def test_mesh_tree_exact(nt, leaf):
    import thea_amd
    print(f"case build mesh nt={nt} leaf={leaf}", flush=True)
    g = np.random.default_rng(nt + leaf)
    v, t = random_mesh(g, nt, size=0.02)
    b = thea_amd.MeshBvh(thea_amd.Mesh(v, t), leaf_size=leaf)
    elo, ehi = _tri_boxes(v, t)
    _check_tree(b, elo, ehi)


def test_point_tree_with_duplicates():
    import thea_amd
    print("case build points with duplicates", flush=True)
    g = np.random.default_rng(7)
    p = g.random((200_000, 3)).astype(np.float32)
    p[100_000:150_000] = p[0]  # 50k copies of one point: equal Morton codes
    p[150_000:] = np.round(p[150_000:] * 8) / 8  # a coarse lattice: many equal codes
    b = thea_amd.PointBvh(p, leaf_size=4)
    _check_tree(b, p, p)


def test_c3_tree_exact_and_refits_agree():
    import subprocess
    import sys
    import thea_amd
    from thea_amd import workloads as W
    print("case build C3 torus", flush=True)
    v, t, _, _, _ = W.torus_c3(n_queries=0, n_rays=0)
    b = thea_amd.MeshBvh(thea_amd.Mesh(v, t))
    elo, ehi = _tri_boxes(v, t)
    _check_tree(b, elo, ehi)
    # the serial atomic refit (THG_REFIT=1, read once per process) builds the same tree
    code = ("import sys, numpy as np, thea_amd; from thea_amd import workloads as W;"
            "v,t,_,_,_=W.torus_c3(n_queries=0,n_rays=0); b=thea_amd.MeshBvh(thea_amd.Mesh(v,t));"
            "f,i,ids=b.export(); np.save(sys.stdout.buffer, i, allow_pickle=False)")
    env = dict(os.environ, THG_REFIT="1")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, timeout=120, check=True,
                         cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__)))).stdout
    import io
    other = np.load(io.BytesIO(out), allow_pickle=False)
    f, i, _ = b.export()
    assert np.array_equal(other, i)


def test_level_refit_large_point_tree():
    """Over 4M elements (4096 spans of 1024) the refit finishes the top tree level by level, one
    launch per level; clustered and duplicated points make the tree unbalanced and deep."""
    import thea_amd
    print("case build 4.6M points (level refit)", flush=True)
    g = np.random.default_rng(23)
    n = 4_600_000
    p = g.random((n, 3)).astype(np.float32)
    p[: n // 4] = (p[: n // 4] * 1e-3 + 0.5).astype(np.float32)  # a dense cluster
    p[n // 4 : n // 4 + 200_000] = p[0]  # one point repeated: equal Morton codes
    b = thea_amd.PointBvh(p, leaf_size=4)
    _check_tree(b, p, p)

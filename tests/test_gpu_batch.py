"""GPU tests of the batched many-mesh plan: every mesh's closest-point results equal the single-mesh
entry point bit for bit; graph replays with new inputs stay exact; autotuned leaf sizes are valid."""
import numpy as np
import pytest

from conftest import random_mesh

pytestmark = pytest.mark.gpu


def _batch(seed, sizes, qsizes):
    from thea_amd import workloads as W
    g = np.random.default_rng(seed)
    meshes, queries = [], []
    for i, (nt, nq) in enumerate(zip(sizes, qsizes)):
        if nt == 0:
            v, t = np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32)
        elif i % 3 == 0 and nt >= 64:
            v, t = W.uv_sphere(max(2, int(np.sqrt(nt / 4))), 2 * max(2, int(np.sqrt(nt / 4))))
            v = (v * g.uniform(0.5, 2) + g.uniform(-5, 5, size=3)).astype(np.float32)
        else:
            v, t = random_mesh(g, nt, spread=1.0, size=0.05)
        meshes.append((v.astype(np.float32), t.astype(np.int32)))
        lo = v.min(0) - 0.3 if len(v) else np.zeros(3)
        hi = v.max(0) + 0.3 if len(v) else np.ones(3)
        queries.append((lo + (hi - lo) * g.random((nq, 3))).astype(np.float32))
    return meshes, queries


def _check(meshes, queries, idx, dist):
    import thea_amd
    off = 0
    for (v, t), q in zip(meshes, queries):
        n = len(q)
        if len(t) == 0:
            assert np.all(idx[off:off + n] == -1)
        elif n:
            b = thea_amd.MeshBvh(thea_amd.Mesh(v, t))
            ri, rd = b.closestElementWithDistance(q)
            assert np.array_equal(idx[off:off + n], ri), f"mesh with {len(t)} triangles"
            assert np.array_equal(dist[off:off + n].view(np.uint32), rd.view(np.uint32))
        off += n


def test_batch_matches_single_mesh_and_replays():
    import torch
    from thea_amd.batch import BatchPlan, concat
    sizes = [0, 1, 2, 3, 5, 17, 100, 1000, 5000, 20000, 4, 60000, 7, 300]
    qsizes = [10, 100, 0, 50, 300, 1000, 2000, 5000, 20000, 30000, 64, 40000, 65, 1]
    print("case batch plan", flush=True)
    meshes, queries = _batch(1, sizes, qsizes)
    nv, nt, nq, verts, tris, qs = concat(meshes, queries)
    V, T, Q = (torch.from_numpy(x).cuda() for x in (verts, tris, qs))
    leaf = np.array([4, 4, 4, 4, 4, 4, 2, 8, 16, 4, 4, 6, 1, 12], np.int32)
    plan = BatchPlan(nv, nt, nq, leaf_sizes=leaf)
    idx, dist, (tb, tq) = plan.run(V, T, Q, timed=True)
    torch.cuda.synchronize()
    _check(meshes, queries, idx.cpu().numpy(), dist.cpu().numpy())
    print(f"eager: build {tb:.3f} ms, query {tq:.3f} ms")
    # graph capture and two replays with new queries
    for seed in (2, 3):
        _, queries2 = _batch(seed, sizes, qsizes)
        Q2 = torch.from_numpy(np.concatenate(queries2)).cuda()
        idx2, dist2 = plan.run(V, T, Q2)
        torch.cuda.synchronize()
        _check(meshes, queries2, idx2.cpu().numpy(), dist2.cpu().numpy())


def test_autotune_and_one_shot():
    import ctypes as C
    import torch
    from thea_amd import _capi as capi
    from thea_amd import default_context
    from thea_amd.batch import autotune_leaf, concat
    sizes = [3000, 12000, 50000, 700]
    qsizes = [20000, 5000, 30000, 1000]
    print("case batch autotune", flush=True)
    meshes, queries = _batch(4, sizes, qsizes)
    nv, nt, nq, verts, tris, qs = concat(meshes, queries)
    V, T, Q = (torch.from_numpy(x).cuda() for x in (verts, tris, qs))
    leaf = autotune_leaf(nv, nt, nq, V, T, Q)
    assert np.all((leaf >= 1) & (leaf <= 16))
    n = int(nq.sum())
    idx = torch.empty(n, dtype=torch.int32, device="cuda")
    dist = torch.empty(n, dtype=torch.float32, device="cuda")
    ctx = default_context()
    capi.check(capi.lib().thg_batch_closest(ctx.h, len(nt), nv.ctypes.data_as(C.c_void_p), nt.ctypes.data_as(C.c_void_p),
                                            nq.ctypes.data_as(C.c_void_p), C.c_void_p(V.data_ptr()),
                                            C.c_void_p(T.data_ptr()), C.c_void_p(Q.data_ptr()), 0,
                                            C.c_void_p(idx.data_ptr()), C.c_void_p(dist.data_ptr()), None, None))
    torch.cuda.synchronize()
    _check(meshes, queries, idx.cpu().numpy(), dist.cpu().numpy())
    print(f"autotuned leaf sizes {leaf.tolist()}")


def test_c4_full_size_oracle(oracle):
    """The 512-mesh workload the benchmark runs (42.7M triangles, 20.4M queries), autotuned leaf sizes,
    through the captured graph: the first 256 queries of every mesh bit-exact against the oracle's
    kd-tree (canonical ties), and for every query the reported index is a triangle of its mesh whose
    float distance (Thea's arithmetic) is exactly the reported distance."""
    from concurrent.futures import ThreadPoolExecutor

    import torch
    from thea_amd import workloads as W
    from thea_amd.batch import BatchPlan, autotune_plan, concat
    meshes, queries = W.many_meshes_c4()
    nv, nt, nq, verts, tris, qs = concat(meshes, queries)
    print(f"case C4 full size: {len(nt)} meshes, {int(nt.sum())} tris, {int(nq.sum())} queries", flush=True)
    V, T, Q = (torch.from_numpy(x).cuda() for x in (verts, tris, qs))
    leaf, order, launch = autotune_plan(nv, nt, nq, V, T, Q)
    assert set(np.unique(order)) <= {0, 1, 2} and launch[0] in (1, 2, 4, 8) and launch[1] in (0, 1)
    plan = BatchPlan(nv, nt, nq, leaf_sizes=leaf, child_orders=order, launch=launch)
    plan.run(V, T, Q)  # captures the graph
    idx, dist = plan.run(V, T, Q)  # replays it
    torch.cuda.synchronize()
    idx, dist = idx.cpu().numpy(), dist.cpu().numpy()
    plan.close()
    qoff = np.concatenate([[0], np.cumsum(nq)])
    head = 256

    def check(m):
        v, t = meshes[m]
        q = queries[m]
        i, d = idx[qoff[m]:qoff[m + 1]], dist[qoff[m]:qoff[m + 1]]
        assert i.min() >= 0 and i.max() < len(t), f"mesh {m}: index out of range"
        d2 = oracle.d2_to(v, t, q, i)
        assert np.array_equal(d, np.sqrt(d2.astype(np.float64)).astype(np.float32)), f"mesh {m}: distance != index's"
        ridx, _, rd2, _ = oracle.KdTree(v, t).closest(q[:head], mode=0, nthreads=1, want_cp=False)
        assert np.array_equal(i[:head], ridx), f"mesh {m}: nearest triangle differs from the oracle"
        assert np.array_equal(d2[:head].view(np.uint32), rd2.view(np.uint32))
        return len(t)

    with ThreadPoolExecutor(max_workers=8) as ex:
        done = list(ex.map(check, range(len(meshes))))
    assert len(done) == 512
    print(f"C4: {len(done)} meshes checked; leaf sizes {dict(zip(*np.unique(leaf, return_counts=True)))}, "
          f"child orders {dict(zip(*np.unique(order, return_counts=True)))}, launch {launch}")


def test_child_orders_and_launch_shapes_change_no_result():
    """Every child visiting order and query launch shape the tuner may pick gives the same bits."""
    import torch
    from thea_amd.batch import BatchPlan, concat
    sizes = [17, 1000, 5000, 20000, 60000, 300]
    qsizes = [1000, 5000, 20000, 30000, 40000, 1]
    print("case batch orders and shapes", flush=True)
    meshes, queries = _batch(3, sizes, qsizes)
    nv, nt, nq, verts, tris, qs = concat(meshes, queries)
    V, T, Q = (torch.from_numpy(x).cuda() for x in (verts, tris, qs))
    plan = BatchPlan(nv, nt, nq, leaf_sizes=np.full(len(nt), 8, np.int32))
    ref = None
    for order in (0, 1, 2):
        for bw, spill in ((1, 0), (2, 1), (4, 1), (8, 0), (8, 1)):
            plan.set_launch(np.full(len(nt), order, np.int32), bw, spill)
            plan.run(V, T, Q)
            idx, dist = plan.run(V, T, Q)
            torch.cuda.synchronize()
            got = (idx.cpu().numpy(), dist.cpu().numpy())
            if ref is None:
                ref = got
                _check(meshes, queries, *got)
            assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1].view(np.uint32), ref[1].view(np.uint32)), \
                (order, bw, spill)
    plan.close()


def test_bad_vertex_index_rejected():
    """A triangle index outside its own mesh's vertices fails the plan run and the autotuner with
    THG_E_INVALID before any vertex is read through it; a corrected input then runs."""
    import torch
    from thea_amd import _capi as capi
    from thea_amd.batch import BatchPlan, autotune_leaf, concat
    print("case batch bad index", flush=True)
    meshes, queries = _batch(6, [500, 800, 300], [100, 200, 50])
    nv, nt, nq, verts, tris, qs = concat(meshes, queries)
    for bad in (int(nv[1]), -1):  # one past mesh 1's vertices (a valid index of mesh 2's range), negative
        t_bad = tris.copy()
        t_bad[int(nt[0]) + 7, 1] = bad
        V, T, Q = (torch.from_numpy(x).cuda() for x in (verts, t_bad, qs))
        plan = BatchPlan(nv, nt, nq)
        with pytest.raises(capi.TheaGpuError) as e:
            plan.run(V, T, Q)
        assert e.value.code == capi.THG_E_INVALID and "out of range" in str(e.value)
        with pytest.raises(capi.TheaGpuError) as e:
            autotune_leaf(nv, nt, nq, V, T, Q)
        assert e.value.code == capi.THG_E_INVALID
        T_ok = torch.from_numpy(tris).cuda()
        idx, dist = plan.run(V, T_ok, Q)
        torch.cuda.synchronize()
        _check(meshes, queries, idx.cpu().numpy(), dist.cpu().numpy())
        plan.close()

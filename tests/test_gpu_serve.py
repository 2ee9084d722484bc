"""GPU tests of the request server: every result of a mixed, concurrently served request stream is
bit-identical to the same request run alone through the single-call entry points; memory-budget
admission and rejection; cancellation of queued requests."""
import numpy as np
import pytest

from conftest import random_mesh

pytestmark = pytest.mark.gpu


def _meshes():
    from thea_amd import workloads as W
    g = np.random.default_rng(50)
    v1, t1, _ = W.sphere_c1(n_queries=0)
    v2, t2 = random_mesh(g, 5000, size=0.08)
    v3, t3 = W.torus(64, 32, seed=9)
    return [(v1, t1), (v2, t2), (v3, t3)]


def _requests(g, meshes, count):
    import torch
    reqs = []
    for i in range(count):
        kind = g.choice(["closest", "knn", "ray"], p=[0.6, 0.25, 0.15])
        m = int(g.integers(len(meshes)))
        n = 0 if i == 3 else int(np.exp(g.uniform(0, np.log(40000))))
        v = meshes[m][0]
        lo, hi = v.min(0) - 0.2, v.max(0) + 0.2
        q = torch.from_numpy((lo + (hi - lo) * g.random((n, 3))).astype(np.float32)).cuda()
        d = torch.from_numpy(g.normal(size=(n, 3)).astype(np.float32)).cuda() if kind == "ray" else None
        k = int(g.choice([1, 5, 8, 16])) if kind == "knn" else 0
        bound = float(g.choice([-1.0, 0.3])) if kind != "ray" else float(g.choice([-1.0, 1.0]))
        reqs.append(dict(kind=kind, mesh=m, q=q, dirs=d, k=k, bound=bound))
    return reqs


def _alone(bvhs, r):
    b = bvhs[r["mesh"]]
    if r["kind"] == "closest":
        idx, dist, cp, _ = b.closest(r["q"], dist_bound=r["bound"])
        return idx, dist, cp
    if r["kind"] == "knn":
        idx, dist = b.kClosestElements(r["q"], r["k"], dist_bound=r["bound"])
        return idx, dist, None
    idx, t, nrm = b.rayStructureIntersection(r["q"], r["dirs"], max_time=r["bound"])
    return idx, t, nrm


def _same(a, b):
    import torch
    if a is None or b is None:
        return a is None and b is None
    return torch.equal(a.view(torch.int32) if a.dtype == torch.float32 else a,
                       b.view(torch.int32) if b.dtype == torch.float32 else b)


@pytest.mark.parametrize("coalesce", [True, False])
def test_mixed_stream_equals_alone(coalesce):
    import torch
    import thea_amd
    from thea_amd.serve import Server
    from thea_amd import _capi as capi
    print(f"case server mix coalesce={coalesce}", flush=True)
    meshes = _meshes()
    srv = Server(num_streams=4, coalesce=coalesce)
    ids = [srv.add_mesh(v, t) for v, t in meshes]
    bvhs = [thea_amd.MeshBvh(thea_amd.Mesh(v, t)) for v, t in meshes]
    g = np.random.default_rng(51)
    reqs = _requests(g, meshes, 60)
    handles = []
    for r in reqs:
        rid, out = srv.submit(r["kind"], ids[r["mesh"]], r["q"], r["dirs"], k=r["k"], bound=r["bound"])
        handles.append((rid, out))
    for (rid, out), r in zip(handles, reqs):
        rc, lat = srv.wait(rid)
        assert rc == capi.THG_OK, capi.lib().thg_last_error()
    torch.cuda.synchronize()
    for (rid, out), r in zip(handles, reqs):
        idx, dist, extra = _alone(bvhs, r)
        torch.cuda.synchronize()
        assert torch.equal(out["idx"], idx), r["kind"]
        assert _same(out["dist"], dist), r["kind"]
        if r["kind"] != "knn":
            assert _same(out["cp"], extra), r["kind"]
    st = srv.stats()
    print(f"served {st['completed']:.0f} requests in {st['launches']:.0f} launches "
          f"({st['requests_per_launch']:.2f}/launch), p50 {st['p50_us']:.0f} us, p99 {st['p99_us']:.0f} us")
    assert st["completed"] == len(reqs)
    srv.close()


def test_budget_admission_and_rejection():
    import torch
    from thea_amd.serve import Server
    from thea_amd import _capi as capi
    from thea_amd import workloads as W
    print("case server budget", flush=True)
    v, t, _ = W.sphere_c1(n_queries=0)
    budget = 16 << 19  # 8 MiB
    srv = Server(memory_budget_bytes=budget, num_streams=2)
    mid = srv.add_mesh(v, t)
    resident = srv.stats()["resident_bytes"]
    # admission of a mesh uses the bound records + nodes + leaf bounds + arrays (~7.5 MB here); once
    # built it is charged its actual size (~6.0 MB: arrays, 128-B, 48-B and 32-B records, 32-B leaf
    # bounds, binary and 4-wide nodes), so a second copy needs more than the 8 MiB budget holds
    assert 5.5e6 < resident < 6.4e6, resident
    with pytest.raises(capi.TheaGpuError) as e:
        srv.add_mesh(v, t)
    assert e.value.code == capi.THG_E_BUDGET
    big = torch.zeros((400_000, 3), dtype=torch.float32, device="cuda")  # needs ~6.4 MB of sort scratch
    with pytest.raises(capi.TheaGpuError) as e:
        srv.submit("closest", mid, big)
    assert e.value.code == capi.THG_E_BUDGET
    # requests that fit one at a time (~1.6 MB each of the ~2.0 MB left) are all served: admission
    # waits for memory to free up
    q = torch.rand((100_000, 3), device="cuda") * 3 - 1.5
    rids = [srv.submit("closest", mid, q)[0] for _ in range(6)]
    for rid in rids:
        rc, _ = srv.wait(rid)
        assert rc == capi.THG_OK
    st = srv.stats()
    assert st["rejected"] == 1
    assert 0 < st["peak_inflight_bytes"] <= budget - st["resident_bytes"]
    srv.close()


def test_cancel_queued_request():
    import torch
    from thea_amd.serve import Server
    from thea_amd import _capi as capi
    from thea_amd import workloads as W
    print("case server cancel", flush=True)
    v, t, _, _, _ = W.torus_c3(n_queries=0, n_rays=0, nu=256, nv=128)
    srv = Server(num_streams=1, max_batch_requests=1)
    mid = srv.add_mesh(v, t)
    qbig = torch.rand((4_000_000, 3), device="cuda") * 3 - 1.5
    first, _ = srv.submit("closest", mid, qbig)
    small = [srv.submit("closest", mid, qbig[:1000])[0] for _ in range(4)]
    assert srv.cancel(small[-1]) == capi.THG_OK
    rc, _ = srv.wait(small[-1])
    assert rc == capi.THG_E_CANCELLED
    for rid in [first] + small[:-1]:
        assert srv.wait(rid)[0] == capi.THG_OK
    assert srv.cancel(first) == capi.THG_E_INVALID  # finished requests cannot be cancelled
    srv.close()


def test_c5_full_size_oracle(oracle):
    """The served workload at its full size: 64 resident meshes (41.2M triangles, log-uniform 10k-4M),
    96 requests drawn like the benchmark's trace (uniform mesh, log-uniform 256..1M points, 60/25/15
    closest / 8-NN / first-hit rays), all queued at once so the workers coalesce them. Every result
    equals the same request run alone through the single-call entry points, and the first 64 queries
    of every request match the oracle's kd-tree (canonical ties) bit for bit."""
    from concurrent.futures import ThreadPoolExecutor

    import torch
    import thea_amd
    from thea_amd.serve import Server
    from thea_amd import _capi as capi
    from thea_amd import workloads as W
    meshes = W.resident_c5()
    print(f"case C5 full size: {len(meshes)} meshes, {sum(len(t) for _, t in meshes)} tris", flush=True)
    srv = Server(num_streams=4)
    ids = [srv.add_mesh(v, t) for v, t in meshes]
    g = W.rng(5, W.STREAM_TRACE)
    R = 96
    kinds = g.choice(["closest", "knn", "ray"], size=R, p=[0.6, 0.25, 0.15])
    mesh_of = g.integers(0, len(meshes), size=R)
    sizes = W.log_uniform_int(g, 256, 1 << 20, R)
    reqs = []
    for i in range(R):
        m, n = int(mesh_of[i]), int(sizes[i])
        lo, hi = W.inflated_box(meshes[m][0], 1.2)
        q = (lo + (hi - lo) * g.random((n, 3))).astype(np.float32)
        d = g.normal(size=(n, 3)).astype(np.float32) if kinds[i] == "ray" else None
        reqs.append(dict(kind=str(kinds[i]), mesh=m, q=q, dirs=d, k=8 if kinds[i] == "knn" else 0))
    dev = [dict(q=torch.from_numpy(r["q"]).cuda(), dirs=None if r["dirs"] is None else torch.from_numpy(r["dirs"]).cuda())
           for r in reqs]
    torch.cuda.synchronize()
    srv.stats(reset=True)
    handles = [srv.submit(r["kind"], ids[r["mesh"]], d["q"], d["dirs"], k=r["k"]) for r, d in zip(reqs, dev)]
    for rid, _ in handles:
        rc, _ = srv.wait(rid)
        assert rc == capi.THG_OK, capi.lib().thg_last_error()
    st = srv.stats()
    print(f"C5: {st['completed']:.0f} requests, {sum(sizes)} points in {st['launches']:.0f} worker batches "
          f"({st['requests_per_launch']:.2f} requests per batch), p99 {st['p99_us'] / 1e3:.1f} ms", flush=True)
    assert st["completed"] == R
    # alone, through the single-call entry points on separately built BVHs (the server's leaf size 4)
    bvhs = {}
    for r, d, (rid, out) in zip(reqs, dev, handles):
        m = r["mesh"]
        if m not in bvhs:
            bvhs[m] = thea_amd.MeshBvh(thea_amd.Mesh(*meshes[m]))
        alone = _alone(bvhs, dict(kind=r["kind"], mesh=m, q=d["q"], dirs=d["dirs"], k=r["k"], bound=-1.0))
        torch.cuda.synchronize()
        assert torch.equal(out["idx"], alone[0]), r["kind"]
        assert _same(out["dist"], alone[1]), r["kind"]
        if r["kind"] != "knn":
            assert _same(out["cp"], alone[2]), r["kind"]
    srv.close()
    del bvhs
    # the oracle on the first 64 queries of every request
    head = 64
    used = sorted(set(int(m) for m in mesh_of))
    with ThreadPoolExecutor(max_workers=8) as ex:
        kds = dict(zip(used, ex.map(lambda m: oracle.KdTree(*meshes[m]), used)))
    for r, (rid, out) in zip(reqs, handles):
        kd, q = kds[r["mesh"]], r["q"][:head]
        idx = out["idx"][:head].cpu().numpy()
        val = out["dist"][:head].cpu().numpy()
        if r["kind"] == "closest":
            ridx, rcp, rd2, _ = kd.closest(q, mode=0, nthreads=1)
            assert np.array_equal(idx, ridx)
            assert np.array_equal(val, np.sqrt(rd2.astype(np.float64)).astype(np.float32))
            assert np.array_equal(out["cp"][:head].cpu().numpy().view(np.uint32), rcp.view(np.uint32))
        elif r["kind"] == "knn":
            ridx, rd2, _ = kd.knn(q, 8, mode=0, nthreads=1)
            assert np.array_equal(idx, ridx)
            assert np.array_equal(val, np.sqrt(rd2.astype(np.float64)).astype(np.float32))
        else:
            ridx, rt, _ = kd.ray(q, r["dirs"][:head], mode=0, nthreads=1)
            assert np.array_equal(idx, ridx)
            assert np.array_equal(val.view(np.uint32), rt.view(np.uint32))
    print(f"C5: {R} requests equal alone and in the mix; first {head} queries of each match the oracle")


def test_add_mesh_while_requests_run():
    """Meshes added from another thread while requests on earlier meshes are queued and running: the
    workers use the BVH each request captured at submit, so growing the mesh table does not disturb
    them, and every result equals the request run alone."""
    import threading

    import torch
    import thea_amd
    from thea_amd.serve import Server
    from thea_amd import _capi as capi
    from thea_amd import workloads as W
    print("case server add_mesh during requests", flush=True)
    base = _meshes()
    srv = Server(num_streams=4)
    ids = [srv.add_mesh(v, t) for v, t in base]
    bvhs = [thea_amd.MeshBvh(thea_amd.Mesh(v, t)) for v, t in base]
    g = np.random.default_rng(52)
    reqs = _requests(g, base, 120)
    extra = [W.torus(48 + 8 * i, 24 + 4 * i, seed=20 + i) for i in range(12)]
    added = []

    def adder():
        for v, t in extra:
            added.append(srv.add_mesh(v, t))
    th = threading.Thread(target=adder)
    th.start()
    handles = [srv.submit(r["kind"], ids[r["mesh"]], r["q"], r["dirs"], k=r["k"], bound=r["bound"]) for r in reqs]
    th.join()
    for rid, _ in handles:
        rc, _ = srv.wait(rid)
        assert rc == capi.THG_OK, capi.lib().thg_last_error()
    assert added == list(range(len(base), len(base) + len(extra)))
    for (rid, out), r in zip(handles, reqs):
        idx, dist, extra_out = _alone(bvhs, r)
        torch.cuda.synchronize()
        assert torch.equal(out["idx"], idx), r["kind"]
        assert _same(out["dist"], dist), r["kind"]
    # the new meshes serve requests too
    v, t = extra[-1]
    q = torch.from_numpy((v.min(0) + (v.max(0) - v.min(0)) * g.random((5000, 3))).astype(np.float32)).cuda()
    rid, out = srv.submit("closest", added[-1], q)
    assert srv.wait(rid)[0] == capi.THG_OK
    ref = thea_amd.MeshBvh(thea_amd.Mesh(v, t)).closestElement(q)
    torch.cuda.synchronize()
    assert torch.equal(out["idx"], ref)
    srv.close()

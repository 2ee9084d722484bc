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
    budget = 6 << 20
    srv = Server(memory_budget_bytes=budget, num_streams=2)
    mid = srv.add_mesh(v, t)
    resident = srv.stats()["resident_bytes"]
    # admission of a mesh uses the bound records + nodes + arrays (~4.3 MB here); once built it is
    # charged its actual size (~3.5 MB), so a second copy needs more than the 6 MB budget holds
    assert 3.0e6 < resident < 4.3e6, resident
    with pytest.raises(capi.TheaGpuError) as e:
        srv.add_mesh(v, t)
    assert e.value.code == capi.THG_E_BUDGET
    big = torch.zeros((400_000, 3), dtype=torch.float32, device="cuda")  # needs ~6.4 MB of sort scratch
    with pytest.raises(capi.TheaGpuError) as e:
        srv.submit("closest", mid, big)
    assert e.value.code == capi.THG_E_BUDGET
    # requests that fit one at a time (~1.6 MB each of the ~2.5 MB left) are all served: admission
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

"""Benchmark of the BVH query path on one MI355X.

Default workload (the headline metric): closest point on a displaced torus of 1,048,576 triangles
for 10,000,000 query points uniform in its 20%-inflated bounding box (float32). One step = one
closest-point pass over the whole 10M-query batch (spatial ordering of the queries + traversal,
writing triangle index, closest point and distance per query) with the mesh and its BVH resident
in HBM and the queries already in device memory. BVH build time is reported separately.

Prints one JSON line. `roofline` uses the algorithmic bytes of Thea's own traversal on the same
inputs (tests/golden/workload_counts.json) over the traversal kernel's average duration, measured
with HIP events on the kernel's stream. `cpu_baseline` times the oracle (a C++ restatement of
Thea's kd-tree and traversal) on a bounded sample of the same queries on this host.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="c3", choices=["c3", "c1", "c3ray", "c2knn", "c2sample"])
    p.add_argument("--queries", type=int, default=None, help="override the query count")
    p.add_argument("--leaf", type=int, default=4)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-threads", type=int, default=None)
    return p.parse_args()


def load_counts():
    path = os.path.join(ROOT, "tests", "golden", "workload_counts.json")
    with open(path) as f:
        return json.load(f)


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", f"{name}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def cpu_threads(args):
    if args.cpu_threads:
        return args.cpu_threads
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def main():
    args = parse()
    if args.gpus != 1:
        sys.stderr.write("bench.py: this build runs on one GPU only (the query path shards by replicating the "
                         "mesh and splitting queries; no multi-GPU run is implemented). Use --gpus 1.\n")
        sys.exit(2)

    import numpy as np
    import torch
    import thea_amd
    from thea_amd import _capi as capi
    from thea_amd import workloads as W

    dev = torch.device("cuda", 0)
    ctx = thea_amd.default_context(0)
    counts = load_counts()
    wl = args.workload
    t_gen = time.time()
    info = {}
    if wl in ("c3", "c3ray"):
        nq = args.queries or (10_000_000 if wl == "c3" else 4_000_000)
        v, t, q, ro, rd = W.torus_c3(n_queries=nq if wl == "c3" else 0, n_rays=nq if wl == "c3ray" else 0)
    elif wl == "c1":
        nq = args.queries or 100_000
        v, t, q = W.sphere_c1(n_queries=nq)
    else:
        nq = args.queries or 1_000_000
        v, t = W.ellipsoids_c2()
    gen_s = time.time() - t_gen

    mesh = thea_amd.Mesh(v, t)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bvh = thea_amd.MeshBvh(mesh, leaf_size=args.leaf)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3

    L = capi.lib()
    stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    mem = capi.THG_MEM_DEVICE
    P = lambda x: C.c_void_p(x.data_ptr())  # noqa: E731

    if wl in ("c3", "c1"):
        qd = torch.from_numpy(q).to(dev)
        idx = torch.empty(nq, dtype=torch.int32, device=dev)
        dist = torch.empty(nq, dtype=torch.float32, device=dev)
        cp = torch.empty((nq, 3), dtype=torch.float32, device=dev)

        def step():
            capi.check(L.thg_closest_point(bvh.h, P(qd), nq, -1.0, P(idx), P(cp), P(dist), None, mem, stream))
        kernel = "closest"
        metric_counts = counts["c3_closest" if wl == "c3" else "c1_closest"]
        bytes_per_query = 32 + 24 * metric_counts["V_box"] + 36 * metric_counts["V_tri"]
        metric = "closest-pt queries/s (1M tri, 10M q)" if wl == "c3" else "closest-pt queries/s (C1 sphere)"
        unit = "queries/s"
        workload = ("closest point: displaced torus 1,048,576 tris x 10M queries in the 20%-inflated bbox"
                    if wl == "c3" else "closest point: perturbed geodesic sphere 20,480 tris x 100k queries")
    elif wl == "c3ray":
        od, dd = torch.from_numpy(ro).to(dev), torch.from_numpy(rd).to(dev)
        idx = torch.empty(nq, dtype=torch.int32, device=dev)
        tt = torch.empty(nq, dtype=torch.float32, device=dev)

        def step():
            capi.check(L.thg_ray_first_hit(bvh.h, P(od), P(dd), nq, -1.0, P(idx), P(tt), None, mem, stream))
        kernel = "ray"
        metric_counts = counts["c3_ray"]
        bytes_per_query = 32 + 24 * metric_counts["V_box"] + 36 * metric_counts["V_tri"]
        metric, unit = "ray first-hit rays/s (1M tri, 4M rays)", "rays/s"
        workload = "ray first hit: displaced torus 1,048,576 tris x 4M rays, origins in the inflated bbox"
    elif wl == "c2knn":
        sampler = thea_amd.MeshSampler(mesh)
        pos, _, _ = sampler.sampleEvenlyByArea(nq, seed=2, stream_id=W.STREAM_SAMPLES, device=True)
        pb = thea_amd.PointBvh(pos, leaf_size=args.leaf)
        k = 16
        idx = torch.empty((nq, k), dtype=torch.int32, device=dev)
        dist = torch.empty((nq, k), dtype=torch.float32, device=dev)

        def step():
            capi.check(L.thg_k_closest(pb.h, P(pos), nq, k, -1.0, 1, P(idx), P(dist), None, mem, stream))
        kernel = "knn"
        metric_counts = counts["c2_knn16"]
        bytes_per_query = 12 + 8 * k + 24 * metric_counts["V_box"] + 12 * metric_counts["V_tri"]
        metric, unit = "k-NN (k=16) queries/s (1M area samples of 64 ellipsoids)", "queries/s"
        workload = "16-NN graph without self over 1M area-weighted samples of 64 ellipsoids (128k tris)"
    else:  # c2sample
        sampler = thea_amd.MeshSampler(mesh)
        pos = torch.empty((nq, 3), dtype=torch.float32, device=dev)
        nrm = torch.empty((nq, 3), dtype=torch.float32, device=dev)
        tri = torch.empty(nq, dtype=torch.int32, device=dev)
        got = C.c_int64()
        mesh.total_area()  # builds the CDF once, outside the timed region

        def step():
            capi.check(L.thg_sample_area(mesh.h, nq, C.c_uint64(2), C.c_uint64(3), P(pos), P(nrm), P(tri),
                                         C.byref(got), mem, stream))
        kernel = "sample"
        bytes_per_query = 52 + 8 * int(np.ceil(np.log2(len(t))))
        metric_counts = {}
        metric, unit = "area samples/s (64 ellipsoids, 128k tris)", "samples/s"
        workload = "area-weighted sampling, 1M samples per step on 64 ellipsoids (128k tris)"

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.kernel_time(kernel, reset=True)
    ctx.kernel_time("order", reset=True)
    ctx.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ctx.set_timing(False)
    k_ms, k_n = ctx.kernel_time(kernel, reset=True)
    o_ms, o_n = ctx.kernel_time("order", reset=True)
    ms_per_step = elapsed * 1e3 / args.steps
    value = nq * args.steps / elapsed
    k_avg = k_ms / max(k_n, 1)
    achieved = bytes_per_query * nq / (k_avg * 1e-3) / 1e9
    traffic = load_traffic(f"{kernel}_{wl}")

    result = {
        "metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload, "triangles": int(len(t)), "queries_per_step": int(nq),
                   "leaf_size": bvh.info["leaf_size"], "bvh_nodes": int(bvh.info["nodes"]),
                   "bvh_max_depth": int(bvh.info["max_depth"])},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "kernel": f"{kernel} traversal", "kernel_avg_ms": k_avg, "kernel_launches": k_n,
                     "bytes_per_query": bytes_per_query, "thea_traversal_counts": metric_counts},
        "build_ms": build_ms,
        "order_kernels_avg_ms": o_ms / max(o_n, 1) if o_n else 0.0,
        "input_generation_s": gen_s,
    }

    if wl in ("c3", "c1"):
        st = (C.c_int64 * 8)()
        capi.check(L.thg_closest_stats(bvh.h, P(qd), nq, stream, st))
        waves = max(st[0], 1)
        result["gpu_traversal"] = {
            "per_wave": {"nodes": st[2] / waves, "box_tests": st[3] / waves, "elements_loaded": st[4] / waves,
                         "elements_evaluated": st[5] / waves, "plane_admitted": st[6] / waves,
                         "near_best": st[7] / waves},
            "queries_per_wave": st[1] / waves,
            "note": "wave-cooperative traversal: every lane evaluates each element its wave evaluates"}
    if not args.no_cpu_baseline and wl in ("c3", "c1", "c3ray", "c2knn"):
        result["cpu_baseline"] = cpu_baseline(args, wl, v, t, q if wl in ("c3", "c1") else None,
                                              (ro, rd) if wl == "c3ray" else None, unit)
    print(json.dumps(result), flush=True)


def cpu_baseline(args, wl, v, t, q, rays, unit):
    """The oracle (C++ restatement of Thea's kd-tree build and traversal, Thea traversal order) timed on
    a bounded sample of the same inputs, on all usable host cores and on one core."""
    import numpy as np
    from oracle import oracle as O
    nth = cpu_threads(args)
    t0 = time.perf_counter()
    if wl == "c2knn":
        from thea_amd import workloads as W
        pos, _, _ = O.sample_area(v, t, 1_000_000, 2, W.STREAM_SAMPLES)
        kd = O.KdTree(points=pos)
    else:
        kd = O.KdTree(v, t)
    build_s = time.perf_counter() - t0
    if wl in ("c3", "c1"):
        n_multi = min(len(q), 60_000 * nth) if wl == "c3" else len(q)
        n_single = min(len(q), 60_000) if wl == "c3" else len(q)
        run = lambda n, th: kd.closest(q[:n], mode=1, nthreads=th, want_cp=True)  # noqa: E731
    elif wl == "c3ray":
        o, d = rays
        n_multi, n_single = min(len(o), 2_000_000), min(len(o), 300_000)
        run = lambda n, th: kd.ray(o[:n], d[:n], mode=1, nthreads=th)  # noqa: E731
    else:
        n_multi, n_single = 400_000, 50_000
        run = lambda n, th: kd.knn(pos[:n], 16, mode=1, exclude_self=True, nthreads=th)  # noqa: E731
    t0 = time.perf_counter()
    run(n_multi, nth)
    multi_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    run(n_single, 1)
    single_s = time.perf_counter() - t0
    return {"value": n_multi / multi_s, "unit": unit, "cores": nth, "kind": "port",
            "sample": f"first {n_multi} queries on {nth} threads ({multi_s:.1f} s) and first {n_single} on 1 thread "
                      f"of the same batch; kd-tree build {build_s:.2f} s excluded",
            "single_thread_value": n_single / single_s, "build_s": build_s}


if __name__ == "__main__":
    main()

"""Benchmark of the BVH query path on one MI355X.

Default workload (the headline metric): closest point on a displaced torus of 1,048,576 triangles
for 10,000,000 query points uniform in its 20%-inflated bounding box (float32). One step = one
closest-point pass over the whole 10M-query batch (spatial ordering of the queries + traversal,
writing triangle index, closest point and distance per query) with the mesh and its BVH resident
in HBM and the queries already in device memory. BVH build time is reported separately.

Prints one JSON line. `roofline` prices the traversal kernel against the ceiling that binds it
(vector ALU issue or HBM bandwidth, from rocprofv3 counters of the same kernel under profiles/)
over its average duration, measured with HIP events on the kernel's stream; the bytes Thea's own
traversal reads on the same inputs (tests/golden/workload_counts.json) ride along as a labelled
diagnostic. `cpu_baseline` times the oracle (a C++ restatement of Thea's kd-tree and traversal) on
a bounded sample of the same queries on this host. `served` is the second half of the headline
metric: the request stream over 64 resident meshes (served requests/s, p50/p99 latency and
requests per launch, against plain one-call-per-request dispatch), as `--workload c5` runs it.
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
# BVH leaf size per workload, the fastest of 1, 2, 4, 6, 8, 12, 16 measured on one MI355X for that query
# type (closest point prefers 12 since whole leaves are rejected at once, k-NN over points 16,
# incoherent rays 1); the library default is 4.
LEAF_DEFAULT = {"c3": 12, "c1": 16, "c3ray": 2, "c2knn": 16, "c2sample": 4, "icp": 8, "graph": 12, "sdf": 1,
                "vis": 1}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="c3", choices=["c3", "c1", "c3ray", "c2knn", "c2sample", "c4", "c5", "copy", "icp", "graph", "sdf", "sep", "vis"])
    p.add_argument("--requests", type=int, default=400, help="c5: requests in the trace")
    p.add_argument("--load", type=float, default=0.7, help="c5: Poisson rate as a fraction of plain capacity")
    p.add_argument("--queries", type=int, default=None, help="override the query count")
    p.add_argument("--leaf", type=int, default=None,
                   help="BVH leaf size (default: the measured best per workload, LEAF_DEFAULT)")
    p.add_argument("--c5-leaf", type=int, default=4, help="c5 (and the served half of the default line): leaf size "
                   "of the resident meshes, for the server and plain dispatch alike")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-served", action="store_true", help="c3: leave out the served-stream half of the line")
    p.add_argument("--kabsch", action="store_true", help="icp: least-squares fit instead of Thea's")
    p.add_argument("--f64", action="store_true",
                   help="icp: double-precision target (ShapeAlign's BvhN<Vector3d>): double positions and search")
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


def load_counters(name):
    path = os.path.join(ROOT, "profiles", f"{name}_counters.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


# Vector ALU ceiling: each of a CU's 4 SIMDs executes 16 lanes of a 32-bit VALU operation per cycle
# (a wave64 instruction every 4 cycles) at MI355X's 2.4 GHz peak engine clock: 256 CUs x 4 x 16 x
# 2.4 GHz = 39.3 T lane-operations/s (the published 157.3 TFLOP/s FP32 counts an FMA as 2 and packed
# FP32 as 2 more; these kernels do neither).
ENGINE_GHZ = 2.4
SIMDS_PER_CU = 4
LANES_PER_SIMD = 16
CYCLES_PER_VALU = 4


def measured_roofline(kernel, wl, kernel_avg_ms, alg_bytes_per_launch, num_cus, model_note):
    """Roofline object of a traversal kernel against the resource that binds it.

    Two ceilings are priced with rocprofv3 counters of the same kernel on the same workload
    (profiles/<kernel>_<wl>_counters.json and _traffic.json; the counters are fixed per launch for a
    fixed input) over the kernel's average duration measured here with HIP events on its stream:
      - vector ALU: lane operations per launch = SQ_INSTS_VALU wave instructions x 64 x the active-lane
        fraction (VALUUtilization), against CUs x 4 SIMDs x 16 lanes x 2.4 GHz. Beside it, the issue
        slots the kernel fills: SQ_INSTS_VALU x 4 cycles / (CUs x 4 SIMDs x kernel cycles) (this can
        pass 1 where transcendental instructions issue alongside other vector work);
      - HBM: measured bytes per launch (FETCH_SIZE and WRITE_SIZE scaled by a plain copy of known
        size) against 8 TB/s;
      - vector memory data path: the fraction of GPU-active cycles the texture data units (TD, one
        per CU) are busy returning load data (TD_TD_BUSY over GRBM_GUI_ACTIVE), the ceiling of
        gather-heavy traversals whose loads mostly hit in cache.
    `bound` names the largest fraction. Thea's own traversal bytes (what its kd-tree walk tests on the
    same inputs) are kept as a labelled diagnostic: most of them are L2/MALL re-reads, not HBM."""
    sec = kernel_avg_ms * 1e-3
    traffic = load_traffic(f"{kernel}_{wl}")
    ctr = load_counters(f"{kernel}_{wl}") or {}
    valu = ctr.get("SQ_INSTS_VALU_per_dispatch")
    out = {"kernel": ctr.get("kernel", kernel), "kernel_avg_ms": kernel_avg_ms}
    hbm = None
    if traffic:
        gbps = traffic / sec / 1e9
        hbm = {"achieved": gbps, "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": gbps / HBM_PEAK_GBPS,
               "bytes_per_launch": traffic, "source": f"profiles/{kernel}_{wl}_traffic.json"}
    issue = None
    util = ctr.get("VALUUtilization")
    if valu and util:
        rate = valu * 64 * (util / 100.0) / sec / 1e12
        peak = num_cus * SIMDS_PER_CU * LANES_PER_SIMD * ENGINE_GHZ / 1e3
        issue = {"achieved": rate, "peak": peak, "unit": "T VALU lane-ops/s", "frac": rate / peak,
                 "issue_slot_fraction": valu * CYCLES_PER_VALU / (num_cus * SIMDS_PER_CU * sec * ENGINE_GHZ * 1e9),
                 "active_lane_fraction": util / 100.0, "valu_instructions_per_launch": valu,
                 "counter_run_kernel_ms": ctr.get("counter_run_kernel_ms"), "l2_hit_rate": ctr.get("l2_hit_rate"),
                 "source": f"profiles/{kernel}_{wl}_counters.json"}
    vmem = None
    if ctr.get("TD_busy_fraction") is not None:
        vmem = {"achieved": ctr["TD_busy_fraction"], "peak": 1.0, "unit": "TD busy fraction",
                "frac": ctr["TD_busy_fraction"], "ta_busy_fraction": ctr.get("TA_busy_fraction"),
                "source": f"profiles/{kernel}_{wl}_counters.json"}
    cands = [(c["frac"], name, c) for name, c in (("valu", issue), ("hbm", hbm), ("vmem", vmem)) if c]
    if cands:
        _, bound, main = max(cands, key=lambda x: x[0])
    else:
        bound, main = None, None
    if main:
        out.update({"bound": bound, "achieved": main["achieved"], "peak": main["peak"], "unit": main["unit"],
                    "frac": main["frac"]})
    else:
        out.update({"bound": None, "achieved": None, "peak": None, "unit": None, "frac": None})
    out["traffic"] = traffic
    out["valu_issue"] = issue
    out["hbm"] = hbm
    out["vmem"] = vmem
    # the other two figures rocprof reports for the kernel: L2 hit rate and wave occupancy
    out["l2_hit_rate"] = ctr.get("l2_hit_rate")
    out["occupancy"] = None if ctr.get("occupancy_percent") is None else {
        "percent_of_max": ctr["occupancy_percent"], "mean_waves_per_simd": ctr.get("mean_waves_per_simd"),
        "note": ctr.get("occupancy_note"), "source": f"profiles/{kernel}_{wl}_counters.json"}
    alg_gbps = alg_bytes_per_launch / sec / 1e9
    out["thea_bytes_model"] = {"bytes_per_launch": alg_bytes_per_launch, "GB/s": alg_gbps,
                               "ratio_to_hbm_peak": alg_gbps / HBM_PEAK_GBPS, "note": model_note}
    return out


def log(msg):
    """Progress line on stderr (the JSON result is the only line on stdout)."""
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


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
    if args.leaf is None:
        args.leaf = LEAF_DEFAULT.get(wl, 4)
    if wl == "c5":
        run_c5(args)
        return
    if wl == "c4":
        run_c4(args, counts)
        return
    if wl == "copy":
        run_copy(args)
        return
    if wl == "icp":
        run_icp(args)
        return
    if wl == "graph":
        run_graph(args)
        return
    if wl == "sdf":
        run_sdf(args)
        return
    if wl == "sep":
        run_sep(args)
        return
    if wl == "vis":
        run_vis(args)
        return
    # The served half of the default line runs first, while the device holds nothing this process
    # allocated and freed: resident meshes placed in memory that an earlier allocation of the same
    # process released serve 10-20% fewer requests/s (measured with a 1 GB tensor or one C3 BVH
    # allocated and freed beforehand), while plain dispatch is unaffected.
    c5 = None
    if wl == "c3" and not args.no_served:
        log("served stream (64 resident meshes)")
        c5 = c5_measure(args)
        torch.cuda.empty_cache()
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

    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"warmup step {i + 1}/{args.warmup} done")
    ctx.kernel_time(kernel, reset=True)
    ctx.kernel_time("order", reset=True)
    ctx.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    log(f"{args.steps} timed steps done in {elapsed:.3f} s")
    ctx.set_timing(False)
    k_ms, k_n = ctx.kernel_time(kernel, reset=True)
    o_ms, o_n = ctx.kernel_time("order", reset=True)
    ms_per_step = elapsed * 1e3 / args.steps
    value = nq * args.steps / elapsed
    k_avg = k_ms / max(k_n, 1)
    roofline = measured_roofline(
        kernel, wl, k_avg, bytes_per_query * nq, torch.cuda.get_device_properties(dev).multi_processor_count,
        "bytes Thea's own traversal reads per query (the boxes and elements its kd-tree tests) on the same "
        f"inputs: {bytes_per_query:.1f} B/query; most are re-read from L2/MALL, so this is no HBM figure")
    roofline["bytes_per_query"] = bytes_per_query
    roofline["thea_traversal_counts"] = metric_counts
    roofline["kernel_launches"] = k_n

    result = {
        "metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload, "triangles": int(len(t)), "queries_per_step": int(nq),
                   "leaf_size": bvh.info["leaf_size"], "bvh_nodes": int(bvh.info["nodes"]),
                   "bvh_max_depth": int(bvh.info["max_depth"])},
        "roofline": roofline,
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
                         "exact_eval_lane_slots": st[5] / waves, "pairs_passing_rejects": st[6] / waves,
                         "near_best": st[7] / waves},
            "queries_per_wave": st[1] / waves,
            "lane_utilisation_of_exact_evals": st[6] / max(st[5], 1),
            "note": "one wave walks one tree for its 64 queries; exact evaluation lane slots vs the "
                    "(lane, element) pairs that needed one"}
    if not args.no_cpu_baseline and wl in ("c3", "c1", "c3ray", "c2knn", "c2sample"):
        result["cpu_baseline"] = cpu_baseline(args, wl, v, t, q if wl in ("c3", "c1") else None,
                                              (ro, rd) if wl == "c3ray" else None, unit)
    if c5 is not None:
        result["metric"] = "closest-pt queries/s (1M tri, 10M q); served p50/p99 + qps vs plain dispatch"
        result["served"] = {k: c5[k] for k in ("metric", "value", "unit", "config", "saturation", "poisson",
                                               "requests_per_launch")}
    print(json.dumps(result), flush=True)


def run_graph(args):
    """SampleGraph's proximity graph (PointSet3::updateGraph) over the C2 workload's 1,000,000 area
    samples with their face normals, max degree 16: one step = the separation estimate (100 picks),
    the neighbour sets of every sample and the final average separation. The point BVH is built once,
    outside the timed region."""
    import numpy as np
    import torch
    import thea_amd
    from thea_amd import workloads as W
    v, t = W.ellipsoids_c2()
    n = args.queries or 1_000_000
    pos, nrm, _ = thea_amd.MeshSampler(thea_amd.Mesh(v, t)).sampleEvenlyByArea(n, seed=2, stream_id=W.STREAM_SAMPLES,
                                                                               device=True)
    b = thea_amd.PointBvh(pos, leaf_size=args.leaf)
    for i in range(args.warmup):
        b.proximityGraph(normals=nrm, max_degree=16, seed=4, average=False)
        log(f"warmup graph {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        idx, sep, deg, _ = b.proximityGraph(normals=nrm, max_degree=16, seed=4, average=False)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    log(f"{args.steps} graphs: {sec * 1e3:.1f} ms each")
    t0 = time.perf_counter()
    _, _, _, avg = b.proximityGraph(normals=nrm, max_degree=16, seed=4)
    avg_sec = time.perf_counter() - t0
    deg = deg.cpu().numpy()
    result = {
        "metric": "SampleGraph proximity graph samples/s (1M area samples of 64 ellipsoids, max degree 16)",
        "value": n / sec, "unit": "samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sec * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "proximity graph of C2 area samples with face normals", "samples": n,
                   "max_degree": 16, "leaf_size": args.leaf},
        "mean_degree": float(deg.mean()), "separation": [float(avg[0]), float(avg[1])],
        "with_final_average_ms": avg_sec * 1e3,
        "note": "value excludes the final average separation (a sequential float sum of all separations "
                "on the host, as Thea computes it); with_final_average_ms is one call including it",
    }
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        t0 = time.perf_counter()
        O.proximity_graph(pos.cpu().numpy(), nrm.cpu().numpy(), 16, O.graph_picks(n, 100, 4))
        csec = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": n / csec, "unit": "samples/s", "cores": 1, "kind": "port",
                                  "sample": f"the whole graph on one thread ({csec:.1f} s; grid ball queries)"}
    print(json.dumps(result), flush=True)


def run_sdf(args):
    """Shape diameter function (ShapeDiameter::compute) at 1,000,000 area samples of the C3 torus with
    their face normals: 30 cone rays per point (30M first-hit rays) and Thea's per-point reduction.
    One step = one batch; the mesh BVH is built once."""
    import numpy as np
    import torch
    import thea_amd
    from thea_amd import workloads as W
    v, t, _, _, _ = W.torus_c3(n_queries=0, n_rays=0)
    mesh = thea_amd.Mesh(v, t)
    n = args.queries or 1_000_000
    pos, nrm, _ = thea_amd.MeshSampler(mesh).sampleEvenlyByArea(n, seed=9, stream_id=W.STREAM_SAMPLES, device=True)
    b = thea_amd.MeshBvh(mesh, leaf_size=args.leaf)
    scale = float(np.linalg.norm(v.max(0) - v.min(0)))
    for i in range(args.warmup):
        b.shapeDiameter(pos, nrm, scale=scale)
        log(f"warmup sdf {i + 1}/{args.warmup} done")
    b.ctx.kernel_time("ray", reset=True)
    b.ctx.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sdf = b.shapeDiameter(pos, nrm, scale=scale)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    b.ctx.set_timing(False)
    k_ms, k_n = b.ctx.kernel_time("ray", reset=True)
    log(f"{args.steps} batches: {sec * 1e3:.1f} ms each")
    result = {
        "metric": "shape diameter points/s (C3 torus, 1M surface points, 30 rays each)", "value": n / sec,
        "unit": "points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 rays, f64 reduction",
        "data": "synthetic",
        "config": {"workload": "ShapeDiameter at C3 area samples with face normals", "points": n, "rays": 30 * n,
                   "leaf_size": args.leaf, "scale": scale},
        "roofline": ray_roofline("c3_sdf_ray", 30 * n, k_ms / max(args.steps, 1), k_n / max(args.steps, 1), "ray",
                                 "sdf"),
        "ray_kernel_avg_ms": k_ms / max(k_n, 1), "rays_per_s_kernel": 30 * n / (k_ms / max(k_n, 1) * 1e-3),
        "median_sdf": float(sdf[sdf >= 0].median()),
    }
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        m = min(n, 100_000)
        kd = O.KdTree(v, t)
        tn = O.tri_props(v, t)[:, :3]
        t0 = time.perf_counter()
        O.shape_diameter(kd, tn, pos[:m].cpu().numpy(), nrm[:m].cpu().numpy(), scale)
        csec = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": m / csec, "unit": "points/s", "cores": cpu_threads(args), "kind": "port",
                                  "sample": f"the first {m} points ({csec:.1f} s): rays on all threads, generation "
                                            f"and reduction on one"}
    print(json.dumps(result), flush=True)


def ray_roofline(counts_key, rays, kernel_ms, launches, kernel, wl):
    """Roofline object of a ray feature's traversal kernel (kernel_ms: its total time for the batch,
    over `launches` launches, HIP events on its stream), against the measured ceilings of
    measured_roofline (profiles/<kernel>_<wl>_*.json). Thea's per-ray box and triangle tests on the same
    rays (tests/golden/workload_counts.json) priced at 32 + 24 V_box + 36 V_tri bytes per ray are
    kept as a labelled diagnostic."""
    import torch
    c = load_counts().get(counts_key)
    if not c or kernel_ms <= 0 or launches <= 0:
        return None
    bpr = 32 + 24 * c["V_box"] + 36 * c["V_tri"]
    out = measured_roofline(kernel, wl, kernel_ms / launches, bpr * rays / launches,
                            torch.cuda.get_device_properties(0).multi_processor_count,
                            f"bytes Thea's own traversal reads per ray (the boxes and triangles its kd-tree tests, "
                            f"sample of the batch): {bpr:.1f} B/ray; most are re-read from L2/MALL")
    out["bytes_per_ray"] = bpr
    out["thea_traversal_counts"] = c
    out["kernel_launches_per_batch"] = launches
    return out


def run_vis(args):
    """External visibility (Visibility::compute) at 1,000,000 area samples of the C3 torus, 100 rays
    each (100M any-hit rays, processed in 2^24-ray chunks). One step = one batch; the BVH is built
    once."""
    import numpy as np
    import torch
    import thea_amd
    from thea_amd import workloads as W
    v, t, _, _, _ = W.torus_c3(n_queries=0, n_rays=0)
    mesh = thea_amd.Mesh(v, t)
    n = args.queries or 1_000_000
    pos, _, _ = thea_amd.MeshSampler(mesh).sampleEvenlyByArea(n, seed=10, stream_id=W.STREAM_SAMPLES, device=True)
    b = thea_amd.MeshBvh(mesh, leaf_size=args.leaf)
    for i in range(args.warmup):
        b.visibility(pos, num_rays=100, seed=1)
        log(f"warmup vis {i + 1}/{args.warmup} done")
    b.ctx.kernel_time("any_hit", reset=True)
    b.ctx.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        vis = b.visibility(pos, num_rays=100, seed=1)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    b.ctx.set_timing(False)
    k_ms, k_n = b.ctx.kernel_time("any_hit", reset=True)
    log(f"{args.steps} batches: {sec * 1e3:.1f} ms each")
    result = {
        "metric": "visibility points/s (C3 torus, 1M surface points, 100 rays each)", "value": n / sec,
        "unit": "points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 rays, f64 fractions",
        "data": "synthetic",
        "config": {"workload": "Visibility at C3 area samples", "points": n, "rays": 100 * n, "leaf_size": args.leaf},
        "roofline": ray_roofline("c3_vis_ray", 100 * n, k_ms / max(args.steps, 1), k_n / max(args.steps, 1),
                                 "any_hit", "vis"),
        "any_hit_ms_per_batch": k_ms / max(args.steps, 1), "any_hit_launches_per_batch": k_n / max(args.steps, 1),
        "rays_per_s_kernel": 100 * n * args.steps / (k_ms * 1e-3) if k_ms > 0 else None,
        "mean_visibility": float(vis.mean()),
    }
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        m = min(n, 200_000)
        kd = O.KdTree(v, t)
        lo, hi = b.getBounds()
        e = hi - lo
        scale = float(np.sqrt(e[0] * e[0] + (e[1] * e[1] + e[2] * e[2])))
        hpos = pos[:m].cpu().numpy()
        t0 = time.perf_counter()
        O.visibility(kd, hpos, 100, scale, 1)
        csec = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": m / csec, "unit": "points/s", "cores": cpu_threads(args), "kind": "port",
                                  "sample": f"the first {m} points ({csec:.1f} s): rays traced on all threads, "
                                            f"generated on one"}
    print(json.dumps(result), flush=True)


def run_sep(args):
    """MeshSample's separation mode (MeshSampler::sampleEvenlyBySeparation) on the C2 mesh (64
    ellipsoids, 128k triangles): 10,000 samples picked by geodesic furthest-point sampling among
    30,000 area samples. One step = one call: area oversampling, point BVH, proximity graph, the
    picks and the gather."""
    import numpy as np
    import torch
    import thea_amd
    from thea_amd import workloads as W
    v, t = W.ellipsoids_c2()
    mesh = thea_amd.Mesh(v, t)
    sampler = thea_amd.MeshSampler(mesh)
    n = args.queries or 10_000
    for i in range(args.warmup):
        sampler.sampleEvenlyBySeparation(n, seed=2, stream_id=W.STREAM_SAMPLES, device=True)
        log(f"warmup sep {i + 1}/{args.warmup} done")
    mesh.ctx.kernel_time("fps", reset=True)
    mesh.ctx.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pos, _, _ = sampler.sampleEvenlyBySeparation(n, seed=2, stream_id=W.STREAM_SAMPLES, device=True)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    mesh.ctx.set_timing(False)
    k_ms, k_n = mesh.ctx.kernel_time("fps", reset=True)
    log(f"{args.steps} calls: {sec * 1e3:.1f} ms each")
    result = {
        "metric": "separation samples/s (C2 mesh, 10k geodesic furthest-point picks among 30k area samples)",
        "value": n / sec, "unit": "samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sec * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 positions, f64 geodesic distances", "data": "synthetic",
        "config": {"workload": "sampleEvenlyBySeparation on the C2 mesh", "samples": n, "oversampling": 3 * n,
                   "max_degree": 8},
        "fps_kernel_avg_ms": k_ms / max(k_n, 1), "samples_returned": int(len(pos)),
    }
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        over = 3 * n
        opos = O.sample_area(v, t, over, 2, W.STREAM_SAMPLES)[0]
        idx, sep, _, _ = O.proximity_graph(opos, None, 8, O.graph_picks(over, 100, 0))
        m = min(n, 1500)
        t0 = time.perf_counter()
        O.fps(idx, sep, m)
        csec = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": m / csec, "unit": "samples/s", "cores": 1, "kind": "port",
                                  "sample": f"the first {m} picks over the same {over}-point graph ({csec:.1f} s; "
                                            f"each pick is one multi-source Dijkstra over the whole graph, as in "
                                            f"Thea, so the rate holds for all {n})"}
    print(json.dumps(result), flush=True)


def cube_rotations():
    """The 24 rotations that map the coordinate axes onto signed axes (ShapeAlign's axis-aligned
    starting rotations, Tools/ShapeAlign/ShapeAlign.cpp:350-366)."""
    import itertools
    import numpy as np
    out = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            m = np.zeros((3, 3))
            for i in range(3):
                m[i, perm[i]] = signs[i]
            if np.linalg.det(m) > 0:
                out.append(m)
    return np.array(out)


def run_icp(args):
    """ShapeAlign's rotation search: ICP (Icp3<double>::align, Thea's defaults: at most 10 iterations)
    of 1,000,000 source points onto a point BVH of 1,000,000 target points, from each of the 24
    axis-aligned rotations about the source centroid, all 24 in one batched call. Both point sets are
    area samples of the C3 torus (different seeds); the source is moved by 0.1 rad about (1, 2, 3)
    and (0.05, -0.03, 0.02). One step = one whole search."""
    import numpy as np
    import torch
    import thea_amd
    from thea_amd import workloads as W

    v, t, _, _, _ = W.torus_c3(n_queries=0, n_rays=0)
    mesh = thea_amd.Mesh(v, t)
    sampler = thea_amd.MeshSampler(mesh)
    n = args.queries or 1_000_000
    to, _, _ = sampler.sampleEvenlyByArea(n, seed=6, stream_id=W.STREAM_SAMPLES)
    frm, _, _ = sampler.sampleEvenlyByArea(n, seed=7, stream_id=W.STREAM_SAMPLES)
    ax = np.array([1, 2, 3.0]) / np.sqrt(14)
    Kx = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    R = np.eye(3) + np.sin(0.1) * Kx + (1 - np.cos(0.1)) * Kx @ Kx
    src = frm.astype(np.float64) @ R.T + np.array([0.05, -0.03, 0.02])
    c = src.mean(0)
    inits = np.stack([np.hstack([m, (c - m @ c).reshape(3, 1)]) for m in cube_rotations()])
    bvh = (thea_amd.PointBvhF64(to.astype(np.float64), leaf_size=args.leaf) if args.f64
           else thea_amd.PointBvh(to, leaf_size=args.leaf))
    icp = thea_amd.Icp3(kabsch=args.kabsch)
    for i in range(args.warmup):
        icp.alignBatch(src, bvh, initial=inits)
        log(f"warmup search {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tr, err, its = icp.alignBatch(src, bvh, initial=inits)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    log(f"{args.steps} searches: {sec * 1e3:.1f} ms each")
    # nearest-point queries: one initial pass per problem, then one per iteration after the first
    queries = int(n * (len(inits) + np.maximum(its - 1, 0).sum()))
    best = int(np.argmin(err))
    result = {
        "metric": "ShapeAlign 24-rotation ICP search, 1M source onto 1M target points (seconds per search)",
        "value": sec, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec * 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 fit, f64 nearest" if args.f64 else "f64 fit, f32 nearest", "data": "synthetic",
        "config": {"workload": "icp rotation search on C3 torus samples", "source_points": n, "target_points": n,
                   "starting_rotations": len(inits), "fit": "least squares" if args.kabsch else "Thea (U V^T)",
                   "target": "double point BVH" if args.f64 else "float point BVH", "leaf_size": args.leaf},
        "nearest_queries_per_s": queries / sec,
        "iterations": its.tolist(), "errors": err.tolist(), "best_rotation": best,
        "best_transform": (tr[best][:, :3] @ inits[best][:, :3]).tolist(),
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = icp_cpu_baseline(args, to, src, inits[0], sec, len(inits), n)
    print(json.dumps(result), flush=True)


def icp_cpu_baseline(args, to, src, init, gpu_sec, nrot, n):
    """The oracle ICP (numpy fit, kd-tree nearest points on all threads) on one starting rotation with
    a 100k-point prefix of the source, scaled to the whole search."""
    import numpy as np
    from oracle import icp as I
    from oracle import oracle as O
    nth = cpu_threads(args)
    kd = O.KdTree(points=to)
    m = min(100_000, n)
    t0 = time.perf_counter()
    _, _, its = I.icp_align(lambda x: kd.closest(x, mode=1, nthreads=nth)[1], I.xf_apply(init, src[:m]),
                            kabsch=args.kabsch)
    sec = time.perf_counter() - t0
    est = sec * (n / m) * nrot
    return {"value": est, "unit": "s", "cores": nth, "kind": "port",
            "sample": f"one of the {nrot} rotations on the first {m} source points ({sec:.2f} s, {its} iterations), "
                      f"scaled by {n / m:.0f} x {nrot} to the whole search",
            "speedup_of_gpu": est / gpu_sec}


def run_copy(args):
    """Device-to-device copy of 1 GiB (hipMemcpyAsync through torch): the measured HBM ceiling, and
    the plain copy of known size that scales the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = 1 << 28  # float32 elements = 1 GiB
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        y.copy_(x)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    gbps = 2 * n * 4 / (ms * 1e-3) / 1e9
    print(json.dumps({"metric": "device copy GB/s (read + write)", "value": gbps, "unit": "GB/s", "n_gpus": 1,
                      "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                      "config": {"workload": "1 GiB device-to-device copy", "bytes_read": n * 4,
                                 "bytes_written": n * 4},
                      "frac_of_peak": gbps / HBM_PEAK_GBPS}), flush=True)


def run_c4(args, counts):
    """Batched many-mesh build + query: 512 meshes (perturbed spheres and tori, log-uniform 1k-500k
    triangles), log-uniform 1k-200k queries per mesh uniform in each 20%-inflated bounding box. One
    step = one replay of the plan's captured HIP graph, which builds every mesh's BVH and answers
    every query (nearest triangle index and distance). Compared with plain dispatch of the same work:
    one thg_mesh_create + thg_bvh_build_mesh + thg_closest_point per mesh, with the same leaf sizes."""
    import numpy as np
    import torch
    import thea_amd
    from thea_amd import _capi as capi
    from thea_amd import workloads as W
    from thea_amd.batch import BatchPlan, autotune_plan, concat

    dev = torch.device("cuda", 0)
    t0 = time.time()
    meshes, queries = W.many_meshes_c4()
    nv, nt, nq, verts, tris, qs = concat(meshes, queries)
    del meshes, queries
    gen_s = time.time() - t0
    log(f"generated {len(nt)} meshes in {gen_s:.1f} s")
    V, T, Q = (torch.from_numpy(x).to(dev) for x in (verts, tris, qs))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    leaf, order, launch = autotune_plan(nv, nt, nq, V, T, Q)
    tune_s = time.perf_counter() - t0
    log(f"autotuned leaf sizes, child orders and launch shape {launch} in {tune_s:.1f} s")
    t0 = time.perf_counter()
    plan = BatchPlan(nv, nt, nq, leaf_sizes=leaf, child_orders=order, launch=launch)
    plan_s = time.perf_counter() - t0
    phases = []
    for _ in range(3):  # eager runs with per-phase HIP events on the plan's stream
        idx, dist, ph = plan.run(V, T, Q, timed=True)
        phases.append(ph)
        log(f"eager plan run: build {ph[0]:.2f} ms, query {ph[1]:.2f} ms")
    build_ms = float(np.mean([p[0] for p in phases[1:]]))
    query_ms = float(np.mean([p[1] for p in phases[1:]]))
    for _ in range(args.warmup):  # the first graph run captures, the rest replay
        idx, dist = plan.run(V, T, Q)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        idx, dist = plan.run(V, T, Q)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    log(f"{args.steps} graph replays: {ms_per_step:.2f} ms each")
    NQ, NT = int(nq.sum()), int(nt.sum())

    # plain dispatch of the same work, one mesh at a time
    L = capi.lib()
    stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    P = lambda x: C.c_void_p(x.data_ptr())  # noqa: E731
    pidx = torch.empty(NQ, dtype=torch.int32, device=dev)
    pdist = torch.empty(NQ, dtype=torch.float32, device=dev)
    voff = np.concatenate([[0], np.cumsum(nv)])
    toff = np.concatenate([[0], np.cumsum(nt)])
    qoff = np.concatenate([[0], np.cumsum(nq)])

    def plain(m):
        mesh = thea_amd.Mesh(V[voff[m]:voff[m + 1]], T[toff[m]:toff[m + 1]])
        b = thea_amd.MeshBvh(mesh, leaf_size=int(leaf[m]))
        n = int(nq[m])
        capi.check(L.thg_closest_point(b.h, P(Q[qoff[m]:]), n, -1.0, P(pidx[qoff[m]:]), None, P(pdist[qoff[m]:]),
                                       None, capi.THG_MEM_DEVICE, stream))
        torch.cuda.current_stream(dev).synchronize()
        b.close()
        mesh.close()
    for m in range(4):
        plain(m)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for m in range(len(nt)):
        plain(m)
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    log(f"plain dispatch over {len(nt)} meshes: {plain_s:.2f} s")
    mism = int((pidx != idx).sum().item()) + int((pdist.view(torch.int32) != dist.view(torch.int32)).sum().item())

    c = counts["c4_closest"]
    bq = 32 + 24 * c["V_box"] + 36 * c["V_tri"]
    build_gbps = 244.0 * NT / (build_ms * 1e-3) / 1e9
    # The query phase (query sort + one multi-segment closest-point launch) against the measured
    # ceilings of its kernel; the build phase's bytes and Thea's traversal bytes are diagnostics.
    roof = measured_roofline("closest_multi", "c4", query_ms, bq * NQ, torch.cuda.get_device_properties(dev).multi_processor_count,
                             f"bytes Thea's own traversal reads per query: {bq:.1f} B/query (per-mesh means "
                             f"weighted by query count); most are re-read from L2/MALL")
    roof["kernel_avg_ms_note"] = "query phase: query-key sort and the multi-segment launch, HIP events"
    roof["build_phase"] = {"ms": build_ms, "bytes_per_triangle": 244.0, "GB/s": build_gbps}
    roof["query_phase"] = {"ms": query_ms, "bytes_per_query": bq, "thea_traversal_counts": c}
    result = {
        "metric": "batched many-mesh build+query queries/s (512 meshes, 1k-500k tris, 1k-200k queries each)",
        "value": NQ / (ms_per_step * 1e-3), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched build + closest point over 512 meshes in one captured graph",
                   "meshes": int(len(nt)), "triangles_total": NT, "queries_total": NQ,
                   "leaf_sizes": {str(k): int(v) for k, v in zip(*np.unique(leaf, return_counts=True))},
                   "child_orders": {["majority", "nearest", "build"][k]: int(v)
                                    for k, v in zip(*np.unique(order, return_counts=True))},
                   "launch": {"waves_per_workgroup": launch[0], "eight_wave_build": bool(launch[1])}},
        "roofline": roof,
        "graph_replay_ms": ms_per_step, "eager_ms": build_ms + query_ms,
        "plain_dispatch": {"seconds": plain_s, "queries_per_s": NQ / plain_s, "speedup": plain_s / (ms_per_step * 1e-3),
                           "mismatches_vs_batched": mism},
        "autotune_s": tune_s, "plan_create_s": plan_s, "input_generation_s": gen_s,
    }
    print(json.dumps(result), flush=True)
    plan.close()


def run_c5(args):
    print(json.dumps(c5_measure(args)), flush=True)


def c5_measure(args):
    """Served request stream: 64 resident meshes (log-uniform 10k-4M triangles), requests with a
    uniform mesh, a log-uniform batch of 256-1M points and a 60/25/15 mix of closest point, k-NN
    (k = 8) and first-hit rays. Compares the server (continuous batching, multi-batch launches on 4
    streams) with plain dispatch (one C ABI call per request on one stream, in arrival order):
    saturation throughput with every request queued at once, then p50/p99 latency under Poisson
    arrivals at `--load` times the plain path's capacity."""
    import threading

    import numpy as np
    import torch
    import thea_amd
    from thea_amd import _capi as capi
    from thea_amd import workloads as W
    from thea_amd.serve import Server

    dev = torch.device("cuda", 0)
    t0 = time.time()
    meshes = W.resident_c5()
    gen_s = time.time() - t0
    srv = Server(num_streams=4)
    t0 = time.perf_counter()
    ids = [srv.add_mesh(v, t, leaf_size=args.c5_leaf) for v, t in meshes]
    srv_build_s = time.perf_counter() - t0
    bvhs = [thea_amd.MeshBvh(thea_amd.Mesh(v, t), leaf_size=args.c5_leaf) for v, t in meshes]
    g = W.rng(5, W.STREAM_TRACE)
    R = args.requests
    nmax = 1 << 20
    pools = []
    for v, _ in meshes:
        lo, hi = W.inflated_box(v, 1.2)
        lo_t = torch.tensor(lo, dtype=torch.float32, device=dev)
        hi_t = torch.tensor(hi, dtype=torch.float32, device=dev)
        pools.append(lo_t + (hi_t - lo_t) * torch.rand((nmax, 3), device=dev))
    dirs = torch.nn.functional.normalize(torch.randn((nmax, 3), device=dev), dim=1).contiguous()
    kinds = g.choice(["closest", "knn", "ray"], size=R, p=[0.6, 0.25, 0.15])
    mesh_of = g.integers(0, len(meshes), size=R)
    sizes = W.log_uniform_int(g, 256, 1 << 20, R)
    reqs = []
    for i in range(R):
        n = int(sizes[i])
        m = int(mesh_of[i])
        k = 8 if kinds[i] == "knn" else 0
        out = {"idx": torch.empty((n, k) if k else n, dtype=torch.int32, device=dev),
               "dist": torch.empty((n, k) if k else n, dtype=torch.float32, device=dev),
               "cp": None if k else torch.empty((n, 3), dtype=torch.float32, device=dev)}
        reqs.append(dict(kind=str(kinds[i]), mesh=m, n=n, k=k, q=pools[m][:n], dirs=dirs[:n] if kinds[i] == "ray"
                         else None, out=out))
    total_q = int(sizes.sum())
    L = capi.lib()
    stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    P = lambda x: None if x is None else C.c_void_p(x.data_ptr())  # noqa: E731

    def plain_call(r):
        b = bvhs[r["mesh"]]
        o = r["out"]
        if r["kind"] == "closest":
            capi.check(L.thg_closest_point(b.h, P(r["q"]), r["n"], -1.0, P(o["idx"]), P(o["cp"]), P(o["dist"]), None,
                                           capi.THG_MEM_DEVICE, stream))
        elif r["kind"] == "knn":
            capi.check(L.thg_k_closest(b.h, P(r["q"]), r["n"], r["k"], -1.0, 0, P(o["idx"]), P(o["dist"]), None,
                                       capi.THG_MEM_DEVICE, stream))
        else:
            capi.check(L.thg_ray_first_hit(b.h, P(r["q"]), P(r["dirs"]), r["n"], -1.0, P(o["idx"]), P(o["dist"]),
                                           P(o["cp"]), capi.THG_MEM_DEVICE, stream))
        torch.cuda.current_stream(dev).synchronize()

    def plain_run(arrivals):
        lat = []
        t_start = time.perf_counter()
        for r, a in zip(reqs, arrivals):
            while time.perf_counter() - t_start < a:
                pass
            plain_call(r)
            lat.append(time.perf_counter() - t_start - a)
        return time.perf_counter() - t_start, np.array(lat) * 1e6

    def server_run(arrivals, stats=None):
        srv.stats(reset=True)
        rids = [None] * len(arrivals)
        t_start = time.perf_counter()

        def submitter():
            for i, (r, a) in enumerate(zip(reqs, arrivals)):
                while time.perf_counter() - t_start < a:
                    pass
                rids[i] = srv.submit(r["kind"], ids[r["mesh"]], r["q"], r["dirs"], k=r["k"], outputs=r["out"])[0]
        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        lat = []
        for rid in rids:
            rc, l_us = srv.wait(rid)
            capi.check(rc)
            lat.append(l_us)
        elapsed = time.perf_counter() - t_start
        if stats is not None:
            stats.update(srv.stats())
        return elapsed, np.array(lat)

    log(f"{len(meshes)} meshes resident, {R} requests ({total_q} points) prepared")
    zero = np.zeros(R)
    # warm up both paths on the whole trace: every kernel variant the trace uses is loaded and the
    # stream-ordered memory pools have grown before anything is timed
    plain_run(zero)
    server_run(zero)
    log("warm-up pass over the trace done (plain dispatch and server)")
    t_plain, _ = plain_run(zero)
    log(f"plain dispatch, all queued: {t_plain:.2f} s")
    st_sat = {}
    t_srv, _ = server_run(zero, st_sat)
    log(f"server, all queued: {t_srv:.2f} s")
    cap_plain = R / t_plain
    lam = args.load * cap_plain
    arrivals = np.cumsum(g.exponential(1.0 / lam, size=R))
    tp, lat_p = plain_run(arrivals)
    log(f"plain dispatch, Poisson arrivals: {tp:.2f} s")
    st = {}
    ts, lat_s = server_run(arrivals, st)
    log(f"server, Poisson arrivals: {ts:.2f} s")

    def summ(t_total, lat):
        return {"requests_per_s": R / t_total, "queries_per_s": total_q / t_total,
                "p50_ms": float(np.percentile(lat, 50)) / 1e3, "p99_ms": float(np.percentile(lat, 99)) / 1e3}

    result = {
        "metric": "served mixed requests/s (64 resident meshes; closest 60%, k-NN k=8 25%, rays 15%)",
        "value": R / t_srv, "unit": "requests/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": t_srv * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "served request stream", "meshes": len(meshes),
                   "triangles_total": int(sum(len(t) for _, t in meshes)), "requests": R,
                   "queries_total": total_q, "batch": "log-uniform 256..1M points", "streams": 4},
        "saturation": {"server": summ(t_srv, np.zeros(1)) | {"p50_ms": None, "p99_ms": None},
                       "plain_dispatch": summ(t_plain, np.zeros(1)) | {"p50_ms": None, "p99_ms": None},
                       "speedup": t_plain / t_srv},
        "poisson": {"rate_requests_per_s": lam, "load_vs_plain_capacity": args.load,
                    "server": summ(ts, lat_s), "plain_dispatch": summ(tp, lat_p)},
        "requests_per_launch": {"saturation": st_sat.get("requests_per_launch"),
                                "poisson": st.get("requests_per_launch"),
                                "note": "requests a worker takes from the queue and serves with one multi-batch "
                                        "launch per (type, traversal, k) group"},
        "server_stats": {"saturation": st_sat, "poisson": st}, "mesh_generation_s": gen_s,
        "server_build_s": srv_build_s,
    }
    srv.close()
    for b in bvhs:
        b.close()
    return result


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
    elif wl != "c2sample":
        kd = O.KdTree(v, t)
    build_s = time.perf_counter() - t0
    # The whole batch on all threads when that is about 30 s of CPU work or less (C3: ~25 s on 16
    # threads), and a prefix of it on one thread.
    if wl in ("c3", "c1"):
        n_multi = min(len(q), 1_000_000 * nth)
        n_single = min(len(q), 60_000)
        run = lambda n, th: kd.closest(q[:n], mode=1, nthreads=th, want_cp=True)  # noqa: E731
    elif wl == "c3ray":
        o, d = rays
        n_multi, n_single = len(o), min(len(o), 1_000_000)
        run = lambda n, th: kd.ray(o[:n], d[:n], mode=1, nthreads=th)  # noqa: E731
    elif wl == "c2knn":
        n_multi, n_single = len(pos), 100_000
        run = lambda n, th: kd.knn(pos[:n], 16, mode=1, exclude_self=True, nthreads=th)  # noqa: E731
    else:  # c2sample: Thea's sampler is sequential (one CDF, one RNG stream)
        t0 = time.perf_counter()
        O.sample_area(v, t, 1_000_000, 2, 3)
        sec = time.perf_counter() - t0
        return {"value": 1_000_000 / sec, "unit": unit, "cores": 1, "kind": "port",
                "sample": f"1,000,000 samples of the same mesh on one thread ({sec:.2f} s, CDF included)"}
    t0 = time.perf_counter()
    run(n_multi, nth)
    multi_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    run(n_single, 1)
    single_s = time.perf_counter() - t0
    return {"value": n_multi / multi_s, "unit": unit, "cores": nth, "kind": "port",
            "sample": f"first {n_multi} queries on {nth} threads ({multi_s:.1f} s) and first {n_single} on 1 thread "
                      f"of the same batch; kd-tree build {build_s:.2f} s excluded",
            "single_thread_value": n_single / single_s, "build_s": build_s, "host": host_info()}


def host_info():
    """CPU model and the cores this process may use (the GPU box's share, not the machine's)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "usable_cores": usable, "nproc": os.cpu_count()}


if __name__ == "__main__":
    main()

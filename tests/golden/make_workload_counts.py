"""Writes tests/golden/workload_counts.json: per-query box and element test counts (V_box, V_tri)
made by the oracle's restatement of Thea's own traversal (kd-tree, <= 5 elements per leaf, x1.05
padding, nearest child first, prune when the lower bound is >= the best) on the seeded workloads.

These counts fix the algorithmic-bytes model of the roofline (bytes per query = query + result +
24 B per box test + 36 B per triangle test, or 12 B per point test), so it does not depend on
the GPU tree. Run: python tests/golden/make_workload_counts.py [--c4 | --rays]
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from thea_amd import workloads as W  # noqa: E402

SAMPLE = 1_000_000


def per_query(stats, n):
    return {"V_box": float(stats[0]) / n, "V_tri": float(stats[1]) / n, "V_node": float(stats[2]) / n}


def c4_counts(per_mesh=2000):
    """Batched many-mesh workload: counts on the first `per_mesh` queries of every mesh, weighted by
    each mesh's full query count."""
    meshes, queries = W.many_meshes_c4()
    nq = np.array([len(q) for q in queries], np.float64)
    vb = np.zeros(len(meshes))
    vt = np.zeros(len(meshes))
    for m, ((v, t), q) in enumerate(zip(meshes, queries)):
        kd = O.KdTree(v, t)
        sub = q[:per_mesh]
        _, _, _, st = kd.closest(sub, mode=1, nthreads=8)
        vb[m], vt[m] = float(st[0]) / len(sub), float(st[1]) / len(sub)
    w = nq / nq.sum()
    return {"V_box": float((vb * w).sum()), "V_tri": float((vt * w).sum()), "queries": int(nq.sum()),
            "triangles": int(sum(len(t) for _, t in meshes)), "sample_per_mesh": per_mesh,
            "note": "per-mesh means on the first queries of each mesh, weighted by its query count"}


def ray_feature_counts(n_sdf=20_000, n_vis=10_000):
    """The shape diameter's and visibility's rays on the C3 torus, generated as the bench generates
    them (area samples with the bench's seeds; SDF scale = the vertex bounding-box diagonal,
    visibility scale = the diagonal of the BVH bounds grown 1.05x, 100 rays and seed 1), on the
    first points of the batch. Thea's rayIntersects runs the first-hit traversal
    (BvhN.hpp:1266-1269), so both use first-hit counts."""
    v, t, _, _, _ = W.torus_c3(n_queries=0, n_rays=0)
    kd = O.KdTree(v, t)
    out = {}
    pos, nrm, _ = O.sample_area(v, t, n_sdf, 9, W.STREAM_SAMPLES)
    scale = np.float32(np.linalg.norm((v.max(0) - v.min(0)).astype(np.float32)))
    org = np.zeros((30 * n_sdf, 3), np.float32)
    dr = np.zeros((30 * n_sdf, 3), np.float32)
    O.lib().or_sdf_rays(pos, nrm, n_sdf, scale, org, dr)
    _, _, st = kd.ray(org, dr, mode=1)
    out["c3_sdf_ray"] = per_query(st, len(org)) | {"queries": len(org), "points": n_sdf}
    pos, _, _ = O.sample_area(v, t, n_vis, 10, W.STREAM_SAMPLES)
    lo, hi = v.min(0), v.max(0)
    c = np.float32(0.5) * (hi + lo)
    f = np.float32(1.05)
    e = (f * (hi - c) + c) - (f * (lo - c) + c)
    vscale = np.float32(np.sqrt(e[0] * e[0] + (e[1] * e[1] + e[2] * e[2])))
    org, dr = O.vis_rays(pos, 100, vscale, 1)
    _, _, st = kd.ray(org, dr, mode=1)
    out["c3_vis_ray"] = per_query(st, len(org)) | {"queries": len(org), "points": n_vis}
    return out


def main():
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "workload_counts.json")
    if "--rays" in sys.argv:  # add the ray features' counts to the existing file
        with open(path) as f:
            out = json.load(f)
        out.update(ray_feature_counts())
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
        print(json.dumps({k: out[k] for k in ("c3_sdf_ray", "c3_vis_ray")}, indent=1))
        return
    if "--c4" in sys.argv:  # add the batched many-mesh counts to the existing file
        with open(path) as f:
            out = json.load(f)
        out["c4_closest"] = c4_counts()
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
        print(json.dumps(out["c4_closest"], indent=1))
        return
    out = {"note": "Thea-order traversal counts from the oracle (mode 1) on the first SAMPLE queries",
           "sample": SAMPLE}
    t0 = time.time()
    v, t, q = W.sphere_c1()
    kd = O.KdTree(v, t)
    _, _, _, st = kd.closest(q, mode=1)
    out["c1_closest"] = per_query(st, len(q)) | {"queries": len(q)}

    v, t, q, ro, rd = W.torus_c3(n_queries=SAMPLE, n_rays=SAMPLE)
    kd = O.KdTree(v, t)
    _, _, _, st = kd.closest(q, mode=1)
    out["c3_closest"] = per_query(st, len(q)) | {"queries": len(q)}
    _, _, st = kd.ray(ro, rd, mode=1)
    out["c3_ray"] = per_query(st, len(ro)) | {"queries": len(ro)}

    v, t = W.ellipsoids_c2()
    pos, _, _ = O.sample_area(v, t, 1_000_000, 2, W.STREAM_SAMPLES)
    kp = O.KdTree(points=pos)
    sub = pos[:200_000]
    _, _, st = kp.knn(sub, 16, mode=1)
    out["c2_knn16"] = per_query(st, len(sub)) | {"queries": len(sub), "note": "V_tri counts point tests"}
    out["seconds"] = time.time() - t0
    out["c4_closest"] = c4_counts()
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()

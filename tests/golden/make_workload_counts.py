"""Writes tests/golden/workload_counts.json: per-query box and element test counts (V_box, V_tri)
made by the oracle's restatement of Thea's own traversal (kd-tree, <= 5 elements per leaf, x1.05
padding, nearest child first, prune when the lower bound is >= the best) on the seeded workloads.

These counts fix the algorithmic-bytes model of the roofline (bytes per query = query + result +
24 B per box test + 36 B per triangle test, or 12 B per point test), so it does not depend on
the GPU tree. Run: python tests/golden/make_workload_counts.py [--c4]
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


def main():
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "workload_counts.json")
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

"""TEST INFRASTRUCTURE ONLY -- numpy restatement of Thea's Icp3<double> (Algorithms/Icp3.hpp).

Only tests/ and bench.py's cpu_baseline leg may import this module, and only as the checker.

The nearest points come from the oracle's kd-tree (canonical mode) over the target, queried with the
source positions rounded to float32, as the GPU path does on a float BVH, or (f64=True) from the
oracle's double brute force with the positions in double, as on a double point BVH; every other
quantity is double. The 3x3
fit uses numpy's SVD (LAPACK) for U V^T, an algorithm independent of the library's Jacobi sweeps.
Sums are numpy's (pairwise), so results agree with the GPU to rounding, not bit for bit.
"""
import numpy as np


def xf_apply(tr, p):
    """tr (3, 4) applied to points p (n, 3): ((L_i0 x + L_i1 y) + L_i2 z) + t_i, elementwise in double
    (the operation order the GPU path uses, so positions round to the same floats)."""
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    out = np.empty_like(p)
    for i in range(3):
        out[:, i] = ((tr[i, 0] * x + tr[i, 1] * y) + tr[i, 2] * z) + tr[i, 3]
    return out


def compose(a, b):
    """a * b, b applied first (AffineTransformN.hpp:154-157)."""
    r = np.zeros((3, 4))
    r[:, :3] = a[:, :3] @ b[:, :3]
    r[:, 3] = a[:, :3] @ b[:, 3] + a[:, 3]
    return r


def identity():
    return np.hstack([np.eye(3), np.zeros((3, 1))])


def align_one_step(p, q, wt=None, wr=None, up=None, kabsch=False):
    """Icp3::alignOneStep with a known mapping p[i] -> q[i] (Icp3.hpp:288-375), no symmetry planes.
    Thea takes rot = U V^T of cov = sum dp dq^T (Icp3.hpp:349-350), the transpose of the rotation that
    minimises sum |rot dp - dq|^2 (V U^T); kabsch=True returns the minimiser instead."""
    if wt is None:
        pm, qm = p.sum(0) / len(p), q.sum(0) / len(p)
    else:
        pm, qm = (wt[:, None] * p).sum(0) / wt.sum(), (wt[:, None] * q).sum(0) / wt.sum()
    dp, dq = p - pm, q - qm
    if up is not None:  # Icp3.hpp:333-337
        dp = dp - np.outer(dp @ up, up)
        dq = dq - np.outer(dq @ up, up)
    if wr is not None:
        dp, dq = dp * wr[:, None], dq * wr[:, None]
    cov = dp.T @ dq
    u, _, vt = np.linalg.svd(cov)
    rot = u @ vt
    if kabsch:
        rot = rot.T
    tr = np.zeros((3, 4))
    tr[:, :3] = rot
    tr[:, 3] = qm - rot @ pm
    return tr


def measure_error(tr, p, q, wt=None, wr=None):
    """Icp3::measureError with a known mapping (Icp3.hpp:422-465)."""
    if wt is None:
        pm, qm = p.sum(0) / len(p), q.sum(0) / len(p)
    else:
        pm, qm = (wt[:, None] * p).sum(0) / wt.sum(), (wt[:, None] * q).sum(0) / wt.sum()
    d = tr[:, :3] @ pm + tr[:, 3] - qm
    err = float(d @ d)
    r = (p - pm) @ tr[:, :3].T - (q - qm)
    e = (r * r).sum(1)
    return err + float((e if wr is None else wr * wr * e).sum())


def icp_align(nearest, src, frac_thr=0.0001, min_it=3, max_it=10, wt=None, wr=None, up=None, kabsch=False,
              f64=False):
    """Icp3::align (Icp3.hpp:193-263). nearest(points (n, 3)) -> closest points (n, 3), given the
    positions rounded to float32, or in double with f64=True (Icp3<double> over a BvhN<Vector3d>,
    ShapeAlign.cpp:339); src: the source points (double). Returns (transform (3, 4), error,
    iterations)."""
    if up is not None:
        up = np.asarray(up, np.float64) / np.linalg.norm(up)

    def mapping(pts):
        return np.asarray(nearest(pts if f64 else pts.astype(np.float32)), np.float64)

    from_points = src.copy()
    old_tr = identity()
    to_points = mapping(from_points)
    old_err = measure_error(old_tr, from_points, to_points, wt, wr)
    if old_err <= np.finfo(np.float64).tiny:
        return old_tr, old_err, 0
    new_err, new_tr, its = old_err, old_tr, 0
    for i in range(max_it):
        inc = align_one_step(from_points, to_points, wt, wr, up, kabsch)
        its = i + 1
        new_tr = compose(inc, old_tr)
        to_points = mapping(from_points)  # measureError(inc, from_points, to): same positions
        new_err = measure_error(inc, from_points, to_points, wt, wr)
        if i < max_it - 1:
            frac = (old_err - new_err) / old_err
            if i >= min_it and frac < frac_thr:
                return (new_tr, new_err, its) if frac > 0 else (old_tr, old_err, its)
            from_points = xf_apply(new_tr, src)
        old_err, old_tr = new_err, new_tr
    return new_tr, new_err, its

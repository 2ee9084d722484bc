"""Pins the numpy restatement of Thea's Icp3 (oracle/icp.py) with closed forms: the least-squares
fit recovers exact rigid transforms (with weights and with an up vector) and Thea's fit returns its
transposed rotation (rot = U V^T of sum dp dq^T, Icp3.hpp:349-350), the error of a pure translation
is its squared length, identical sets stop at once, and the iteration structure (the fit of step i
uses the mapping of the positions before step i - 1 moved them) is Thea's."""
import numpy as np

from oracle import icp as I


def _rot(g):
    q, r = np.linalg.qr(g.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def _rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])


def test_one_step_recovers_rigid_transform():
    g = np.random.default_rng(1)
    for _ in range(20):
        R, t = _rot(g), g.normal(size=3)
        p = g.normal(size=(500, 3))
        q = p @ R.T + t
        tr = I.align_one_step(p, q, kabsch=True)
        assert np.allclose(tr[:, :3], R, atol=1e-12) and np.allclose(tr[:, 3], t, atol=1e-12)
        wt, wr = g.uniform(0.1, 2, 500), g.uniform(0.1, 2, 500)
        tr = I.align_one_step(p, q, wt, wr, kabsch=True)
        assert np.allclose(tr[:, :3], R, atol=1e-12) and np.allclose(tr[:, 3], t, atol=1e-12)
        assert I.measure_error(tr, p, q) < 1e-20 * len(p) * (1 + np.abs(q).max()) ** 2
        # Thea's fit: the transposed rotation, translation q_mean - rot p_mean
        th = I.align_one_step(p, q, wt, wr)
        pm = (wt[:, None] * p).sum(0) / wt.sum()
        assert np.allclose(th[:, :3], R.T, atol=1e-12)
        assert np.allclose(th[:, 3], R @ pm + t - R.T @ pm, atol=1e-12)


def test_one_step_with_up_vector_rotates_about_it():
    g = np.random.default_rng(2)
    p = g.normal(size=(400, 3))
    R = _rz(0.7)
    q = p @ R.T + np.array([0.3, -0.2, 0.5])
    for kabsch, want in ((True, R), (False, R.T)):
        tr = I.align_one_step(p, q, up=np.array([0, 0, 1.0]), kabsch=kabsch)
        assert np.allclose(tr[:, :3], want, atol=1e-12)
        # a tilted target: the constrained fit still fixes the up axis
        T = _rot(g)
        tr = I.align_one_step(p, p @ T.T, up=np.array([0, 0, 1.0]), kabsch=kabsch)
        assert np.allclose(tr[2, :3], [0, 0, 1], atol=1e-12) and np.allclose(tr[:3, 2], [0, 0, 1], atol=1e-12)


def test_error_of_translation_and_identity():
    g = np.random.default_rng(3)
    p = g.normal(size=(300, 3))
    c = np.array([0.5, -1.0, 2.0])
    assert np.isclose(I.measure_error(I.identity(), p, p + c), c @ c, rtol=1e-12)
    w = g.uniform(0.5, 1.5, 300)
    assert np.isclose(I.measure_error(I.identity(), p, p + c, wt=w, wr=w), c @ c, rtol=1e-12)


def test_loop_identical_sets_and_structure():
    g = np.random.default_rng(4)
    to = g.normal(size=(2000, 3)).astype(np.float32).astype(np.float64)

    def nearest(x):
        d = ((x[:, None, :].astype(np.float64) - to[None, :, :]) ** 2).sum(-1)
        return to[d.argmin(1)]
    tr, err, its = I.icp_align(nearest, to.copy())
    assert its == 0 and err == 0 and np.array_equal(tr, I.identity())
    # one step: the fit to the initial mapping, and its error on that mapping
    R = _rot(g)
    src = (to[:400] - to[:400].mean(0)) @ (np.eye(3) + 0.05 * (R - np.eye(3))).T + 0.01
    m0 = nearest(src.astype(np.float32))
    tr1, err1, its1 = I.icp_align(nearest, src, max_it=1)
    inc = I.align_one_step(src, m0)
    assert its1 == 1 and np.allclose(tr1, inc, atol=1e-14) and np.isclose(err1, I.measure_error(inc, src, m0))
    # two steps: Thea fits the moved points to the mapping measured before they moved
    tr2, _, its2 = I.icp_align(nearest, src, max_it=2, min_it=5)
    moved = I.xf_apply(tr1, src)
    inc2 = I.align_one_step(moved, m0)
    assert its2 == 2 and np.allclose(tr2, I.compose(inc2, tr1), atol=1e-12)


def test_loop_converges_on_small_rotation():
    """The least-squares fit converges on an ellipsoid moved by a small oblique rotation; Thea's
    transposed fit turns the other way, its error grows, and the loop returns its previous transform."""
    import scipy.spatial
    g = np.random.default_rng(5)
    u = g.normal(size=(20000, 3))
    to = (u / np.linalg.norm(u, axis=1)[:, None] * np.array([1.0, 0.6, 0.3])).astype(np.float32).astype(np.float64)
    tree = scipy.spatial.cKDTree(to)

    def nearest(x):
        return to[tree.query(x.astype(np.float64))[1]]
    ax = np.array([1, 2, 3.0]) / np.sqrt(14)
    Kx = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    R = np.eye(3) + np.sin(0.08) * Kx + (1 - np.cos(0.08)) * Kx @ Kx
    src = to[::10] @ R.T + np.array([0.02, -0.01, 0.03])
    e0 = I.measure_error(I.identity(), src, nearest(src))
    tr, err, its = I.icp_align(nearest, src, max_it=30, frac_thr=1e-9, kabsch=True)
    assert err < 0.1 * e0 and np.abs(tr[:, :3] @ R - np.eye(3)).max() < 0.03
    tr, err, its = I.icp_align(nearest, src, max_it=30, frac_thr=1e-9)
    assert its == 4 and err > e0


def test_double_nearest_point_oracle_against_kdtree():
    """The oracle's double brute-force nearest point (BvhN<Vector3d>::closestElement semantics)
    against scipy's kd-tree, an independent search: same indices and distances to the last bit on
    random points (exact ties aside), the strict bound, and lowest index on exact duplicates."""
    from scipy.spatial import cKDTree
    from oracle import oracle as O
    g = np.random.default_rng(41)
    pts = g.normal(size=(4000, 3))
    q = g.normal(size=(3000, 3)) * 1.5
    idx, d2 = O.brute_closest_points_f64(pts, q)
    dd, ii = cKDTree(pts).query(q)
    assert np.array_equal(idx, ii)
    d = pts[idx] - q
    assert np.array_equal(d2, (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
    assert np.allclose(np.sqrt(d2), dd, rtol=1e-15, atol=0)
    # bound: admitted iff d^2 < bound^2
    b = float(np.median(np.sqrt(d2)))
    bidx, bd2 = O.brute_closest_points_f64(pts, q, b)
    inside = d2 < b * b
    assert np.array_equal(bidx[inside], idx[inside]) and np.all(bidx[~inside] == -1) and np.all(bd2[~inside] == -1)
    # duplicates: lowest index
    dup = np.concatenate([pts, pts])
    didx, _ = O.brute_closest_points_f64(dup, q)
    assert np.array_equal(didx, idx)

"""Seeded synthetic workloads for the BVH-query and sampling path (float32 throughout).

  sphere_c1      geodesic sphere, 5 subdivisions of an icosahedron (20,480 triangles, 10,242
                 vertices; Thea's GeodesicSphere3, Algorithms/GeodesicSphere3.cpp:28-134), vertex
                 radii perturbed by N(0, 0.02); 100,000 queries uniform in [-1.5, 1.5]^3.
  ellipsoids_c2  64 ellipsoids, semi-axes U[0.05, 0.3], centres U[-1, 1]^3, 21 stacks x 50 slices
                 (2,000 triangles each, 128,000 total); area-weighted samples and k-NN over them.
  torus_c3       torus R = 1, r = 0.3 on a 1024 x 512 grid (1,048,576 triangles), minor radius
                 displaced by 0.05 x periodic gradient noise; queries uniform in the bounding box
                 inflated 20%; rays with origins in that box and directions uniform on the sphere.
  many_meshes_c4 512 perturbed spheres/tori, triangle counts log-uniform in [1k, 500k], query
                 counts log-uniform in [1k, 200k] uniform in each inflated box.
  resident_c5    64 meshes, log-uniform 10k-4M triangles, for the request server.

Randomness: numpy Philox streams keyed by (seed, stream id), so every array is reproducible.
"""
import math

import numpy as np

STREAM_MESH, STREAM_QUERIES, STREAM_RAYS, STREAM_SAMPLES, STREAM_TRACE = 0, 1, 2, 3, 4


def rng(seed, stream):
    return np.random.Generator(np.random.Philox(key=np.array([seed, stream], dtype=np.uint64)))


# --------------------------------------------------------------------------------------------------
# Geodesic sphere (restates GeodesicSphere3.cpp:28-134 in float32, Eigen-order normalisation)
# --------------------------------------------------------------------------------------------------
_ICO_X = np.float32(0.525731112119133606)
_ICO_Z = np.float32(0.850650808352039932)
_ICO_TRIS = [(1, 4, 0), (4, 9, 0), (4, 5, 9), (8, 5, 4), (1, 8, 4), (1, 10, 8), (10, 3, 8), (8, 3, 5), (3, 2, 5),
             (3, 7, 2), (3, 10, 7), (10, 6, 7), (6, 11, 7), (6, 0, 11), (6, 1, 0), (10, 1, 6), (11, 0, 9),
             (2, 11, 9), (5, 2, 9), (11, 2, 7)]


def _normalized32(s):
    x, y, z = s
    n = x * x + (y * y + z * z)
    if n > 0:
        r = np.sqrt(n)
        return (x / r, y / r, z / r)
    return s


def geodesic_sphere(depth):
    X, Z, O = _ICO_X, _ICO_Z, np.float32(0)
    verts = [(-X, O, Z), (X, O, Z), (-X, O, -Z), (X, O, -Z), (O, Z, X), (O, Z, -X), (O, -Z, X), (O, -Z, -X),
             (Z, X, O), (-Z, X, O), (Z, -X, O), (-Z, -X, O)]
    tris = []
    mids = {}

    def midpoint(i, j):
        key = (i, j) if i < j else (j, i)
        if key in mids:
            return mids[key]
        a, b = verts[i], verts[j]
        p = _normalized32((a[0] + b[0], a[1] + b[1], a[2] + b[2]))
        mids[key] = len(verts)
        verts.append(p)
        return len(verts) - 1

    def sub(d, i0, i1, i2):
        if d <= 0:
            tris.append((i0, i1, i2))
            return
        m01 = midpoint(i0, i1)
        m12 = midpoint(i1, i2)
        m20 = midpoint(i2, i0)
        sub(d - 1, i0, m01, m20)
        sub(d - 1, i1, m12, m01)
        sub(d - 1, i2, m20, m12)
        sub(d - 1, m01, m12, m20)

    if depth == 0:
        tris = list(_ICO_TRIS)
    else:
        for t in _ICO_TRIS:
            sub(depth, *t)
    return np.array(verts, dtype=np.float32), np.array(tris, dtype=np.int32)


def sphere_c1(seed=1, n_queries=100_000):
    v, t = geodesic_sphere(5)
    r = rng(seed, STREAM_MESH)
    scale = (1.0 + r.normal(0.0, 0.02, size=len(v))).astype(np.float32)
    v = (v * scale[:, None]).astype(np.float32)
    q = rng(seed, STREAM_QUERIES).uniform(-1.5, 1.5, size=(n_queries, 3)).astype(np.float32)
    return v, t, q


# --------------------------------------------------------------------------------------------------
# Parametric meshes
# --------------------------------------------------------------------------------------------------
def uv_sphere(rings, slices):
    """Unit UV sphere: `rings` latitude rings plus two poles; 2 * slices * rings triangles."""
    th = np.pi * (np.arange(rings) + 1) / (rings + 1)
    ph = 2 * np.pi * np.arange(slices) / slices
    st, ct = np.sin(th)[:, None], np.cos(th)[:, None]
    ring = np.stack([st * np.cos(ph)[None, :], st * np.sin(ph)[None, :], np.broadcast_to(ct, (rings, slices))], -1)
    v = np.concatenate([[[0, 0, 1]], ring.reshape(-1, 3), [[0, 0, -1]]]).astype(np.float64)
    top, bot = 0, 1 + rings * slices
    idx = lambda i, j: 1 + i * slices + (j % slices)  # noqa: E731
    j = np.arange(slices)
    tris = [np.stack([np.full(slices, top), idx(0, j), idx(0, j + 1)], -1)]
    for i in range(rings - 1):
        a, b, c, d = idx(i, j), idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)
        tris.append(np.stack([a, b, c], -1))
        tris.append(np.stack([a, c, d], -1))
    tris.append(np.stack([np.full(slices, bot), idx(rings - 1, j + 1), idx(rings - 1, j)], -1))
    return v, np.concatenate(tris).astype(np.int32)


def _gradient_noise(u, v, gu, gv, r):
    """Periodic 2-D gradient (Perlin-style) noise on a gu x gv lattice; u, v in [0, 1)."""
    ang = r.uniform(0, 2 * np.pi, size=(gu, gv))
    gx, gy = np.cos(ang), np.sin(ang)
    x, y = u * gu, v * gv
    x0, y0 = np.floor(x).astype(np.int64), np.floor(y).astype(np.int64)
    fx, fy = x - x0, y - y0

    def dotg(ix, iy, dx, dy):
        ix, iy = ix % gu, iy % gv
        return gx[ix, iy] * dx + gy[ix, iy] * dy

    fade = lambda t: t * t * t * (t * (t * 6 - 15) + 10)  # noqa: E731
    n00 = dotg(x0, y0, fx, fy)
    n10 = dotg(x0 + 1, y0, fx - 1, fy)
    n01 = dotg(x0, y0 + 1, fx, fy - 1)
    n11 = dotg(x0 + 1, y0 + 1, fx - 1, fy - 1)
    sx, sy = fade(fx), fade(fy)
    nx0 = n00 + sx * (n10 - n00)
    nx1 = n01 + sx * (n11 - n01)
    return nx0 + sy * (nx1 - nx0)


def torus(nu, nv, R=1.0, r=0.3, amp=0.05, seed=3, lattice=(32, 16)):
    """Torus on an nu x nv grid (2 nu nv triangles), minor radius r + amp * noise in [-1, 1]."""
    g = rng(seed, STREAM_MESH)
    u = (np.arange(nu) / nu)[:, None]
    w = (np.arange(nv) / nv)[None, :]
    noise = _gradient_noise(np.broadcast_to(u, (nu, nv)), np.broadcast_to(w, (nu, nv)), lattice[0], lattice[1], g)
    m = np.abs(noise).max()
    rr = r + amp * (noise / m if m > 0 else noise)
    th, ph = 2 * np.pi * u, 2 * np.pi * w
    x = (R + rr * np.cos(ph)) * np.cos(th)
    y = (R + rr * np.cos(ph)) * np.sin(th)
    z = rr * np.sin(ph)
    v = np.stack([x, y, z], -1).reshape(-1, 3).astype(np.float32)
    i = np.arange(nu)[:, None]
    j = np.arange(nv)[None, :]
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    t = np.stack([np.stack([a, b, c], -1), np.stack([a, c, d], -1)], 2).reshape(-1, 3).astype(np.int32)
    return v, t


def inflated_box(v, factor=1.2):
    lo, hi = v.min(0).astype(np.float64), v.max(0).astype(np.float64)
    c, h = (lo + hi) / 2, (hi - lo) / 2 * factor
    return c - h, c + h


def uniform_in_box(lo, hi, n, g):
    return (lo + (hi - lo) * g.random((n, 3))).astype(np.float32)


def unit_directions(n, g):
    z = 2 * g.random(n) - 1
    ph = 2 * np.pi * g.random(n)
    s = np.sqrt(np.maximum(0.0, 1 - z * z))
    return np.stack([s * np.cos(ph), s * np.sin(ph), z], -1).astype(np.float32)


def torus_c3(seed=3, n_queries=10_000_000, n_rays=4_000_000, nu=1024, nv=512):
    v, t = torus(nu, nv, seed=seed)
    lo, hi = inflated_box(v, 1.2)
    q = uniform_in_box(lo, hi, n_queries, rng(seed, STREAM_QUERIES)) if n_queries else None
    if n_rays:
        g = rng(seed, STREAM_RAYS)
        ro = uniform_in_box(lo, hi, n_rays, g)
        rd = unit_directions(n_rays, g)
    else:
        ro = rd = None
    return v, t, q, ro, rd


def ellipsoids_c2(seed=2, n_ellipsoids=64, rings=20, slices=50):
    g = rng(seed, STREAM_MESH)
    sv, st = uv_sphere(rings, slices)
    vs, ts = [], []
    for e in range(n_ellipsoids):
        ax = g.uniform(0.05, 0.3, size=3)
        c = g.uniform(-1, 1, size=3)
        vs.append((sv * ax + c).astype(np.float32))
        ts.append(st + e * len(sv))
    return np.concatenate(vs), np.concatenate(ts).astype(np.int32)


def _mesh_of_size(T, kind, g):
    """A perturbed sphere (kind 0) or torus (kind 1) with about T triangles."""
    if kind == 0:
        rings = max(2, int(round(math.sqrt(T / 4))))
        v, t = uv_sphere(rings, 2 * rings)
        v = v * g.uniform(0.5, 2.0)
        v = v * (1 + 0.02 * g.normal(size=len(v)))[:, None]
    else:
        nv = max(3, int(round(math.sqrt(T / 4))))
        v, t = torus(2 * nv, nv, R=1.0, r=g.uniform(0.15, 0.4), amp=0.03, seed=int(g.integers(1 << 31)))
        v = v * g.uniform(0.5, 2.0)
    v = v + g.uniform(-10, 10, size=3)
    return v.astype(np.float32), t.astype(np.int32)


def log_uniform_int(g, lo, hi, n):
    return np.exp(g.uniform(math.log(lo), math.log(hi), size=n)).round().astype(np.int64)


def many_meshes_c4(seed=4, n_meshes=512, tri_range=(1_000, 500_000), q_range=(1_000, 200_000)):
    g = rng(seed, STREAM_MESH)
    tcount = log_uniform_int(g, *tri_range, n_meshes)
    qcount = log_uniform_int(g, *q_range, n_meshes)
    meshes, queries = [], []
    gq = rng(seed, STREAM_QUERIES)
    for m in range(n_meshes):
        v, t = _mesh_of_size(int(tcount[m]), m % 2, g)
        lo, hi = inflated_box(v, 1.2)
        meshes.append((v, t))
        queries.append(uniform_in_box(lo, hi, int(qcount[m]), gq))
    return meshes, queries


def resident_c5(seed=5, n_meshes=64, tri_range=(10_000, 4_000_000)):
    g = rng(seed, STREAM_MESH)
    tcount = log_uniform_int(g, *tri_range, n_meshes)
    return [_mesh_of_size(int(tcount[m]), m % 2, g) for m in range(n_meshes)]

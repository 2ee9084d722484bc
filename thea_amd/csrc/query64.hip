// Nearest point in double precision: BvhN<Vector3d, 3, double>::closestElement over a point set
// (Algorithms/BvhN.hpp:1050-1063 with MetricL2's point-point distance, MetricL2.hpp:176-193), the
// target structure of ShapeAlign's ICP (Tools/ShapeAlign/ShapeAlign.cpp:339, Icp3.hpp:388-404).
//
// The tree is the float LBVH of the points rounded to float (build_bvh); the leaves carry the
// caller's double coordinates (PointRecD, leaf order). Node boxes only prune: the query is widened
// to a box of half-width 2^-16 max(|q|, scene scale) as in the float kernels, which covers the
// float rounding of both the query and the points (2^-24 relative each), so a box lower bound never
// exceeds the double squared distance of a point inside. The answer is the lexicographic minimum of
// (double d^2, index), d^2 summed in Eigen's fixed-size order for three doubles, (x0 + x1) + x2
// (Redux.h vectorises Vector3d with a packet of two and adds the third term last).
//
// One wave walks the tree for 64 spatially sorted queries (as the float wave kernels do): a child
// is entered if some lane's bound admits it, and an entered node's leaves (at most 2 x 16 points)
// arrive with one vector load per lane while the next node's record is loaded.
#include "kernels_common.hpp"
#include "query.hpp"
#include "query64.hpp"
#include "traverse.hpp"

namespace thg {

namespace {

constexpr int B64 = 256;
constexpr int B64_WAVES = B64 / WAVE;

__device__ __forceinline__ int32_t uni64(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float unif64(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ int32_t lread(int32_t x, int32_t l) { return __builtin_amdgcn_readlane(x, l); }
__device__ __forceinline__ float lreadf(float x, int32_t l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double lreadd(double x, int32_t l) {
  const int64_t b = __double_as_longlong(x);
  const int32_t lo = __builtin_amdgcn_readlane((int32_t)(uint32_t)b, l);
  const int32_t hi = __builtin_amdgcn_readlane((int32_t)(b >> 32), l);
  return __longlong_as_double(((int64_t)hi << 32) | (uint32_t)lo);
}

__device__ __forceinline__ bool better64(double d, int32_t i, double bd, int32_t bi) {
  return d < bd || (d == bd && i < bi);
}

// Smallest float >= x (x finite or +inf; -1 stays -1).
__device__ __forceinline__ float float_up(double x) {
  float f = (float)x;
  if ((double)f < x) f = __int_as_float(__float_as_int(f) + (f >= 0.0f ? 1 : -1));
  return f;
}

__global__ void __launch_bounds__(256) k_point_records64(const double *__restrict__ xyz,
                                                         const PointRec *__restrict__ pr, int64_t n,
                                                         PointRecD *__restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const int32_t id = __float_as_int(pr[k].p.w);
  PointRecD r;
  r.x = xyz[3 * (int64_t)id];
  r.y = xyz[3 * (int64_t)id + 1];
  r.z = xyz[3 * (int64_t)id + 2];
  r.id = id;
  r.pad = 0;
  out[k] = r;
}

__global__ void __launch_bounds__(256) k_round_points(const double *__restrict__ xyz, int64_t n3,
                                                      float *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n3) out[i] = (float)xyz[i];
}

__device__ __forceinline__ uint32_t spread10_64(uint32_t v) {
  v &= 0x3FF;
  v = (v | (v << 16)) & 0x030000FF;
  v = (v | (v << 8)) & 0x0300F00F;
  v = (v | (v << 4)) & 0x030C30C3;
  v = (v | (v << 2)) & 0x09249249;
  return v;
}

// Morton key of each double query (rounded to float) in the scene box, for the spatial sort.
__global__ void __launch_bounds__(256) k_query_keys64(const double *__restrict__ q, int64_t n, float3 lo, float3 inv,
                                                      uint32_t *__restrict__ keys, uint32_t *__restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float px = (float)q[3 * i], py = (float)q[3 * i + 1], pz = (float)q[3 * i + 2];
  const float x = (px - lo.x) * inv.x, y = (py - lo.y) * inv.y, z = (pz - lo.z) * inv.z;
  const uint32_t ix = (uint32_t)fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
  const uint32_t iy = (uint32_t)fminf(fmaxf(y * 1024.0f, 0.0f), 1023.0f);
  const uint32_t iz = (uint32_t)fminf(fmaxf(z * 1024.0f, 0.0f), 1023.0f);
  keys[i] = (spread10_64(ix) << 2) | (spread10_64(iy) << 1) | spread10_64(iz);
  vals[i] = (uint32_t)i;
}

__global__ void __launch_bounds__(B64) k_closest64(const BNode *__restrict__ nodes, const PointRecD *__restrict__ pr,
                                                   const double *__restrict__ q, const uint32_t *__restrict__ order,
                                                   int64_t n, uint32_t nb_real, double best_init, float scale,
                                                   int32_t *__restrict__ out_idx, double *__restrict__ out_cp,
                                                   double *__restrict__ out_dist, double *__restrict__ out_d2) {
  const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t wave = (int64_t)blk * B64_WAVES + (threadIdx.x / WAVE);
  const int lane = lane_id();
  const int64_t t = wave * WAVE + lane;
  const bool inrange = blk < nb_real && t < n;
  bool active = inrange;
  int64_t qi = 0;
  double qx = 0, qy = 0, qz = 0;
  if (active) {
    qi = order ? (int64_t)order[t] : t;
    qx = q[3 * qi];
    qy = q[3 * qi + 1];
    qz = q[3 * qi + 2];
    active = !(isnan(qx) || isnan(qy) || isnan(qz));
  }
  const uint64_t amask = __ballot(active);
  const int nact = __popcll(amask);
  const float fx = (float)qx, fy = (float)qy, fz = (float)qz;
  const float m = fmaxf(scale, fmaxf(fabsf(fx), fmaxf(fabsf(fy), fabsf(fz))));
  const float sl = m * SLACK_REL;
  const float3 qlo = f3(fx - sl, fy - sl, fz - sl), qhi = f3(fx + sl, fy + sl, fz + sl);
  double bd = active ? best_init : -1.0;  // an inactive lane admits nothing (every d^2 >= 0)
  int32_t bi = 0x7fffffff;
  int32_t bslot = -1;
  // Seed: the nearest leaf of each lane's own descent gives a real bound before the shared walk.
  if (active) {
    int32_t f, cnt;
    seed_leaf(nodes, 0, qlo, qhi, f, cnt);
    for (int32_t e = f; e < f + cnt; ++e) {
      const PointRecD r = pr[e];
      const double dx = r.x - qx, dy = r.y - qy, dz = r.z - qz;
      const double d = (dx * dx + dy * dy) + dz * dz;
      if (better64(d, r.id, bd, bi)) { bd = d; bi = r.id; bslot = e; }
    }
  }
  int32_t s_ref = 0;
  float s_lx = 0, s_ly = 0, s_lz = 0, s_hx = 0, s_hy = 0, s_hz = 0;
  int32_t sp = 0;
  int32_t node = 0;
  BNode nd = nodes[0];
  while (amask != 0) {
    node = uni64(node);
    const float a0 = unif64(nd.a.x), a1 = unif64(nd.a.y), a2 = unif64(nd.a.z), a3 = unif64(nd.a.w);
    const float b0 = unif64(nd.b.x), b1 = unif64(nd.b.y), b2 = unif64(nd.b.z), b3 = unif64(nd.b.w);
    const float c0 = unif64(nd.c.x), c1 = unif64(nd.c.y), c2 = unif64(nd.c.z), c3 = unif64(nd.c.w);
    const int32_t r0 = uni64(nd.d.x), r1 = uni64(nd.d.y), n0 = uni64(nd.d.z), n1 = uni64(nd.d.w);
    const float bf = float_up(bd);
    const float lb0 = box_lb(qlo, qhi, a0, a1, a2, a3, b0, b1);
    const float lb1 = box_lb(qlo, qhi, b2, b3, c0, c1, c2, c3);
    const uint64_t m0 = __ballot(lb0 <= bf), m1 = __ballot(lb1 <= bf);
    const bool first1 = 2 * __popcll(__ballot(lb1 < lb0) & amask) > nact;
    int32_t next = -1;
    {
      const uint64_t mf = first1 ? m1 : m0, ms = first1 ? m0 : m1;
      const int32_t rf = first1 ? r1 : r0, rs = first1 ? r0 : r1;
      next = (mf != 0 && rf >= 0) ? rf : ((ms != 0 && rs >= 0) ? rs : -1);
    }
    int32_t lf0 = 0, lc0 = 0, lf1 = 0, lc1 = 0;
    if (m0 != 0 && r0 < 0) { lf0 = ~r0; lc0 = n0; }
    if (m1 != 0 && r1 < 0) {
      if (lc0 == 0) { lf0 = ~r1; lc0 = n1; }
      else { lf1 = ~r1; lc1 = n1; }
    }
    const int32_t total = lc0 + lc1;  // <= 2 x 16 (leaf_size <= 16)
    double px = 0, py = 0, pz = 0;
    int32_t pid = 0;
    if (lane < total) {
      const PointRecD r = pr[lane < lc0 ? lf0 + lane : lf1 + (lane - lc0)];
      px = r.x;
      py = r.y;
      pz = r.z;
      pid = r.id;
    }
    BNode nd_next;
    if (next >= 0) nd_next = nodes[next];
    bool entered = false;
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
      const int c = first1 ? 1 - cc : cc;
      const uint64_t mc = c == 0 ? m0 : m1;
      const int32_t r = c == 0 ? r0 : r1;
      if (mc == 0 || r < 0) continue;
      if (!entered && r == next) {
        entered = true;
      } else if (sp < WAVE) {  // never full: depth < 64 is checked on the host
        if (lane == sp) {
          s_ref = r;
          if (c == 0) { s_lx = a0; s_ly = a1; s_lz = a2; s_hx = a3; s_hy = b0; s_hz = b1; }
          else { s_lx = b2; s_ly = b3; s_lz = c0; s_hx = c1; s_hy = c2; s_hz = c3; }
        }
        ++sp;
      }
    }
    for (int32_t j = 0; j < total; ++j) {
      const double x = lreadd(px, j), y = lreadd(py, j), z = lreadd(pz, j);
      const int32_t id = lread(pid, j);
      const double dx = x - qx, dy = y - qy, dz = z - qz;
      const double d = (dx * dx + dy * dy) + dz * dz;
      if (better64(d, id, bd, bi)) {
        bd = d;
        bi = id;
        bslot = j < lc0 ? lf0 + j : lf1 + (j - lc0);
      }
    }
    if (next >= 0) { node = next; nd = nd_next; continue; }
    bool found = false;
    const float bfn = float_up(bd);
    while (sp > 0) {
      --sp;
      const float lx = lreadf(s_lx, sp), ly = lreadf(s_ly, sp), lz = lreadf(s_lz, sp);
      const float hx = lreadf(s_hx, sp), hy = lreadf(s_hy, sp), hz = lreadf(s_hz, sp);
      if (__ballot(box_lb(qlo, qhi, lx, ly, lz, hx, hy, hz) <= bfn)) {
        node = lread(s_ref, sp);
        found = true;
        break;
      }
    }
    if (!found) break;
    nd = nodes[uni64(node)];
  }
  if (!inrange) return;
  const bool hit = active && bi != 0x7fffffff;
  out_idx[qi] = hit ? bi : -1;
  if (out_d2) out_d2[qi] = hit ? bd : -1.0;
  if (out_dist) out_dist[qi] = hit ? sqrt(bd) : -1.0;
  if (out_cp) {
    double cx = 0, cy = 0, cz = 0;
    if (hit) {
      const PointRecD r = pr[bslot];
      cx = r.x;
      cy = r.y;
      cz = r.z;
    }
    out_cp[3 * qi] = cx;
    out_cp[3 * qi + 1] = cy;
    out_cp[3 * qi + 2] = cz;
  }
}

__global__ void k_fill_missing64(int64_t n, int32_t *idx, double *cp, double *dist, double *d2) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  idx[i] = -1;
  if (dist) dist[i] = -1.0;
  if (d2) d2[i] = -1.0;
  if (cp) { cp[3 * i] = 0; cp[3 * i + 1] = 0; cp[3 * i + 2] = 0; }
}

}  // namespace

void build_points64(Bvh &b, const double *xyz, int64_t n, int leaf_size, hipStream_t s) {
  require(leaf_size >= 1 && leaf_size <= 16, "leaf_size must be in 1..16");
  DevBuf rounded;
  if (n > 0) {
    rounded.alloc(n * 3 * sizeof(float));
    hipLaunchKernelGGL(k_round_points, dim3(ceil_div(n * 3, 256)), dim3(256), 0, s, xyz, n * 3,
                       rounded.as<float>());
    launch_check("k_round_points");
  }
  build_bvh(b, rounded.as<float>(), nullptr, n, true, leaf_size, s);
  b.f64 = true;
  if (n > 0) {
    b.elems64.alloc(n * sizeof(PointRecD));
    hipLaunchKernelGGL(k_point_records64, dim3(ceil_div(n, 256)), dim3(256), 0, s, xyz, b.elems.as<PointRec>(), n,
                       b.elems64.as<PointRecD>());
    launch_check("k_point_records64");
  }
  THG_HIP(hipStreamSynchronize(s));
}

void order_queries64(const Bvh &b, const double *q, int64_t n, QueryScratch &ws, hipStream_t s) {
  ws.order = nullptr;
  if (n < ORDER_MIN_QUERIES) return;
  const QueryScratch::Sort m = ws.sort_memory(n, s);
  const float3 ext = f3(b.hi.x - b.lo.x, b.hi.y - b.lo.y, b.hi.z - b.lo.z);
  const float3 inv =
      f3(ext.x > 0 ? 1.0f / ext.x : 0.0f, ext.y > 0 ? 1.0f / ext.y : 0.0f, ext.z > 0 ? 1.0f / ext.z : 0.0f);
  hipLaunchKernelGGL(k_query_keys64, dim3(ceil_div(n, 256)), dim3(256), 0, s, q, n, b.lo, inv, m.keys, m.vals);
  launch_check("k_query_keys64");
  radix_sort_pairs(m.keys, m.vals, m.tk, m.tv, m.hist, n, 6, 30, s);
  ws.order = m.vals;
}

double bound_to_best64(double dist_bound) {
  if (dist_bound < 0) return INFINITY;
  return std::nextafter(dist_bound * dist_bound, -INFINITY);  // d^2 < bound^2 (BvhN.hpp:446-466)
}

void launch_closest64(const Bvh &b, const double *q, int64_t n, double dist_bound, int32_t *idx, double *cp,
                      double *dist, double *d2, QueryScratch &ws, hipStream_t s) {
  if (n <= 0) return;
  require(b.f64 && b.points, "not a double-precision point BVH");
  if (b.n == 0) {
    hipLaunchKernelGGL(k_fill_missing64, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, idx, cp, dist, d2);
    launch_check("k_fill_missing64");
    return;
  }
  require(b.max_depth < WAVE, "point BVH deeper than the 64-entry wave stack");
  order_queries64(b, q, n, ws, s);
  const uint32_t nb = ceil_div(n, B64);
  KernelTimer kt(ws.timer_ctx(b.ctx), s, "closest64");
  hipLaunchKernelGGL(k_closest64, dim3(round_up8(nb)), dim3(B64), 0, s, b.nodes.as<BNode>(), b.elems64.as<PointRecD>(),
                     q, ws.order, n, nb, bound_to_best64(dist_bound), b.scale, idx, cp, dist, d2);
  launch_check("k_closest64");
  kt.stop();
}

}  // namespace thg

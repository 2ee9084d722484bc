// 4-wide quantized ray tree (QNode4) folded from a built binary BVH.
//
// Every binary node at even depth (the root included) becomes a QNode4 whose children are its
// grandchildren through inner children and its leaf children directly, so a ray visits about half
// as many nodes and reads 16 bytes per child box instead of 32. Child boxes are stored as bytes on
// a per-node, per-axis power-of-two grid and rounded outward, so each decoded box contains the
// exact one and culling stays conservative. Leaves keep their RayTri ranges.
#include "kernels_common.hpp"
#include "query.hpp"

namespace thg {

constexpr int QB_BLOCK = 256;

__global__ void __launch_bounds__(QB_BLOCK) k_q_parents(const BNode *__restrict__ nodes, int64_t nn,
                                                        int32_t *__restrict__ parent) {
  const int64_t i = (int64_t)blockIdx.x * QB_BLOCK + threadIdx.x;
  if (i >= nn) return;
  if (i == 0) parent[0] = -1;
  const int4 d = nodes[i].d;
  if (d.x >= 0) parent[d.x] = (int32_t)i;
  if (d.y >= 0) parent[d.y] = (int32_t)i;
}

// flag[i] = 1 for binary nodes at even depth (walking the parents up to the root).
__global__ void __launch_bounds__(QB_BLOCK) k_q_flags(const int32_t *__restrict__ parent, int64_t nn,
                                                      uint32_t *__restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * QB_BLOCK + threadIdx.x;
  if (i >= nn) return;
  int depth = 0;
  for (int32_t p = parent[i]; p >= 0; p = parent[p]) ++depth;
  flag[i] = (depth & 1) == 0 ? 1u : 0u;
}

// Smallest power-of-two exponent e (within the normal range) with fmaf(255, 2^e, lo) >= hi.
__device__ int q_exponent(float lo, float hi) {
  const double E = (double)hi - (double)lo;
  int e;
  if (!(E > 0.0)) {
    e = -126;
  } else if (!(E < 1e38)) {
    e = 127;
  } else {
    int x;
    frexp(E / 255.0, &x);  // E / 255 = m 2^x, m in [0.5, 1)
    e = x < -126 ? -126 : (x > 127 ? 127 : x);
  }
  while (e < 127 && !(fmaf(255.0f, ldexpf(1.0f, e), lo) >= hi)) ++e;
  return e;
}

// Byte offsets of [lo, hi] on the grid (origin, 2^e), rounded outward against the decode fmaf.
__device__ void q_round(float lo, float hi, float origin, int e, uint32_t &qlo, uint32_t &qhi) {
  const float s = ldexpf(1.0f, e);
  int a = 0, b = 255;
  if (lo == lo) {
    const double t = floor(((double)lo - (double)origin) / (double)s);
    a = t < 0.0 ? 0 : (t > 255.0 ? 255 : (int)t);
    while (a > 0 && fmaf((float)a, s, origin) > lo) --a;
  }
  if (hi == hi) {
    const double t = ceil(((double)hi - (double)origin) / (double)s);
    b = t < 0.0 ? 0 : (t > 255.0 ? 255 : (int)t);
    while (b < 255 && fmaf((float)b, s, origin) < hi) ++b;
  }
  qlo = (uint32_t)a;
  qhi = (uint32_t)b;
}

__global__ void __launch_bounds__(QB_BLOCK) k_q_emit(const BNode *__restrict__ nodes, int64_t nn,
                                                     const uint32_t *__restrict__ flag,
                                                     const uint32_t *__restrict__ qidx, QNode4 *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * QB_BLOCK + threadIdx.x;
  if (i >= nn || !flag[i]) return;
  float lo[4][3], hi[4][3];
  int32_t ref[4] = {INT32_MIN, INT32_MIN, INT32_MIN, INT32_MIN};
  uint32_t cnt[4] = {0, 0, 0, 0};
  int nc = 0;
  const BNode nd = nodes[i];
  auto add = [&](float lx, float ly, float lz, float hx, float hy, float hz, int32_t r, uint32_t c) {
    lo[nc][0] = lx; lo[nc][1] = ly; lo[nc][2] = lz;
    hi[nc][0] = hx; hi[nc][1] = hy; hi[nc][2] = hz;
    ref[nc] = r;
    cnt[nc] = c;
    ++nc;
  };
  // Child k of a binary node: box, ref, count.
  auto child = [](const BNode &n, int k, float *l, float *h, int32_t &r, int32_t &c) {
    if (k == 0) {
      l[0] = n.a.x; l[1] = n.a.y; l[2] = n.a.z; h[0] = n.a.w; h[1] = n.b.x; h[2] = n.b.y;
      r = n.d.x; c = n.d.z;
    } else {
      l[0] = n.b.z; l[1] = n.b.w; l[2] = n.c.x; h[0] = n.c.y; h[1] = n.c.z; h[2] = n.c.w;
      r = n.d.y; c = n.d.w;
    }
  };
  for (int k = 0; k < 2; ++k) {
    float l[3], h[3];
    int32_t r, c;
    child(nd, k, l, h, r, c);
    if (r >= 0) {
      const BNode g = nodes[r];
      for (int j = 0; j < 2; ++j) {
        float gl[3], gh[3];
        int32_t gr, gc;
        child(g, j, gl, gh, gr, gc);
        if (gr >= 0) add(gl[0], gl[1], gl[2], gh[0], gh[1], gh[2], (int32_t)qidx[gr], 0);
        else if (gc > 0) add(gl[0], gl[1], gl[2], gh[0], gh[1], gh[2], gr, (uint32_t)gc);
      }
    } else if (c > 0) {
      add(l[0], l[1], l[2], h[0], h[1], h[2], r, (uint32_t)c);
    }
  }
  float nlo[3] = {INFINITY, INFINITY, INFINITY}, nhi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int k = 0; k < nc; ++k)
    for (int a = 0; a < 3; ++a) {
      nlo[a] = fminf(nlo[a], lo[k][a]);
      nhi[a] = fmaxf(nhi[a], hi[k][a]);
    }
  if (nc == 0)
    for (int a = 0; a < 3; ++a) nlo[a] = nhi[a] = 0.0f;
  int e[3];
  uint32_t ql[3] = {0, 0, 0}, qh[3] = {0, 0, 0};
  for (int a = 0; a < 3; ++a) {
    e[a] = q_exponent(nlo[a], nhi[a]);
    for (int k = 0; k < nc; ++k) {
      uint32_t x, y;
      q_round(lo[k][a], hi[k][a], nlo[a], e[a], x, y);
      ql[a] |= x << (8 * k);
      qh[a] |= y << (8 * k);
    }
  }
  QNode4 q;
  const uint32_t meta = (uint32_t)(e[0] + 127) | (uint32_t)(e[1] + 127) << 8 | (uint32_t)(e[2] + 127) << 16 |
                        (uint32_t)nc << 24;
  q.a = make_float4(nlo[0], nlo[1], nlo[2], __uint_as_float(meta));
  q.b = make_uint4(ql[0], ql[1], ql[2], qh[0]);
  q.c = make_int4((int)qh[1], (int)qh[2], ref[0], ref[1]);
  q.d = make_int4(ref[2], ref[3], (int)(cnt[0] | cnt[1] << 16), (int)(cnt[2] | cnt[3] << 16));
  out[qidx[i]] = q;
}

void build_qbvh4(Bvh &b, hipStream_t s) {
  const int64_t nn = b.num_nodes;
  if (nn <= 0) return;
  // Leaf counts are kept in 16 bits (leaves hold at most leaf_size elements).
  require(b.leaf_size <= 65535, "leaf size too large for the ray tree");
  DevBuf parent(nn * 4), flag(nn * 4), qidx((nn + 1) * 4), tmp(scan_tmp_words(nn) * 4);
  const uint32_t grid = ceil_div(nn, QB_BLOCK);
  hipLaunchKernelGGL(k_q_parents, dim3(grid), dim3(QB_BLOCK), 0, s, b.nodes.as<BNode>(), nn, parent.as<int32_t>());
  launch_check("k_q_parents");
  hipLaunchKernelGGL(k_q_flags, dim3(grid), dim3(QB_BLOCK), 0, s, parent.as<int32_t>(), nn, flag.as<uint32_t>());
  launch_check("k_q_flags");
  scan_excl_u32(flag.as<uint32_t>(), qidx.as<uint32_t>(), nn, tmp.as<uint32_t>(), s);
  uint32_t nq = 0;
  THG_HIP(hipMemcpyAsync(&nq, qidx.as<uint32_t>() + nn, 4, hipMemcpyDeviceToHost, s));
  THG_HIP(hipStreamSynchronize(s));
  b.qnodes.alloc((size_t)nq * sizeof(QNode4));
  hipLaunchKernelGGL(k_q_emit, dim3(grid), dim3(QB_BLOCK), 0, s, b.nodes.as<BNode>(), nn, flag.as<uint32_t>(),
                     qidx.as<uint32_t>(), b.qnodes.as<QNode4>());
  launch_check("k_q_emit");
  THG_HIP(hipStreamSynchronize(s));
  b.num_qnodes = nq;
  // Binary internal levels d = 0 .. max_depth - 1 fold into levels 0, 2, 4, ...
  b.q_depth = b.max_depth / 2 + 1;
}

}  // namespace thg

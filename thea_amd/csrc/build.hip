// LBVH construction on the GPU (replaces Thea's kd-tree build, Algorithms/BvhN.hpp:1366-1525).
//
//  1. k_elem_prep     element boxes and centres (triangle: Thea's centroid formula (v0+v1+v2)/3,
//                     Triangle3.hpp:148; point: the point), per-block bounds reduced by one more
//                     small kernel (ordered-int min/max).
//  2. k_morton        30-bit Morton code of each centre in the centre bounds.
//  3. radix sort      (code, element) pairs, stable, so equal codes keep element order.
//  4. k_karras        binary radix tree over the sorted codes (Karras, "Maximizing Parallelism in
//                     the Construction of BVHs, Octrees, and k-d Trees", HPG 2012), duplicate codes
//                     disambiguated by position.
//  5. k_refit         bottom-up boxes; the second child to arrive at a node (atomic counter) merges.
//  6. k_collapse/emit subtrees of <= leaf_size elements become leaves; remaining nodes are
//                     compacted into 64-byte BNodes holding both children's boxes.
//  7. k_records       element records in leaf order (TriRec with Thea's per-triangle constants).
// Boxes are exact float min/max of element coordinates; the query kernels add their own rounding
// slack, so no padding is applied here.
#include "kernels_common.hpp"
#include "query.hpp"
#include "tri_math.hpp"

namespace thg {

constexpr int B_BLOCK = 256;

struct BoundsAcc {  // ordered-int encoded: elem lo xyz, elem hi xyz, centre lo xyz, centre hi xyz
  int v[12];
};

__device__ __forceinline__ float3 ld_vert(const float *__restrict__ verts, int64_t i) {
  return f3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
}

__global__ void __launch_bounds__(B_BLOCK) k_elem_prep(const float *__restrict__ verts, const int32_t *__restrict__ tris,
                                                       int64_t n, int points, float4 *__restrict__ center,
                                                       float4 *__restrict__ blo, float4 *__restrict__ bhi,
                                                       int *__restrict__ part) {
  float mn[6] = {INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY};
  float mx[6] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = (int64_t)blockIdx.x * B_BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * B_BLOCK) {
    float3 lo, hi, c;
    if (points) {
      c = ld_vert(verts, i);
      lo = c;
      hi = c;
    } else {
      float3 v0 = ld_vert(verts, tris[3 * i]), v1 = ld_vert(verts, tris[3 * i + 1]), v2 = ld_vert(verts, tris[3 * i + 2]);
      lo = f3(fminf(v0.x, fminf(v1.x, v2.x)), fminf(v0.y, fminf(v1.y, v2.y)), fminf(v0.z, fminf(v1.z, v2.z)));
      hi = f3(fmaxf(v0.x, fmaxf(v1.x, v2.x)), fmaxf(v0.y, fmaxf(v1.y, v2.y)), fmaxf(v0.z, fmaxf(v1.z, v2.z)));
      float3 s = add3(add3(v0, v1), v2);
      c = f3(s.x / 3.0f, s.y / 3.0f, s.z / 3.0f);
    }
    center[i] = make_float4(c.x, c.y, c.z, 0.0f);
    blo[i] = make_float4(lo.x, lo.y, lo.z, 0.0f);
    bhi[i] = make_float4(hi.x, hi.y, hi.z, 0.0f);
    mn[0] = fminf(mn[0], lo.x); mn[1] = fminf(mn[1], lo.y); mn[2] = fminf(mn[2], lo.z);
    mx[0] = fmaxf(mx[0], hi.x); mx[1] = fmaxf(mx[1], hi.y); mx[2] = fmaxf(mx[2], hi.z);
    mn[3] = fminf(mn[3], c.x); mn[4] = fminf(mn[4], c.y); mn[5] = fminf(mn[5], c.z);
    mx[3] = fmaxf(mx[3], c.x); mx[4] = fmaxf(mx[4], c.y); mx[5] = fmaxf(mx[5], c.z);
  }
  // Block partials in BoundsAcc's slot order (no same-address atomics: they serialise in L2).
  __shared__ float red[B_BLOCK / WAVE][12];
  const int w = threadIdx.x / WAVE;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const float a = wave_min(mn[k]), b = wave_max(mx[k]);
    if (lane_id() == 0) {
      red[w][k < 3 ? k : 6 + (k - 3)] = a;
      red[w][k < 3 ? 3 + k : 9 + (k - 3)] = b;
    }
  }
  __syncthreads();
  if (threadIdx.x < 12) {
    const int k = threadIdx.x;
    const bool lo_slot = k < 3 || (k >= 6 && k < 9);
    float v = red[0][k];
    for (int ww = 1; ww < B_BLOCK / WAVE; ++ww) v = lo_slot ? fminf(v, red[ww][k]) : fmaxf(v, red[ww][k]);
    part[(int64_t)blockIdx.x * 12 + k] = float_to_ord(v);
  }
}

// One wave per slot: the min (lo slots) or max (hi slots) of the block partials.
__global__ void __launch_bounds__(12 * WAVE) k_bounds_reduce(const int *__restrict__ part, int nb,
                                                             BoundsAcc *__restrict__ acc) {
  const int k = threadIdx.x / WAVE, lane = lane_id();
  const bool lo_slot = k < 3 || (k >= 6 && k < 9);
  int v = lo_slot ? float_to_ord(INFINITY) : float_to_ord(-INFINITY);
  for (int b = lane; b < nb; b += WAVE) {
    const int x = part[(int64_t)b * 12 + k];
    v = lo_slot ? min(v, x) : max(v, x);
  }
  for (int off = 32; off >= 1; off >>= 1) {
    const int o = __shfl_xor(v, off, WAVE);
    v = lo_slot ? min(v, o) : max(v, o);
  }
  if (lane == 0) acc->v[k] = v;
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3FF;
  v = (v | (v << 16)) & 0x030000FF;
  v = (v | (v << 8)) & 0x0300F00F;
  v = (v | (v << 4)) & 0x030C30C3;
  v = (v | (v << 2)) & 0x09249249;
  return v;
}

__device__ __forceinline__ uint32_t morton30(float3 c, float3 lo, float3 inv) {
  float x = (c.x - lo.x) * inv.x, y = (c.y - lo.y) * inv.y, z = (c.z - lo.z) * inv.z;
  uint32_t ix = (uint32_t)fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
  uint32_t iy = (uint32_t)fminf(fmaxf(y * 1024.0f, 0.0f), 1023.0f);
  uint32_t iz = (uint32_t)fminf(fmaxf(z * 1024.0f, 0.0f), 1023.0f);
  return (spread10(ix) << 2) | (spread10(iy) << 1) | spread10(iz);
}

__global__ void __launch_bounds__(B_BLOCK) k_morton(const float4 *__restrict__ center, int64_t n, float3 lo, float3 inv,
                                                    uint32_t *__restrict__ keys, uint32_t *__restrict__ vals) {
  int64_t i = (int64_t)blockIdx.x * B_BLOCK + threadIdx.x;
  if (i >= n) return;
  float4 c = center[i];
  keys[i] = morton30(f3(c.x, c.y, c.z), lo, inv);
  vals[i] = (uint32_t)i;
}

// Common prefix length of sorted (code, position) keys i and j; -1 outside [0, n).
__device__ __forceinline__ int delta(const uint32_t *__restrict__ k, int64_t n, int64_t i, int64_t j) {
  if (j < 0 || j >= n) return -1;
  uint32_t a = k[i], b = k[j];
  if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
  return __clz(a ^ b);
}

// Internal node i of n - 1: children (leaf if flag bit 31 set), range, parent links.
__global__ void __launch_bounds__(B_BLOCK) k_karras(const uint32_t *__restrict__ k, int64_t n, int32_t *__restrict__ left,
                                                    int32_t *__restrict__ right, int32_t *__restrict__ rfirst,
                                                    int32_t *__restrict__ rlast, int32_t *__restrict__ node_parent,
                                                    int32_t *__restrict__ leaf_parent) {
  int64_t i = (int64_t)blockIdx.x * B_BLOCK + threadIdx.x;
  if (i >= n - 1) return;
  int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = delta(k, n, i, i - d);
  int64_t lmax = 2;
  while (delta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
  int64_t l = 0;
  for (int64_t t = lmax / 2; t >= 1; t /= 2)
    if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
  int64_t j = i + l * d;
  int dnode = delta(k, n, i, j);
  int64_t s = 0;
  int64_t div = 2;
  int64_t t;
  do {
    t = (l + div - 1) / div;
    if (t >= 1 && delta(k, n, i, i + (s + t) * d) > dnode) s += t;
    div *= 2;
  } while (t > 1);
  int64_t g = i + s * d + (d < 0 ? -1 : 0);
  int64_t first = i < j ? i : j, last = i < j ? j : i;
  int32_t lc, rc;
  if (first == g) { lc = (int32_t)(g | 0x80000000u); leaf_parent[g] = (int32_t)i; }
  else { lc = (int32_t)g; node_parent[g] = (int32_t)i; }
  if (last == g + 1) { rc = (int32_t)((g + 1) | 0x80000000u); leaf_parent[g + 1] = (int32_t)i; }
  else { rc = (int32_t)(g + 1); node_parent[g + 1] = (int32_t)i; }
  left[i] = lc;
  right[i] = rc;
  rfirst[i] = (int32_t)first;
  rlast[i] = (int32_t)last;
  if (i == 0) node_parent[0] = -1;
}

__device__ __forceinline__ float ld_coherent(const float *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coherent(float *p, float v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bottom-up refit over the Karras tree. Node boxes: nbox[6 i + 0..5] = lo xyz, hi xyz.
// `sorted` maps sorted position -> element; element boxes come from blo/bhi.
__global__ void __launch_bounds__(B_BLOCK) k_refit(int64_t n, const uint32_t *__restrict__ sorted,
                                                   const float4 *__restrict__ blo, const float4 *__restrict__ bhi,
                                                   const int32_t *__restrict__ left, const int32_t *__restrict__ right,
                                                   const int32_t *__restrict__ node_parent,
                                                   const int32_t *__restrict__ leaf_parent, uint32_t *__restrict__ flags,
                                                   float *nbox) {
  int64_t k = (int64_t)blockIdx.x * B_BLOCK + threadIdx.x;
  if (k >= n) return;
  int32_t cur = leaf_parent[k];
  while (cur >= 0) {
    __threadfence();
    uint32_t old = atomicAdd(&flags[cur], 1u);
    if (old == 0) return;  // first child to arrive; the sibling's thread continues
    __threadfence();
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    int32_t ch[2] = {left[cur], right[cur]};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      if (ch[c] < 0) {
        uint32_t e = sorted[ch[c] & 0x7fffffff];
        float4 a = blo[e], b = bhi[e];
        lo[0] = fminf(lo[0], a.x); lo[1] = fminf(lo[1], a.y); lo[2] = fminf(lo[2], a.z);
        hi[0] = fmaxf(hi[0], b.x); hi[1] = fmaxf(hi[1], b.y); hi[2] = fmaxf(hi[2], b.z);
      } else {
        const float *q = nbox + 6 * (int64_t)ch[c];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          lo[a] = fminf(lo[a], ld_coherent(q + a));
          hi[a] = fmaxf(hi[a], ld_coherent(q + 3 + a));
        }
      }
    }
    float *o = nbox + 6 * (int64_t)cur;
#pragma unroll
    for (int a = 0; a < 3; ++a) { st_coherent(o + a, lo[a]); st_coherent(o + 3 + a, hi[a]); }
    cur = node_parent[cur];
  }
}

// Two-pass refit. Pass 1 (k_refit_span): a block owns REFIT_SPAN consecutive sorted positions and
// computes every internal node whose element range lies inside them. A node's split position g
// (left child = [first, g], right child = [g + 1, last]) is unique, so LDS slot g - span start
// holds the node's arrival counter and one box per side; the second child to arrive merges, with
// workgroup-scope ordering only. Node boxes go to global memory as they are computed. A child whose
// parent straddles the span appends the parent to the span's own list, entries[span start ...]
// with an LDS counter, and the block stores the list's length in nentries[span]. (One device-wide
// counter for all spans serialised the pass: a few hundred thousand appends on one address.) The
// maximal subtrees inside a span have disjoint ranges in it, so a list never outgrows the span.
// Pass 2 (k_refit_top) finishes the straddling nodes with the device-wide arrival protocol of
// k_refit; there are O(depth) of them per span boundary, so the costly device-scope fences run for
// a small fraction of the nodes.
constexpr int REFIT_SPAN = 1024;
constexpr int REFIT_THREADS = 256;
constexpr int REFIT_TOP_LANES = 16;

__device__ __forceinline__ void elem_box(const uint32_t *__restrict__ sorted, const float4 *__restrict__ blo,
                                         const float4 *__restrict__ bhi, int64_t pos, float b[6]) {
  const uint32_t e = sorted[pos];
  const float4 a = blo[e], c = bhi[e];
  b[0] = a.x; b[1] = a.y; b[2] = a.z; b[3] = c.x; b[4] = c.y; b[5] = c.z;
}

__global__ void __launch_bounds__(REFIT_THREADS) k_refit_span(int64_t n, const uint32_t *__restrict__ sorted,
                                                              const float4 *__restrict__ blo,
                                                              const float4 *__restrict__ bhi,
                                                              const int32_t *__restrict__ rfirst,
                                                              const int32_t *__restrict__ rlast,
                                                              const int32_t *__restrict__ node_parent,
                                                              const int32_t *__restrict__ leaf_parent, float *nbox,
                                                              int32_t *__restrict__ entries,
                                                              uint32_t *__restrict__ nentries) {
  __shared__ float s_l[REFIT_SPAN * 6], s_r[REFIT_SPAN * 6];
  __shared__ uint32_t s_flag[REFIT_SPAN];
  __shared__ uint32_t s_nent;
  const int64_t lo = (int64_t)blockIdx.x * REFIT_SPAN;
  const int64_t hi = lo + REFIT_SPAN < n ? lo + REFIT_SPAN : n;
  for (int i = threadIdx.x; i < REFIT_SPAN; i += REFIT_THREADS) s_flag[i] = 0;
  if (threadIdx.x == 0) s_nent = 0;
  __syncthreads();
  for (int64_t k = lo + threadIdx.x; k < hi; k += REFIT_THREADS) {
    float b[6];
    elem_box(sorted, blo, bhi, k, b);
    int64_t cf = k, cl = k;  // element range of the completed child
    int32_t p = leaf_parent[k];
    while (p >= 0) {
      const int64_t fp = rfirst[p], lp = rlast[p];
      if (fp < lo || lp >= hi) {
        entries[lo + atomicAdd(&s_nent, 1u)] = p;
        break;
      }
      const bool is_left = cf == fp;
      const int slot = (int)((is_left ? cl : cf - 1) - lo);
      float *mine = (is_left ? s_l : s_r) + 6 * slot;
#pragma unroll
      for (int a = 0; a < 6; ++a) mine[a] = b[a];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      if (atomicAdd(&s_flag[slot], 1u) == 0) break;  // the sibling's thread continues
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      const float *other = (is_left ? s_r : s_l) + 6 * slot;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        b[a] = fminf(b[a], other[a]);
        b[3 + a] = fmaxf(b[3 + a], other[3 + a]);
      }
      float *o = nbox + 6 * (int64_t)p;
#pragma unroll
      for (int a = 0; a < 6; ++a) o[a] = b[a];
      cf = fp;
      cl = lp;
      p = node_parent[p];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) nentries[blockIdx.x] = s_nent;
}

// Box of internal node `cur` from its two children: element boxes, or node boxes written before
// (read at L2 with agent-scope atomics, so a concurrent writer's box is never a stale L1 copy).
__device__ __forceinline__ void refit_node(int32_t cur, const uint32_t *__restrict__ sorted,
                                           const float4 *__restrict__ blo, const float4 *__restrict__ bhi,
                                           const int32_t *__restrict__ left, const int32_t *__restrict__ right,
                                           float *nbox) {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  const int32_t ch[2] = {left[cur], right[cur]};
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    float b[6];
    if (ch[c] < 0) {
      elem_box(sorted, blo, bhi, ch[c] & 0x7fffffff, b);
    } else {
      const float *q = nbox + 6 * (int64_t)ch[c];
#pragma unroll
      for (int a = 0; a < 6; ++a) b[a] = ld_coherent(q + a);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], b[a]);
      hi[a] = fmaxf(hi[a], b[3 + a]);
    }
  }
  float *o = nbox + 6 * (int64_t)cur;
#pragma unroll
  for (int a = 0; a < 3; ++a) { st_coherent(o + a, lo[a]); st_coherent(o + 3 + a, hi[a]); }
}

// Fenced walk from `cur` to the root with the device-wide arrival protocol of k_refit.
__device__ __forceinline__ void refit_walk(int32_t cur, const uint32_t *__restrict__ sorted,
                                           const float4 *__restrict__ blo, const float4 *__restrict__ bhi,
                                           const int32_t *__restrict__ left, const int32_t *__restrict__ right,
                                           const int32_t *__restrict__ node_parent, uint32_t *flags, float *nbox) {
  while (cur >= 0) {
    __threadfence();
    if (atomicAdd(&flags[cur], 1u) == 0) break;  // first arrival; the other child's thread continues
    __threadfence();
    refit_node(cur, sorted, blo, bhi, left, right, nbox);
    cur = node_parent[cur];
  }
}

// Pass 2 for small trees: REFIT_TOP_LANES threads per span list walk up with fences.
__global__ void __launch_bounds__(B_BLOCK) k_refit_top(const uint32_t *__restrict__ sorted,
                                                       const float4 *__restrict__ blo, const float4 *__restrict__ bhi,
                                                       const int32_t *__restrict__ left,
                                                       const int32_t *__restrict__ right,
                                                       const int32_t *__restrict__ node_parent,
                                                       const int32_t *__restrict__ entries,
                                                       const uint32_t *__restrict__ nentries, uint32_t spans,
                                                       uint32_t *flags, float *nbox) {
  const uint64_t t = (uint64_t)blockIdx.x * B_BLOCK + threadIdx.x;
  const uint32_t sp = (uint32_t)(t / REFIT_TOP_LANES);
  if (sp >= spans) return;
  const uint32_t count = nentries[sp];
  for (uint32_t e = (uint32_t)(t % REFIT_TOP_LANES); e < count; e += REFIT_TOP_LANES)
    refit_walk(entries[(int64_t)sp * REFIT_SPAN + e], sorted, blo, bhi, left, right, node_parent, flags, nbox);
}

// Pass 2 for large trees, level by level. A device-scope fence writes back the issuing XCD's L2 on
// this GPU, and the fenced walk pays two per node: 5-7 ms on the 512-mesh batched build, where
// the top tree above 41.7k spans has some 10^5 nodes. Instead, k_refit_arrive counts one arrival
// per span-list entry, and a node's second arrival queues it as ready (both children's boxes are
// written); each k_refit_level launch then writes the boxes of the ready nodes and counts one
// arrival at each parent, queueing the parents that became ready for the next launch. Kernel
// boundaries order every box write before its reads, so no fence is needed. After REFIT_LEVELS
// launches (more than a tree's usual height), the last one finishes any remaining nodes with the
// fenced walk, so the result never depends on the height.
constexpr int REFIT_LEVELS = 64;
constexpr uint32_t REFIT_LEVEL_MIN_SPANS = 4096;
constexpr uint32_t REFIT_LEVEL_GRID = 512;

__global__ void k_zero_words(uint32_t *p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

__global__ void __launch_bounds__(B_BLOCK) k_refit_arrive(const int32_t *__restrict__ entries,
                                                          const uint32_t *__restrict__ nentries, uint32_t spans,
                                                          uint32_t *flags, int32_t *__restrict__ out,
                                                          uint32_t *out_cnt) {
  const uint64_t t = (uint64_t)blockIdx.x * B_BLOCK + threadIdx.x;
  const uint32_t sp = (uint32_t)(t / REFIT_TOP_LANES);
  if (sp >= spans) return;
  const uint32_t count = nentries[sp];
  for (uint32_t e = (uint32_t)(t % REFIT_TOP_LANES); e < count; e += REFIT_TOP_LANES) {
    const int32_t p = entries[(int64_t)sp * REFIT_SPAN + e];
    if (atomicAdd(&flags[p], 1u) == 1u) out[wave_append(out_cnt)] = p;
  }
}

__global__ void __launch_bounds__(B_BLOCK) k_refit_level(const uint32_t *__restrict__ sorted,
                                                         const float4 *__restrict__ blo,
                                                         const float4 *__restrict__ bhi,
                                                         const int32_t *__restrict__ left,
                                                         const int32_t *__restrict__ right,
                                                         const int32_t *__restrict__ node_parent,
                                                         const int32_t *__restrict__ in,
                                                         const uint32_t *__restrict__ in_cnt,
                                                         int32_t *__restrict__ out, uint32_t *out_cnt,
                                                         uint32_t *flags, float *nbox, int last) {
  const uint32_t count = *in_cnt;
  for (uint32_t i = blockIdx.x * B_BLOCK + threadIdx.x; i < count; i += gridDim.x * B_BLOCK) {
    const int32_t p = in[i];
    refit_node(p, sorted, blo, bhi, left, right, nbox);
    const int32_t q = node_parent[p];
    if (last) {
      refit_walk(q, sorted, blo, bhi, left, right, node_parent, flags, nbox);
    } else if (q >= 0 && atomicAdd(&flags[q], 1u) == 1u) {
      out[wave_append(out_cnt)] = q;
    }
  }
}

// keep[i] = 1 if internal node i stays an internal node of the collapsed tree (range > leaf_size).
__global__ void __launch_bounds__(B_BLOCK) k_keep(int64_t n, const int32_t *__restrict__ rfirst,
                                                  const int32_t *__restrict__ rlast, int leaf_size,
                                                  uint32_t *__restrict__ keep) {
  int64_t i = (int64_t)blockIdx.x * B_BLOCK + threadIdx.x;
  if (i >= n - 1) return;
  keep[i] = (rlast[i] - rfirst[i] + 1) > leaf_size ? 1u : 0u;
}

__device__ __forceinline__ void child_box(int32_t c, const uint32_t *sorted, const float4 *blo, const float4 *bhi,
                                          const float *nbox, float lo[3], float hi[3]) {
  if (c < 0) {
    uint32_t e = sorted[c & 0x7fffffff];
    float4 a = blo[e], b = bhi[e];
    lo[0] = a.x; lo[1] = a.y; lo[2] = a.z; hi[0] = b.x; hi[1] = b.y; hi[2] = b.z;
  } else {
    const float *q = nbox + 6 * (int64_t)c;
    lo[0] = q[0]; lo[1] = q[1]; lo[2] = q[2]; hi[0] = q[3]; hi[1] = q[4]; hi[2] = q[5];
  }
}

// Emit the collapsed tree. `fidx` = exclusive prefix of keep (final node index of kept nodes, plus
// `base`); a child that is not kept becomes a leaf over its whole element range. `leaf_count`, when
// given, receives the number of leaves (a diagnostic: a full binary tree with `kept` internal nodes
// has kept + 1 leaves, which the builders use instead, since one counter that every wave adds to
// serialises the launch: 7.6 ms of the 512-mesh batched build).
__global__ void __launch_bounds__(B_BLOCK) k_emit(int64_t n, int32_t base, const uint32_t *__restrict__ keep,
                                                  const uint32_t *__restrict__ fidx, const int32_t *__restrict__ left,
                                                  const int32_t *__restrict__ right, const int32_t *__restrict__ rfirst,
                                                  const int32_t *__restrict__ rlast, const uint32_t *__restrict__ sorted,
                                                  const float4 *__restrict__ blo, const float4 *__restrict__ bhi,
                                                  const float *__restrict__ nbox, BNode *__restrict__ out,
                                                  uint32_t *__restrict__ leaf_count) {
  int64_t i = (int64_t)blockIdx.x * B_BLOCK + threadIdx.x;
  uint32_t leaves = 0;
  if (i < n - 1 && keep[i]) {
    int32_t ch[2] = {left[i], right[i]};
    float lo[2][3], hi[2][3];
    int32_t ref[2], cnt[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      child_box(ch[c], sorted, blo, bhi, nbox, lo[c], hi[c]);
      if (ch[c] < 0) {
        ref[c] = ~(ch[c] & 0x7fffffff);
        cnt[c] = 1;
        ++leaves;
      } else if (!keep[ch[c]]) {
        ref[c] = ~rfirst[ch[c]];
        cnt[c] = rlast[ch[c]] - rfirst[ch[c]] + 1;
        ++leaves;
      } else {
        ref[c] = base + (int32_t)fidx[ch[c]];
        cnt[c] = 0;
      }
    }
    BNode nd;
    nd.a = make_float4(lo[0][0], lo[0][1], lo[0][2], hi[0][0]);
    nd.b = make_float4(hi[0][1], hi[0][2], lo[1][0], lo[1][1]);
    nd.c = make_float4(lo[1][2], hi[1][0], hi[1][1], hi[1][2]);
    nd.d = make_int4(ref[0], ref[1], cnt[0], cnt[1]);
    out[base + fidx[i]] = nd;
  }
  if (leaf_count) {
    // one atomic per wave on the shared counter (every lane of the wave takes part in the sum)
    leaves = wave_sum(leaves);
    if (lane_id() == 0 && leaves) atomicAdd(leaf_count, leaves);
  }
}

// Maximum leaf depth of the collapsed tree (number of kept ancestors), via atomicMax. The elements
// of one collapsed leaf share their kept ancestors, so only the leaf's first element walks to the
// root; the others stop after the non-kept nodes above them (all inside that leaf). Ancestors of a
// kept node are kept (their ranges are larger).
__global__ void __launch_bounds__(B_BLOCK) k_depth(int64_t n, const int32_t *__restrict__ leaf_parent,
                                                   const int32_t *__restrict__ node_parent,
                                                   const uint32_t *__restrict__ keep,
                                                   const int32_t *__restrict__ rfirst, int32_t *__restrict__ max_depth) {
  int64_t k = (int64_t)blockIdx.x * B_BLOCK + threadIdx.x;
  int d = 0;
  if (k < n) {
    int32_t cur = leaf_parent[k], top = -1;
    while (cur >= 0 && !keep[cur]) {
      top = cur;
      cur = node_parent[cur];
    }
    if (top < 0 || rfirst[top] == k)
      for (; cur >= 0; cur = node_parent[cur]) ++d;
  }
  int m = (int)wave_max((float)d);  // every lane of the wave takes part in the reduction
  // one atomic per wave, and none once the maximum is known to be at least this wave's
  if (lane_id() == 0 && m > __hip_atomic_load(max_depth, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    atomicMax(max_depth, m);
}

__global__ void __launch_bounds__(B_BLOCK) k_tri_records(const float *__restrict__ verts, const int32_t *__restrict__ tris,
                                                         const uint32_t *__restrict__ sorted, int64_t n,
                                                         TriRec *__restrict__ out) {
  int64_t k = (int64_t)blockIdx.x * B_BLOCK + threadIdx.x;
  if (k >= n) return;
  uint32_t e = sorted[k];
  float3 v0 = ld_vert(verts, tris[3 * (int64_t)e]), v1 = ld_vert(verts, tris[3 * (int64_t)e + 1]),
         v2 = ld_vert(verts, tris[3 * (int64_t)e + 2]);
  out[k] = make_trirec(v0, v1, v2, (int32_t)e);
}

__global__ void __launch_bounds__(B_BLOCK) k_point_records(const float *__restrict__ xyz, const uint32_t *__restrict__ sorted,
                                                           int64_t n, PointRec *__restrict__ out) {
  int64_t k = (int64_t)blockIdx.x * B_BLOCK + threadIdx.x;
  if (k >= n) return;
  uint32_t e = sorted[k];
  PointRec r;
  r.p = make_float4(xyz[3 * (int64_t)e], xyz[3 * (int64_t)e + 1], xyz[3 * (int64_t)e + 2], __int_as_float((int)e));
  out[k] = r;
}

// Screening records from the triangle records, in the same (leaf) order.
__global__ void __launch_bounds__(B_BLOCK) k_screen_records(const TriRec *__restrict__ tr, int64_t n,
                                                            ScreenRec *__restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * B_BLOCK + threadIdx.x;
  if (k >= n) return;
  ScreenRec r;
  r.c = tr[k].sb;
  r.n = make_float4(tr[k].p0.w, tr[k].p1.w, tr[k].p2.w, 0.0f);
  out[k] = r;
}

void launch_screen_records(const TriRec *tr, int64_t n, ScreenRec *out, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_screen_records, dim3(ceil_div(n, B_BLOCK)), dim3(B_BLOCK), 0, s, tr, n, out);
  launch_check("k_screen_records");
}

// Ray records from the triangle records, in the same (leaf) order.
__global__ void __launch_bounds__(B_BLOCK) k_ray_records(const TriRec *__restrict__ tr, int64_t n,
                                                         RayTri *__restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * B_BLOCK + threadIdx.x;
  if (k >= n) return;
  const TriRec &r = tr[k];
  const float3 v0 = f3(r.p0.x, r.p0.y, r.p0.z);
  const float3 e01 = sub3(f3(r.p1.x, r.p1.y, r.p1.z), v0), e02 = sub3(f3(r.p2.x, r.p2.y, r.p2.z), v0);
  RayTri o;
  o.a = make_float4(v0.x, v0.y, v0.z, r.pl.w);
  o.b = make_float4(e01.x, e01.y, e01.z, 0.0f);
  o.c = make_float4(e02.x, e02.y, e02.z, 0.0f);
  out[k] = o;
}

// Flags the last ray record of every non-empty leaf (RayTri::b.w int bits = 1), so ray traversal can
// run a leaf from its first record without its count.
__global__ void __launch_bounds__(B_BLOCK) k_leaf_ends(const BNode *__restrict__ nodes, int64_t num_nodes,
                                                       RayTri *__restrict__ rays) {
  const int64_t i = (int64_t)blockIdx.x * B_BLOCK + threadIdx.x;
  if (i >= num_nodes) return;
  const int4 d = nodes[i].d;
  if (d.x < 0 && d.z > 0) rays[(int64_t)~d.x + d.z - 1].b.w = __int_as_float(1);
  if (d.y < 0 && d.w > 0) rays[(int64_t)~d.y + d.w - 1].b.w = __int_as_float(1);
}

// Leaf bounds of the mesh BVH's leaves (one thread per node; nodes [0, nsmall + *kept), kept may
// be null for 0): the normal is the normalised sum of the leaf's triangle normals and the centre the
// leaf box's centre, both rounded to float; the half-thickness t and the radius r are then the
// largest vertex offsets from them, in double and rounded up, so the leaf's triangles (convex hulls
// of their vertices) lie inside the stored slab and cylinder whatever the normal is.
__global__ void __launch_bounds__(B_BLOCK) k_leaf_bounds(const BNode *__restrict__ nodes, int64_t nsmall,
                                                         const uint32_t *__restrict__ kept, int64_t max_nodes,
                                                         const TriRec *__restrict__ tr, LeafBound *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * B_BLOCK + threadIdx.x;
  const int64_t count = nsmall + (kept ? (int64_t)*kept : 0);
  if (i >= count || i >= max_nodes) return;
  const BNode nd = nodes[i];
  const float lo6[2][3] = {{nd.a.x, nd.a.y, nd.a.z}, {nd.b.z, nd.b.w, nd.c.x}};
  const float hi6[2][3] = {{nd.a.w, nd.b.x, nd.b.y}, {nd.c.y, nd.c.z, nd.c.w}};
  const int32_t ref[2] = {nd.d.x, nd.d.y}, cnt[2] = {nd.d.z, nd.d.w};
  for (int c = 0; c < 2; ++c) {
    if (ref[c] >= 0 || cnt[c] <= 0) continue;
    const int64_t f = ~(int64_t)ref[c];
    double sx = 0, sy = 0, sz = 0;
    for (int64_t e = f; e < f + cnt[c]; ++e) {
      const float4 p0 = tr[e].p0, p1 = tr[e].p1, p2 = tr[e].p2;
      if (isfinite(p0.w) && isfinite(p1.w) && isfinite(p2.w)) { sx += p0.w; sy += p1.w; sz += p2.w; }
    }
    double nn = sqrt(sx * sx + sy * sy + sz * sz);
    if (!(nn > 1e-30)) { sx = 0; sy = 0; sz = 1; nn = 1; }
    const float nx = (float)(sx / nn), ny = (float)(sy / nn), nz = (float)(sz / nn);
    const float cx = (float)(0.5 * ((double)lo6[c][0] + (double)hi6[c][0]));
    const float cy = (float)(0.5 * ((double)lo6[c][1] + (double)hi6[c][1]));
    const float cz = (float)(0.5 * ((double)lo6[c][2] + (double)hi6[c][2]));
    const double len = sqrt((double)nx * nx + (double)ny * ny + (double)nz * nz);
    double t = 0, r2 = 0;
    for (int64_t e = f; e < f + cnt[c]; ++e) {
      const float4 pv[3] = {tr[e].p0, tr[e].p1, tr[e].p2};
      for (int k = 0; k < 3; ++k) {
        const double dx = (double)pv[k].x - cx, dy = (double)pv[k].y - cy, dz = (double)pv[k].z - cz;
        const double s = nx * dx + ny * dy + nz * dz;
        t = fmax(t, fabs(s));
        const double sn = s / len;
        r2 = fmax(r2, dx * dx + dy * dy + dz * dz - sn * sn);
      }
    }
    LeafBound lb;
    lb.a = make_float4(cx, cy, cz, (float)(sqrt(fmax(r2, 0.0)) * (1.0 + 1e-6)));
    lb.b = make_float4(nx, ny, nz, (float)(t * (1.0 + 1e-6)));
    out[f] = lb;
  }
}

void launch_leaf_bounds(const BNode *nodes, int64_t nsmall, const uint32_t *kept, int64_t max_nodes, const TriRec *tr,
                        LeafBound *out, hipStream_t s) {
  if (max_nodes <= 0) return;
  hipLaunchKernelGGL(k_leaf_bounds, dim3(ceil_div(max_nodes, B_BLOCK)), dim3(B_BLOCK), 0, s, nodes, nsmall, kept,
                     max_nodes, tr, out);
  launch_check("k_leaf_bounds");
}

// Single-leaf tree for n <= leaf_size: child 0 holds every element, child 1 is an empty leaf.
__global__ void k_single_leaf(int64_t n, float3 lo, float3 hi, BNode *out) {
  BNode nd;
  nd.a = make_float4(lo.x, lo.y, lo.z, hi.x);
  nd.b = make_float4(hi.y, hi.z, INFINITY, INFINITY);
  nd.c = make_float4(INFINITY, -INFINITY, -INFINITY, -INFINITY);
  nd.d = make_int4(~0, ~0, (int)n, 0);
  out[0] = nd;
}

// Exclusive scan of n uint32 into out (total in out[n]); uses the radix scan kernel's approach.
// (in and out may be the same array)
__global__ void __launch_bounds__(1024) k_excl_scan_u32(const uint32_t *in, uint32_t *out, int64_t n) {
  __shared__ uint32_t part[1024 / WAVE];
  __shared__ uint32_t carry;
  int t = threadIdx.x;
  if (t == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < n; base += 1024) {
    int64_t i = base + t;
    uint32_t v = i < n ? in[i] : 0;
    uint32_t inc = wave_incl_scan(v);
    int w = t / WAVE, l = t % WAVE;
    if (l == WAVE - 1) part[w] = inc;
    __syncthreads();
    if (w == 0) {
      uint32_t pv = l < 1024 / WAVE ? part[l] : 0;
      uint32_t pinc = wave_incl_scan(pv);
      if (l < 1024 / WAVE) part[l] = pinc - pv;
    }
    __syncthreads();
    uint32_t excl = carry + part[w] + inc - v;
    if (i < n) out[i] = excl;
    __syncthreads();
    if (t == 1023) carry = excl + v;
    __syncthreads();
  }
  if (t == 0) out[n] = carry;
}

// Three-phase exclusive scan of n uint32 (out[n] = total): block sums, scan of the sums, block
// rescan. `sums` must hold ceil(n / 1024) + 1 words.
__global__ void __launch_bounds__(1024) k_scan_block_sums(const uint32_t *__restrict__ in, int64_t n,
                                                          uint32_t *__restrict__ sums) {
  __shared__ uint32_t part[1024 / WAVE];
  int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
  uint32_t v = i < n ? in[i] : 0;
  uint32_t s = wave_sum(v);
  if (lane_id() == 0) part[threadIdx.x / WAVE] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < 1024 / WAVE; ++w) t += part[w];
    sums[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(1024) k_scan_block_apply(const uint32_t *__restrict__ in, int64_t n,
                                                           const uint32_t *__restrict__ offs, uint32_t *__restrict__ out) {
  __shared__ uint32_t part[1024 / WAVE];
  int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
  uint32_t v = i < n ? in[i] : 0;
  uint32_t inc = wave_incl_scan(v);
  int w = threadIdx.x / WAVE;
  if (lane_id() == WAVE - 1) part[w] = inc;
  __syncthreads();
  uint32_t off = offs[blockIdx.x];
  for (int k = 0; k < w; ++k) off += part[k];
  if (i < n) out[i] = off + inc - v;
  if (i == n - 1) out[n] = off + inc;
}

size_t scan_tmp_words(int64_t n) { return (size_t)ceil_div((uint64_t)std::max<int64_t>(n, 1), 1024) + 1; }

void scan_excl_u32(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *tmp, hipStream_t s) {
  if (n <= 0) return;
  uint32_t nb = ceil_div((uint64_t)n, 1024);
  hipLaunchKernelGGL(k_scan_block_sums, dim3(nb), dim3(1024), 0, s, in, n, tmp);
  launch_check("k_scan_block_sums");
  hipLaunchKernelGGL(k_excl_scan_u32, dim3(1), dim3(1024), 0, s, tmp, tmp, (int64_t)nb);
  launch_check("k_excl_scan_u32");
  hipLaunchKernelGGL(k_scan_block_apply, dim3(nb), dim3(1024), 0, s, in, n, tmp, out);
  launch_check("k_scan_block_apply");
}

// Entry points for the batched many-mesh pipeline (batch.hip), which reuses these kernels over
// concatenated meshes with global element positions and node indices.
static bool refit_atomic_only() {
  static const bool v = [] {
    const char *e = getenv("THG_REFIT");
    return e && atoi(e) == 1;
  }();
  return v;
}

size_t refit_count_words(int64_t n) {
  return (size_t)std::max<int64_t>(ceil_div(n, (int64_t)REFIT_SPAN), 1) + REFIT_LEVELS + 2;
}
bool refit_needs_list(int64_t n) { return ceil_div(n, (int64_t)REFIT_SPAN) >= REFIT_LEVEL_MIN_SPANS; }

// `flags` (one per internal node) must be zero; `entries` holds n int32, `nentries`
// refit_count_words(n) uint32, and `list` n int32 when refit_needs_list(n) (else it may be null).
void launch_refit(int64_t n, const uint32_t *sorted, const float4 *blo, const float4 *bhi, const int32_t *left,
                  const int32_t *right, const int32_t *rfirst, const int32_t *rlast, const int32_t *nparent,
                  const int32_t *lparent, uint32_t *flags, float *nbox, int32_t *entries, uint32_t *nentries,
                  int32_t *list, hipStream_t s) {
  if (n <= 0) return;
  if (refit_atomic_only()) {
    hipLaunchKernelGGL(k_refit, dim3(ceil_div(n, B_BLOCK)), dim3(B_BLOCK), 0, s, n, sorted, blo, bhi, left, right,
                       nparent, lparent, flags, nbox);
    launch_check("k_refit");
    return;
  }
  const uint32_t spans = ceil_div(n, (int64_t)REFIT_SPAN);
  hipLaunchKernelGGL(k_refit_span, dim3(spans), dim3(REFIT_THREADS), 0, s, n, sorted, blo, bhi, rfirst, rlast, nparent,
                     lparent, nbox, entries, nentries);
  launch_check("k_refit_span");
  const uint32_t grid = (uint32_t)ceil_div((int64_t)spans * REFIT_TOP_LANES, (int64_t)B_BLOCK);
  if (!refit_needs_list(n)) {
    hipLaunchKernelGGL(k_refit_top, dim3(grid), dim3(B_BLOCK), 0, s, sorted, blo, bhi, left, right, nparent, entries,
                       nentries, spans, flags, nbox);
    launch_check("k_refit_top");
    return;
  }
  require(list != nullptr, "refit: level list missing");
  uint32_t *cnt = nentries + spans;  // REFIT_LEVELS + 1 list lengths
  hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(128), 0, s, cnt, REFIT_LEVELS + 1);
  launch_check("k_zero_words");
  hipLaunchKernelGGL(k_refit_arrive, dim3(grid), dim3(B_BLOCK), 0, s, entries, nentries, spans, flags, list, cnt);
  launch_check("k_refit_arrive");
  // level L reads list L (even: `list`, odd: `entries`, whose span lists are consumed) and writes L + 1
  for (int L = 0; L < REFIT_LEVELS; ++L) {
    int32_t *in = (L & 1) ? entries : list, *out = (L & 1) ? list : entries;
    hipLaunchKernelGGL(k_refit_level, dim3(REFIT_LEVEL_GRID), dim3(B_BLOCK), 0, s, sorted, blo, bhi, left, right,
                       nparent, in, cnt + L, out, cnt + L + 1, flags, nbox, L == REFIT_LEVELS - 1 ? 1 : 0);
    launch_check("k_refit_level");
  }
}

void launch_emit(int64_t ni, int32_t base, const uint32_t *keep, const uint32_t *fidx, const int32_t *left,
                 const int32_t *right, const int32_t *rfirst, const int32_t *rlast, const uint32_t *sorted,
                 const float4 *blo, const float4 *bhi, const float *nbox, BNode *out, uint32_t *leaf_count,
                 hipStream_t s) {
  if (ni <= 0) return;
  hipLaunchKernelGGL(k_emit, dim3(ceil_div(ni, B_BLOCK)), dim3(B_BLOCK), 0, s, ni + 1, base, keep, fidx, left, right,
                     rfirst, rlast, sorted, blo, bhi, nbox, out, leaf_count);
  launch_check("k_emit");
}

void launch_depth(int64_t n, const int32_t *lparent, const int32_t *nparent, const uint32_t *keep, const int32_t *rfirst,
                  int32_t *max_depth, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_depth, dim3(ceil_div(n, B_BLOCK)), dim3(B_BLOCK), 0, s, n, lparent, nparent, keep, rfirst,
                     max_depth);
  launch_check("k_depth");
}

// Host driver shared by mesh and point builds.
void build_bvh(Bvh &b, const float *verts, const int32_t *tris, int64_t n, bool points, int leaf_size, hipStream_t s) {
  b.points = points;
  b.n = n;
  b.leaf_size = leaf_size;
  if (n <= 0) {
    b.num_nodes = 0;
    b.num_leaves = 0;
    b.max_depth = 0;
    return;
  }
  uint32_t grid = std::min<uint32_t>(ceil_div(n, B_BLOCK), 4096);
  DevBuf center(n * sizeof(float4)), blo(n * sizeof(float4)), bhi(n * sizeof(float4)), acc(sizeof(BoundsAcc)),
      part((size_t)grid * 12 * sizeof(int));
  hipLaunchKernelGGL(k_elem_prep, dim3(grid), dim3(B_BLOCK), 0, s, verts, tris, n, points ? 1 : 0,
                     center.as<float4>(), blo.as<float4>(), bhi.as<float4>(), part.as<int>());
  launch_check("k_elem_prep");
  hipLaunchKernelGGL(k_bounds_reduce, dim3(1), dim3(12 * WAVE), 0, s, part.as<int>(), (int)grid, acc.as<BoundsAcc>());
  launch_check("k_bounds_reduce");
  BoundsAcc h;
  THG_HIP(hipMemcpyAsync(&h, acc.p, sizeof(h), hipMemcpyDeviceToHost, s));
  THG_HIP(hipStreamSynchronize(s));
  float f[12];
  for (int k = 0; k < 12; ++k) f[k] = ord_to_float(h.v[k]);
  b.lo = f3(f[0], f[1], f[2]);
  b.hi = f3(f[3], f[4], f[5]);
  b.scale = 0;
  for (int k = 0; k < 6; ++k) b.scale = std::max(b.scale, std::fabs(f[k]));
  require(std::isfinite(b.scale), "element coordinates must be finite");

  size_t rec_bytes = points ? sizeof(PointRec) : sizeof(TriRec);
  b.elems.alloc(n * rec_bytes);

  DevBuf keys(n * 4), vals(n * 4), tk(n * 4), tv(n * 4), hist(radix_hist_words(n) * 4);
  float3 clo = f3(f[6], f[7], f[8]);
  float3 cext = f3(f[9] - f[6], f[10] - f[7], f[11] - f[8]);
  float3 inv = f3(cext.x > 0 ? 1.0f / cext.x : 0.0f, cext.y > 0 ? 1.0f / cext.y : 0.0f, cext.z > 0 ? 1.0f / cext.z : 0.0f);
  hipLaunchKernelGGL(k_morton, dim3(ceil_div(n, B_BLOCK)), dim3(B_BLOCK), 0, s, center.as<float4>(), n, clo, inv,
                     keys.as<uint32_t>(), vals.as<uint32_t>());
  launch_check("k_morton");
  radix_sort_pairs(keys.as<uint32_t>(), vals.as<uint32_t>(), tk.as<uint32_t>(), tv.as<uint32_t>(), hist.as<uint32_t>(), n,
                   0, 32, s);
  const uint32_t *sorted = vals.as<uint32_t>();

  if (points)
    hipLaunchKernelGGL(k_point_records, dim3(ceil_div(n, B_BLOCK)), dim3(B_BLOCK), 0, s, verts, sorted, n,
                       b.elems.as<PointRec>());
  else
    hipLaunchKernelGGL(k_tri_records, dim3(ceil_div(n, B_BLOCK)), dim3(B_BLOCK), 0, s, verts, tris, sorted, n,
                       b.elems.as<TriRec>());
  launch_check("k_records");
  if (!points) {
    b.rays.alloc(n * sizeof(RayTri));
    hipLaunchKernelGGL(k_ray_records, dim3(ceil_div(n, B_BLOCK)), dim3(B_BLOCK), 0, s, b.elems.as<TriRec>(), n,
                       b.rays.as<RayTri>());
    launch_check("k_ray_records");
    b.screen.alloc(n * sizeof(ScreenRec));
    launch_screen_records(b.elems.as<TriRec>(), n, b.screen.as<ScreenRec>(), s);
  }

  if (n <= leaf_size) {
    b.nodes.alloc(sizeof(BNode));
    hipLaunchKernelGGL(k_single_leaf, dim3(1), dim3(1), 0, s, n, b.lo, b.hi, b.nodes.as<BNode>());
    launch_check("k_single_leaf");
    if (!points) {
      hipLaunchKernelGGL(k_leaf_ends, dim3(1), dim3(B_BLOCK), 0, s, b.nodes.as<BNode>(), (int64_t)1,
                         b.rays.as<RayTri>());
      launch_check("k_leaf_ends");
      b.lbounds.alloc(n * sizeof(LeafBound));
      launch_leaf_bounds(b.nodes.as<BNode>(), 1, nullptr, 1, b.elems.as<TriRec>(), b.lbounds.as<LeafBound>(), s);
    }
    b.num_nodes = 1;
    b.num_leaves = 1;
    b.max_depth = 1;
    if (!points) build_qbvh4(b, s);
    THG_HIP(hipStreamSynchronize(s));
    return;
  }

  int64_t ni = n - 1;
  DevBuf left(ni * 4), right(ni * 4), rfirst(ni * 4), rlast(ni * 4), nparent(ni * 4), lparent(n * 4);
  DevBuf flags(ni * 4), nbox(ni * 6 * sizeof(float)), keep(ni * 4), fidx((ni + 1) * 4), counters(16);
  DevBuf entries(n * 4), nentries(refit_count_words(n) * 4), rlist(refit_needs_list(n) ? n * 4 : 4);
  THG_HIP(hipMemsetAsync(flags.p, 0, ni * 4, s));
  THG_HIP(hipMemsetAsync(counters.p, 0, 16, s));
  hipLaunchKernelGGL(k_karras, dim3(ceil_div(ni, B_BLOCK)), dim3(B_BLOCK), 0, s, keys.as<uint32_t>(), n,
                     left.as<int32_t>(), right.as<int32_t>(), rfirst.as<int32_t>(), rlast.as<int32_t>(),
                     nparent.as<int32_t>(), lparent.as<int32_t>());
  launch_check("k_karras");
  launch_refit(n, sorted, blo.as<float4>(), bhi.as<float4>(), left.as<int32_t>(), right.as<int32_t>(),
               rfirst.as<int32_t>(), rlast.as<int32_t>(), nparent.as<int32_t>(), lparent.as<int32_t>(),
               flags.as<uint32_t>(), nbox.as<float>(), entries.as<int32_t>(), nentries.as<uint32_t>(), rlist.as<int32_t>(), s);
  hipLaunchKernelGGL(k_keep, dim3(ceil_div(ni, B_BLOCK)), dim3(B_BLOCK), 0, s, n, rfirst.as<int32_t>(),
                     rlast.as<int32_t>(), leaf_size, keep.as<uint32_t>());
  launch_check("k_keep");
  DevBuf stmp(scan_tmp_words(ni) * 4);
  scan_excl_u32(keep.as<uint32_t>(), fidx.as<uint32_t>(), ni, stmp.as<uint32_t>(), s);
  uint32_t kept = 0;
  THG_HIP(hipMemcpyAsync(&kept, fidx.as<uint32_t>() + ni, 4, hipMemcpyDeviceToHost, s));
  THG_HIP(hipStreamSynchronize(s));
  b.nodes.alloc((size_t)kept * sizeof(BNode));
  int32_t *maxd = reinterpret_cast<int32_t *>(counters.as<uint32_t>() + 1);
  hipLaunchKernelGGL(k_emit, dim3(ceil_div(ni, B_BLOCK)), dim3(B_BLOCK), 0, s, n, 0, keep.as<uint32_t>(),
                     fidx.as<uint32_t>(), left.as<int32_t>(), right.as<int32_t>(), rfirst.as<int32_t>(),
                     rlast.as<int32_t>(), sorted, blo.as<float4>(), bhi.as<float4>(), nbox.as<float>(),
                     b.nodes.as<BNode>(), nullptr);
  launch_check("k_emit");
  hipLaunchKernelGGL(k_depth, dim3(ceil_div(n, B_BLOCK)), dim3(B_BLOCK), 0, s, n, lparent.as<int32_t>(),
                     nparent.as<int32_t>(), keep.as<uint32_t>(), rfirst.as<int32_t>(), maxd);
  launch_check("k_depth");
  uint32_t cnt[2];
  THG_HIP(hipMemcpyAsync(cnt, counters.p, 8, hipMemcpyDeviceToHost, s));
  THG_HIP(hipStreamSynchronize(s));
  b.num_nodes = kept;
  b.num_leaves = (int64_t)kept + 1;  // a full binary tree
  b.max_depth = (int32_t)cnt[1];
  if (!points) {
    hipLaunchKernelGGL(k_leaf_ends, dim3(ceil_div(kept, B_BLOCK)), dim3(B_BLOCK), 0, s, b.nodes.as<BNode>(),
                       (int64_t)kept, b.rays.as<RayTri>());
    launch_check("k_leaf_ends");
    b.lbounds.alloc(n * sizeof(LeafBound));
    launch_leaf_bounds(b.nodes.as<BNode>(), (int64_t)kept, nullptr, (int64_t)kept, b.elems.as<TriRec>(),
                       b.lbounds.as<LeafBound>(), s);
    build_qbvh4(b, s);
  }
}

}  // namespace thg

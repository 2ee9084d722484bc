// Query kernels: closest element / closest point, k nearest, ray first hit and any hit.
//
// One query per lane; each lane walks the BVH depth first with its own stack in LDS (node ref and
// the lower bound it was pushed with), visiting the nearer child first. Queries are first put in
// spatial (Morton) order so the 64 lanes of a wave walk nearby paths; blocks are remapped so each
// XCD's L2 serves one contiguous spatial range. Results go to the caller's query index.
//
// Results are defined canonically: the lexicographic minimum of (float d^2, element index) for
// closest queries (and the k smallest such pairs for k-NN), (float t, element index) for rays,
// where d^2 and t come from Thea's own float arithmetic (tri_math.hpp). A node is skipped only when
// its lower bound is strictly above the current best, and the lower bound treats the query as a
// box of half-width 2^-16 max(|q|, scene scale), so rounding in the element arithmetic can never
// make a box prune an element that would have won.
#include "kernels_common.hpp"
#include "query.hpp"
#include "traverse.hpp"
#include "tri_math.hpp"

namespace thg {

constexpr int Q_BLOCK = 64;
// Relative margin of the element reject tests: a bound is compared with sqrt(best) * REJ_REL.
constexpr float REJ_REL = 1.0f + 1.0f / 4096.0f;

// Squared threshold of the triangle reject (tri_lb2) for a lane whose current best (or k-th) squared
// distance is `best`: (sqrt(best) REJ_REL + 2 sl)^2; +inf while the lane has no bound.
__device__ __forceinline__ float rej_sq(float best, float sl) {
  const float t = fmaf(2.0f, sl, sqrtf(best) * REJ_REL);
  return t * t;
}

__device__ __forceinline__ float3 load_q(const float *__restrict__ q, int64_t i) {
  return f3(q[3 * i], q[3 * i + 1], q[3 * i + 2]);
}

__device__ __forceinline__ int64_t query_slot(uint32_t nb_real, int64_t n, const uint32_t *__restrict__ order,
                                              bool &active) {
  uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
  int64_t t = (int64_t)lb * Q_BLOCK + threadIdx.x;
  active = lb < nb_real && t < n;
  if (!active) return 0;
  return order ? (int64_t)order[t] : t;
}

//--------------------------------------------------------------------------------------------------
// Closest element
//--------------------------------------------------------------------------------------------------
// One query per lane, stack in LDS (entry i of lane t at s_ref[i * Q_BLOCK + t]).
// Per-lane traversal stack of node references (and, with LB, their lower bounds): the top REG
// entries live in registers as a shift register ([0] on top), deeper ones in LDS at
// (i - REG) * Q_BLOCK + tid. Registers and dropping the bounds both shrink a block's LDS, so more
// waves fit on a CU; without bounds a popped node is entered and its children are tested instead.
constexpr int LANE_REG = 16;

template <int STACK, int REG> constexpr int lane_lds_entries() { return STACK > REG ? STACK - REG : 1; }

template <int REG, bool LB> struct LaneStack {
  int32_t r[REG > 0 ? REG : 1];
  float l[REG > 0 && LB ? REG : 1];
  // Push onto a stack currently holding sp entries.
  __device__ __forceinline__ void push(int sp, int32_t ref, float lb, int32_t *__restrict__ s_ref,
                                       float *__restrict__ s_lb, int tid) {
    if (REG == 0) {
      s_ref[sp * Q_BLOCK + tid] = ref;
      if (LB) s_lb[sp * Q_BLOCK + tid] = lb;
      return;
    }
    if (sp >= REG) {
      s_ref[(sp - REG) * Q_BLOCK + tid] = r[REG - 1];
      if (LB) s_lb[(sp - REG) * Q_BLOCK + tid] = l[REG - 1];
    }
#pragma unroll
    for (int j = REG - 1; j > 0; --j) {
      r[j] = r[j - 1];
      if (LB) l[j] = l[j - 1];
    }
    r[0] = ref;
    if (LB) l[0] = lb;
  }
  // Pop the top; sp is the count after the pop. Without LB the bound returned is -inf (enter).
  __device__ __forceinline__ void pop(int sp, int32_t &ref, float &lb, const int32_t *__restrict__ s_ref,
                                      const float *__restrict__ s_lb, int tid) {
    if (REG == 0) {
      ref = s_ref[sp * Q_BLOCK + tid];
      lb = LB ? s_lb[sp * Q_BLOCK + tid] : -INFINITY;
      return;
    }
    ref = r[0];
    lb = LB ? l[0] : -INFINITY;
#pragma unroll
    for (int j = 0; j < REG - 1; ++j) {
      r[j] = r[j + 1];
      if (LB) l[j] = l[j + 1];
    }
    if (sp >= REG) {
      r[REG - 1] = s_ref[(sp - REG) * Q_BLOCK + tid];
      if (LB) l[REG - 1] = s_lb[(sp - REG) * Q_BLOCK + tid];
    }
  }
};

// Stack configurations: closest-point lanes keep the bounds (a popped node is pruned without a
// load); k-NN lanes, whose k-list already fills registers, keep references only in LDS.
constexpr int CL_REG = 16;
constexpr bool CL_LB = true;
constexpr int KN_REG = 0;
constexpr bool KN_LB = false;

template <int STACK, bool POINTS>
__device__ __forceinline__ void closest_lane_body(int32_t *__restrict__ s_ref, float *__restrict__ s_lb,
                                                  const BNode *__restrict__ nodes, const void *__restrict__ elems,
                                                  const float *__restrict__ q, int64_t qi, float best_init,
                                                  float scale, int exclude_self, int32_t *__restrict__ out_idx,
                                                  float *__restrict__ out_cp, float *__restrict__ out_dist,
                                                  float *__restrict__ out_d2, int32_t root) {
  const int tid = threadIdx.x;
  float3 p = load_q(q, qi);
  float m = fmaxf(scale, fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
  float sl = m * SLACK_REL;
  float3 qlo = f3(p.x - sl, p.y - sl, p.z - sl), qhi = f3(p.x + sl, p.y + sl, p.z + sl);
  float best = best_init;
  int32_t bidx = 0x7fffffff;
  float3 bcp = f3(0, 0, 0);
  const int32_t self = exclude_self ? (int32_t)qi : -1;
  int sp = 0;
  LaneStack<CL_REG, CL_LB> stk;
  int32_t node = root;
  const TriRec *__restrict__ tr = (const TriRec *)elems;
  const PointRec *__restrict__ pr = (const PointRec *)elems;
  while (true) {
    const BNode nd = nodes[node];
    float lb0 = box_lb(qlo, qhi, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y);
    float lb1 = box_lb(qlo, qhi, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w);
    int32_t r0 = nd.d.x, r1 = nd.d.y, c0 = nd.d.z, c1 = nd.d.w;
    if (lb1 < lb0) {
      float tl = lb0; lb0 = lb1; lb1 = tl;
      int32_t tr_ = r0; r0 = r1; r1 = tr_;
      int32_t tc = c0; c0 = c1; c1 = tc;
    }
    int32_t next = -1;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      float lb = c == 0 ? lb0 : lb1;
      int32_t r = c == 0 ? r0 : r1;
      int32_t cnt = c == 0 ? c0 : c1;
      if (!(lb <= best)) continue;
      if (r < 0) {
        int32_t f = ~r;
        for (int32_t k = f; k < f + cnt; ++k) {
          float d;
          float3 cp;
          int32_t id;
          if (POINTS) {
            float4 e = pr[k].p;
            id = __float_as_int(e.w);
            cp = f3(e.x, e.y, e.z);
            d = sqn3(sub3(cp, p));
          } else {
            const TriRec *t = tr + k;
            id = trirec_id(t->pl);
            d = tri_closest(t->p0, t->p1, t->p2, t->pl, t->e0, t->e1, t->e2, p, cp);
          }
          if (id != self && better(d, id, best, bidx)) {
            best = d;
            bidx = id;
            bcp = cp;
          }
        }
      } else if (next < 0) {
        next = r;
      } else if (sp < STACK) {  // never taken when depth < STACK (checked by every caller)
        stk.push(sp, r, lb, s_ref, s_lb, tid);
        ++sp;
      }
    }
    if (next >= 0) { node = next; continue; }
    bool found = false;
    while (sp > 0) {
      --sp;
      int32_t pr_ref;
      float pr_lb;
      stk.pop(sp, pr_ref, pr_lb, s_ref, s_lb, tid);
      if (pr_lb <= best) { node = pr_ref; found = true; break; }
    }
    if (!found) break;
  }
  bool hit = bidx != 0x7fffffff;
  out_idx[qi] = hit ? bidx : -1;
  if (out_cp) {
    out_cp[3 * qi] = hit ? bcp.x : 0.0f;
    out_cp[3 * qi + 1] = hit ? bcp.y : 0.0f;
    out_cp[3 * qi + 2] = hit ? bcp.z : 0.0f;
  }
  if (out_dist) out_dist[qi] = hit ? sqrtf(best) : -1.0f;
  if (out_d2) out_d2[qi] = hit ? best : -1.0f;
}

template <int STACK, bool POINTS>
__global__ void __launch_bounds__(Q_BLOCK) k_closest(const BNode *__restrict__ nodes, const void *__restrict__ elems,
                                                     const float *__restrict__ q, const uint32_t *__restrict__ order,
                                                     int64_t n, uint32_t nb_real, float best_init, float scale,
                                                     int exclude_self, int32_t *__restrict__ out_idx,
                                                     float *__restrict__ out_cp, float *__restrict__ out_dist,
                                                     float *__restrict__ out_d2) {
  __shared__ int32_t s_ref[lane_lds_entries<STACK, CL_REG>() * Q_BLOCK];
  __shared__ float s_lb[CL_LB ? lane_lds_entries<STACK, CL_REG>() * Q_BLOCK : 1];
  bool active;
  const int64_t qi = query_slot(nb_real, n, order, active);
  if (!active) return;
  closest_lane_body<STACK, POINTS>(s_ref, s_lb, nodes, elems, q, qi, best_init, scale, exclude_self, out_idx, out_cp,
                                   out_dist, out_d2, 0);
}

//--------------------------------------------------------------------------------------------------
// Closest element, wave-cooperative ("packet") traversal
//
// The 64 lanes of a wave hold 64 spatially sorted queries and walk the tree together: a child is
// entered if any lane's lower bound admits it, node and triangle loads are wave-uniform (scalar
// loads), and every lane evaluates every admitted triangle, so there is no divergence in the walk.
// The uniform traversal stack is spread over the wave's lanes: entry i (child ref and its box)
// lives in lane i's registers; a push is one predicated register write, a pop is v_readlane.
// Results are identical to the per-lane kernel: evaluating extra triangles never changes the
// canonical minimum, and a subtree is skipped only if its bound exceeds every lane's best.
//--------------------------------------------------------------------------------------------------
constexpr int W_BLOCK = 256;
constexpr int W_WAVES = W_BLOCK / WAVE;
constexpr int W_STACK = WAVE;  // stack entries, one per lane

__device__ __forceinline__ int32_t uni(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float unif(float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); }
__device__ __forceinline__ int32_t lane_read(int32_t x, int32_t l) { return __builtin_amdgcn_readlane(x, l); }
__device__ __forceinline__ float lane_readf(float x, int32_t l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// G lanes share one query (64 / G queries per wave): lane g of a group tests elements g, g + G, ...
// of each admitted leaf; the group's best distance (a min over its lanes) drives pruning, and the
// group's lexicographic minimum is reduced at the end. With G = 1 element records are wave-uniform
// and come through scalar loads. PRE adds a per-element box test before the exact arithmetic: the
// element is skipped by every lane whose box lower bound exceeds its best (never changes results:
// the box bound never exceeds the element's float distance, see box_lb's slack).
// One wave's work: queries [wave * 64 / G, ...) of a batch (`wave_ok` false for padding waves).
template <bool POINTS, int G, bool PRE, bool STATS>
__device__ __forceinline__ void closest_wave_body(const BNode *__restrict__ nodes, const void *__restrict__ elems,
                                                  const float *__restrict__ q, const uint32_t *__restrict__ order,
                                                  int64_t n, int64_t wave, bool wave_ok, float best_init, float scale,
                                                  int exclude_self, int32_t *__restrict__ out_idx,
                                                  float *__restrict__ out_cp, float *__restrict__ out_dist,
                                                  float *__restrict__ out_d2, unsigned long long *__restrict__ stats,
                                                  int32_t root = 0) {
  // STATS: per wave, count nodes loaded, box tests, leaf elements loaded and elements evaluated.
  uint32_t st_nodes = 0, st_boxes = 0, st_loaded = 0, st_tested = 0, st_plane = 0, st_near = 0;
  constexpr int QW = WAVE / G;  // queries per wave
  const int lane = lane_id();
  const int sub = lane % G;
  const int64_t t = wave * QW + lane / G;
  const bool inrange = wave_ok && t < n;
  bool active = inrange;
  int64_t qi = 0;
  float3 p = f3(0, 0, 0);
  if (active) {
    qi = order ? (int64_t)order[t] : t;
    p = load_q(q, qi);
    active = !(isnan(p.x) || isnan(p.y) || isnan(p.z));
  }
  const uint64_t amask = __ballot(active);
  const int nact = __popcll(amask);
  float m = fmaxf(scale, fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
  float sl = m * SLACK_REL;
  float3 qlo = f3(p.x - sl, p.y - sl, p.z - sl), qhi = f3(p.x + sl, p.y + sl, p.z + sl);
  float best = active ? best_init : -1.0f;  // group best (pruning); inactive lanes admit nothing
  float own = best;                          // this lane's best over the elements it tested
  float sbest = active ? sqrtf(best_init) : 0.0f;  // sqrt(best), for the G == 1 reject tests
  int32_t bidx = 0x7fffffff;
  float3 bcp = f3(0, 0, 0);
  const int32_t self = exclude_self ? (int32_t)qi : -1;
  const TriRec *__restrict__ tr = (const TriRec *)elems;
  const PointRec *__restrict__ pr = (const PointRec *)elems;
  if (G == 1 && active) {
    int32_t f, cnt;
    seed_leaf(nodes, root, qlo, qhi, f, cnt);
    for (int32_t e = f; e < f + cnt; ++e) {
      float d;
      int32_t id;
      float3 cp;
      if (POINTS) {
        const float4 pe = pr[e].p;
        cp = f3(pe.x, pe.y, pe.z);
        id = __float_as_int(pe.w);
        d = sqn3(sub3(cp, p));
      } else {
        const TriRec *tk = tr + e;
        d = tri_closest(tk->p0, tk->p1, tk->p2, tk->pl, tk->e0, tk->e1, tk->e2, p, cp);
        id = trirec_id(tk->pl);
      }
      if (id != self && better(d, id, own, bidx)) {
        own = d;
        bidx = id;
        bcp = cp;
      }
    }
    best = own;
    sbest = sqrtf(own);
  }
  int32_t s_ref = 0;
  float s_lx = 0, s_ly = 0, s_lz = 0, s_hx = 0, s_hy = 0, s_hz = 0;
  int32_t sp = 0;
  int32_t node = root;
  while (amask != 0) {
    node = uni(node);
    if (STATS) { st_nodes++; st_boxes += 2; }
    const BNode nd = nodes[node];
    const float a0 = unif(nd.a.x), a1 = unif(nd.a.y), a2 = unif(nd.a.z), a3 = unif(nd.a.w);
    const float b0 = unif(nd.b.x), b1 = unif(nd.b.y), b2 = unif(nd.b.z), b3 = unif(nd.b.w);
    const float c0 = unif(nd.c.x), c1 = unif(nd.c.y), c2 = unif(nd.c.z), c3 = unif(nd.c.w);
    const int32_t r0 = uni(nd.d.x), r1 = uni(nd.d.y), n0 = uni(nd.d.z), n1 = uni(nd.d.w);
    float lb0 = box_lb(qlo, qhi, a0, a1, a2, a3, b0, b1);
    float lb1 = box_lb(qlo, qhi, b2, b3, c0, c1, c2, c3);
    uint64_t m0 = __ballot(lb0 <= best), m1 = __ballot(lb1 <= best);
    bool first1 = 2 * __popcll(__ballot(lb1 < lb0) & amask) > nact;
    int32_t next = -1;
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
      const int c = first1 ? 1 - cc : cc;
      const uint64_t mc = c == 0 ? m0 : m1;
      if (mc == 0) continue;
      const int32_t r = c == 0 ? r0 : r1;
      const int32_t cnt = c == 0 ? n0 : n1;
      if (r < 0) {
        const int32_t f = ~r;
        if (G == 1 && !POINTS) {
          // Wave-uniform records (scalar loads). Cheap conservative rejects first, exact Thea
          // arithmetic only if some lane can still be improved (or tied) by this triangle.
          for (int32_t k = f; k < f + cnt; ++k) {
            const TriRec *tk = tr + k;
            if (STATS) st_loaded++;
            const float4 sbv = make_float4(unif(tk->sb.x), unif(tk->sb.y), unif(tk->sb.z), unif(tk->sb.w));
            bool sph = active;
            float dc2 = 0.0f;
            if (PRE) {
              const float dx = p.x - sbv.x, dy = p.y - sbv.y, dz = p.z - sbv.z;
              dc2 = dx * dx + dy * dy + dz * dz;
              const float T = sbest * REJ_REL + fabsf(sbv.w) + 2.0f * sl;
              sph = active && !(dc2 > T * T);
              if (__ballot(sph) == 0) continue;
            }
            const float4 p0 = make_float4(unif(tk->p0.x), unif(tk->p0.y), unif(tk->p0.z), unif(tk->p0.w));
            const float4 p1 = make_float4(unif(tk->p1.x), unif(tk->p1.y), unif(tk->p1.z), unif(tk->p1.w));
            const float4 p2 = make_float4(unif(tk->p2.x), unif(tk->p2.y), unif(tk->p2.z), unif(tk->p2.w));
            const float4 pl = make_float4(unif(tk->pl.x), unif(tk->pl.y), unif(tk->pl.z), unif(tk->pl.w));
            // HyperplaneN.hpp:145-160 signedDistance, shared with the exact test below
            const float sd = dot3(p, f3(p0.w, p1.w, p2.w)) - pl.x;
            const float Tp = sbest * REJ_REL + 2.0f * sl;
            if (PRE && __ballot(sph && !(tri_lb2(dc2, sd, sbv.w, sl) > Tp * Tp)) == 0) continue;
            if (STATS) {
              st_tested += WAVE;
              st_plane += __popcll(__ballot(sph && !(tri_lb2(dc2, sd, sbv.w, sl) > Tp * Tp)));
            }
            const float4 e0 = make_float4(unif(tk->e0.x), unif(tk->e0.y), unif(tk->e0.z), unif(tk->e0.w));
            const float4 e1 = make_float4(unif(tk->e1.x), unif(tk->e1.y), unif(tk->e1.z), unif(tk->e1.w));
            const float4 e2 = make_float4(unif(tk->e2.x), unif(tk->e2.y), unif(tk->e2.z), unif(tk->e2.w));
            const int32_t id = trirec_id(pl);
            const int ax = trirec_axis(pl);
            float3 cp;
            float d;
            if (ax == 0) d = tri_closest_ax<0>(p0, p1, p2, pl, e0, e1, e2, p, sd, cp);
            else if (ax == 1) d = tri_closest_ax<1>(p0, p1, p2, pl, e0, e1, e2, p, sd, cp);
            else d = tri_closest_ax<2>(p0, p1, p2, pl, e0, e1, e2, p, sd, cp);
            if (STATS && __ballot(active && d <= best * (1.0f + 1.0f / 256.0f))) st_near++;
            if (active && id != self && better(d, id, own, bidx)) {
              own = d;
              bidx = id;
              bcp = cp;
              best = d;
              sbest = sqrtf(d);
            }
          }
        } else {
          for (int32_t k0 = f; k0 < f + cnt; k0 += G) {
            const int32_t k = k0 + sub;
            bool valid = active && k < f + cnt;
            float d = INFINITY;
            float3 cp;
            int32_t id = 0x7fffffff;
            if (POINTS) {
              float4 e = pr[valid ? k : f].p;
              if (G == 1) e = make_float4(unif(e.x), unif(e.y), unif(e.z), unif(e.w));
              id = __float_as_int(e.w);
              cp = f3(e.x, e.y, e.z);
              d = sqn3(sub3(cp, p));
            } else {
              const TriRec *tk = tr + (valid ? k : f);
              if (STATS) st_loaded++;
              id = trirec_id(tk->pl);
              d = tri_closest(tk->p0, tk->p1, tk->p2, tk->pl, tk->e0, tk->e1, tk->e2, p, cp);
              if (STATS) st_tested++;
            }
            if (valid && id != self && better(d, id, own, bidx)) {
              own = d;
              bidx = id;
              bcp = cp;
            }
          }
        }
        // group best for pruning
        if (G > 1 || POINTS) {
          float gb = own;
#pragma unroll
          for (int off = 1; off < G; off <<= 1) gb = fminf(gb, __shfl_xor(gb, off, WAVE));
          best = active ? fminf(best, gb) : best;
        }
      } else if (next < 0) {
        next = r;
      } else if (sp < W_STACK) {  // never full when depth < 64 (checked by every caller)
        if (lane == sp) {
          s_ref = r;
          if (c == 0) { s_lx = a0; s_ly = a1; s_lz = a2; s_hx = a3; s_hy = b0; s_hz = b1; }
          else { s_lx = b2; s_ly = b3; s_lz = c0; s_hx = c1; s_hy = c2; s_hz = c3; }
        }
        ++sp;
      }
    }
    if (next >= 0) { node = next; continue; }
    bool found = false;
    while (sp > 0) {
      --sp;
      const float lx = lane_readf(s_lx, sp), ly = lane_readf(s_ly, sp), lz = lane_readf(s_lz, sp);
      const float hx = lane_readf(s_hx, sp), hy = lane_readf(s_hy, sp), hz = lane_readf(s_hz, sp);
      if (STATS) st_boxes++;
      if (__ballot(box_lb(qlo, qhi, lx, ly, lz, hx, hy, hz) <= best)) {
        node = lane_read(s_ref, sp);
        found = true;
        break;
      }
    }
    if (!found) break;
  }
  if (STATS && lane == 0 && amask != 0) {
    atomicAdd(&stats[0], 1ull);
    atomicAdd(&stats[1], (unsigned long long)nact / G);
    atomicAdd(&stats[2], (unsigned long long)st_nodes);
    atomicAdd(&stats[3], (unsigned long long)st_boxes);
    atomicAdd(&stats[4], (unsigned long long)st_loaded);
    atomicAdd(&stats[5], (unsigned long long)st_tested);
    atomicAdd(&stats[6], (unsigned long long)st_plane);
    atomicAdd(&stats[7], (unsigned long long)st_near);
  }
  // lexicographic (d, id) minimum over the group's lanes
#pragma unroll
  for (int off = 1; off < G; off <<= 1) {
    float od = __shfl_xor(own, off, WAVE);
    int32_t oi = __shfl_xor(bidx, off, WAVE);
    float ox = __shfl_xor(bcp.x, off, WAVE), oy = __shfl_xor(bcp.y, off, WAVE), oz = __shfl_xor(bcp.z, off, WAVE);
    if (better(od, oi, own, bidx)) { own = od; bidx = oi; bcp = f3(ox, oy, oz); }
  }
  if (!inrange || sub != 0) return;
  bool hit = active && bidx != 0x7fffffff;
  out_idx[qi] = hit ? bidx : -1;
  if (out_cp) {
    out_cp[3 * qi] = hit ? bcp.x : 0.0f;
    out_cp[3 * qi + 1] = hit ? bcp.y : 0.0f;
    out_cp[3 * qi + 2] = hit ? bcp.z : 0.0f;
  }
  if (out_dist) out_dist[qi] = hit ? sqrtf(own) : -1.0f;
  if (out_d2) out_d2[qi] = hit ? own : -1.0f;
}

//--------------------------------------------------------------------------------------------------
// Compacted wave traversal (triangles). The wave walks one tree for its 64 queries as above, but a
// triangle of an admitted leaf is not evaluated by the whole wave: every lane applies the cheap
// conservative rejects (bounding sphere, then plane and in-plane disc distance, tri_lb2) against its
// own best, and each surviving
// (lane, triangle) pair is appended to a per-wave LDS queue. Whenever 64 pairs are queued the wave
// evaluates them one per lane with Thea's exact arithmetic and merges each result into its owner
// lane's best through an LDS 64-bit atomic min on (float bits of d^2) << 32 | triangle id, which
// orders pairs exactly as the lexicographic (d^2, index) rule does (d^2 >= 0; NaN never wins).
// Deferring an evaluation only delays the tightening of the pruning bound, never the result.
//--------------------------------------------------------------------------------------------------
// Resident waves per SIMD the compacted closest-point kernels are compiled for (single-stream
// launches). Their walks are chains of dependent loads, so more waves hide more latency: at 8 (a
// 44-byte spill) the C3 kernel (10M queries on 1M triangles) takes 18.1 ms against 22.4 ms at the
// 6 its 74 registers would otherwise allow. The k-NN wave
// kernels keep the compiler's choice: their K-pair lists spill at higher targets (K = 16 at 8
// waves: 14.6 ms against 3.85 ms on C2; at 6: 3.87 ms).
constexpr int CMP_WAVES = 8;

struct CmpWave {
  float4 qp[WAVE];                  // owner's query point; w = excluded triangle id as int bits
  float4 qcp[WAVE];                 // owner's closest point of its current best
  unsigned long long key[WAVE];     // owner's best (d^2 bits << 32 | id)
  int32_t qe[2 * WAVE];             // queued triangle record slots
  uint8_t qo[2 * WAVE];             // queued owner lanes
};

__device__ __forceinline__ unsigned long long cmp_key(float d, int32_t id) {
  return ((unsigned long long)(uint32_t)__float_as_int(d + 0.0f) << 32) | (uint32_t)id;
}
__device__ __forceinline__ float cmp_key_d(unsigned long long k) { return __int_as_float((int)(uint32_t)(k >> 32)); }

// Evaluates queue entries [0, m) (one per lane), then moves entries [m, qn) to the front.
__device__ __forceinline__ void cmp_flush(CmpWave &W, const TriRec *__restrict__ tr, int lane, int m, int qn) {
  if (lane < m) {
    const int32_t e = W.qe[lane];
    const int o = W.qo[lane];
    const float4 po = W.qp[o];
    const TriRec *tk = tr + e;
    const float4 p0 = tk->p0, p1 = tk->p1, p2 = tk->p2, pl = tk->pl, e0 = tk->e0, e1 = tk->e1, e2 = tk->e2;
    float3 cp;
    const float d = tri_closest(p0, p1, p2, pl, e0, e1, e2, f3(po.x, po.y, po.z), cp);
    const int32_t id = trirec_id(pl);
    unsigned long long k = cmp_key(d, id);
    const bool ok = id != __float_as_int(po.w) && !isnan(d);
    if (ok) atomicMin(&W.key[o], k);
    __builtin_amdgcn_wave_barrier();
    // (d, id) is unique per owner: the pair whose key survived writes the closest point
    if (ok && W.key[o] == k) W.qcp[o] = make_float4(cp.x, cp.y, cp.z, 0.0f);
  }
  __builtin_amdgcn_wave_barrier();
  const int rest = qn - m;
  int32_t e = 0;
  uint8_t o = 0;
  if (lane < rest) { e = W.qe[m + lane]; o = W.qo[m + lane]; }
  __builtin_amdgcn_wave_barrier();
  if (lane < rest) { W.qe[lane] = e; W.qo[lane] = o; }
  __builtin_amdgcn_wave_barrier();
}

// PF (batches too small to fill the GPU): a leaf's screening data (bounding spheres, normals and
// plane offsets) arrives with one vector load per lane and is read per triangle with v_readlane, so
// a wave waits once per leaf instead of once per triangle; large batches hide that latency with
// other waves and keep the scalar loads, which cost no vector instructions.
template <bool STATS, bool PF = false, bool LB = false>
__device__ __forceinline__ void closest_cmp_body(CmpWave &W, const BNode *__restrict__ nodes,
                                                 const TriRec *__restrict__ tr, const float *__restrict__ q,
                                                 const uint32_t *__restrict__ order, int64_t n, int64_t wave,
                                                 bool wave_ok, float best_init, float scale, int exclude_self,
                                                 int32_t *__restrict__ out_idx, float *__restrict__ out_cp,
                                                 float *__restrict__ out_dist, float *__restrict__ out_d2,
                                                 unsigned long long *__restrict__ stats, int32_t root = 0,
                                                 int child_order = 0,
                                                 const LeafBound *__restrict__ lbnd = nullptr,
                                                 const ScreenRec *__restrict__ scr = nullptr) {
  uint32_t st_nodes = 0, st_boxes = 0, st_loaded = 0, st_flush = 0, st_pairs = 0;
  const int lane = lane_id();
  const int64_t t = wave * WAVE + lane;
  const bool inrange = wave_ok && t < n;
  bool active = inrange;
  int64_t qi = 0;
  float3 p = f3(0, 0, 0);
  if (active) {
    qi = order ? (int64_t)order[t] : t;
    p = load_q(q, qi);
    active = !(isnan(p.x) || isnan(p.y) || isnan(p.z));
  }
  const uint64_t amask = ballot(active);
  const int nact = __popcll(amask);
  const float m = fmaxf(scale, fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
  const float sl = m * SLACK_REL;
  const float3 qlo = f3(p.x - sl, p.y - sl, p.z - sl), qhi = f3(p.x + sl, p.y + sl, p.z + sl);
  float best = active ? best_init : -1.0f;  // inactive lanes admit nothing
  float rej2 = active ? rej_sq(best_init, sl) : 0.0f;  // refreshed whenever best changes
  const int32_t self = exclude_self ? (int32_t)qi : -1;
  unsigned long long key0 = cmp_key(active ? best_init : 0.0f, 0x7fffffff);
  float3 cp0 = f3(0, 0, 0);
  // best_init < 0 (a zero distance bound) admits nothing; its key must not be compared as bits
  if (active && best_init >= 0.0f) {
    int32_t f, cnt;
    seed_leaf(nodes, root, qlo, qhi, f, cnt);
    for (int32_t e = f; e < f + cnt; ++e) {
      const TriRec *tk = tr + e;
      float3 cp;
      const float d = tri_closest(tk->p0, tk->p1, tk->p2, tk->pl, tk->e0, tk->e1, tk->e2, p, cp);
      const int32_t id = trirec_id(tk->pl);
      const unsigned long long kk = cmp_key(d, id);
      if (id != self && !isnan(d) && kk < key0) {
        key0 = kk;
        cp0 = cp;
      }
    }
    best = cmp_key_d(key0);
    rej2 = rej_sq(best, sl);
  }
  W.qp[lane] = make_float4(p.x, p.y, p.z, __int_as_float(self));
  W.key[lane] = key0;
  W.qcp[lane] = make_float4(cp0.x, cp0.y, cp0.z, 0.0f);
  __builtin_amdgcn_wave_barrier();
  int qn = 0;  // queued pairs (wave-uniform)
  int32_t s_ref = 0;
  float s_lx = 0, s_ly = 0, s_lz = 0, s_hx = 0, s_hy = 0, s_hz = 0;
  int32_t sp = 0;
  int32_t node = root;
  while (amask != 0) {
    node = uni(node);
    if (STATS) { st_nodes++; st_boxes += 2; }
    const BNode nd = nodes[node];
    const float a0 = unif(nd.a.x), a1 = unif(nd.a.y), a2 = unif(nd.a.z), a3 = unif(nd.a.w);
    const float b0 = unif(nd.b.x), b1 = unif(nd.b.y), b2 = unif(nd.b.z), b3 = unif(nd.b.w);
    const float c0 = unif(nd.c.x), c1 = unif(nd.c.y), c2 = unif(nd.c.z), c3 = unif(nd.c.w);
    const int32_t r0 = uni(nd.d.x), r1 = uni(nd.d.y), n0 = uni(nd.d.z), n1 = uni(nd.d.w);
    const float lb0 = box_lb(qlo, qhi, a0, a1, a2, a3, b0, b1);
    const float lb1 = box_lb(qlo, qhi, b2, b3, c0, c1, c2, c3);
    const uint64_t m0 = ballot(lb0 <= best), m1 = ballot(lb1 <= best);
    // visiting order (CHILD_*): the child most lanes find nearer, the child the wave's first active
    // query finds nearer, or the build's order (ballots only, so the choice costs no registers)
    const uint64_t near1 = ballot(lb1 < lb0) & amask;
    const bool first1 = child_order == CHILD_MAJORITY ? 2 * __popcll(near1) > nact
                        : child_order == CHILD_LEADER ? (near1 & (amask & (~amask + 1))) != 0
                                                      : false;
    int32_t next = -1;
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
      const int c = first1 ? 1 - cc : cc;
      const uint64_t mc = c == 0 ? m0 : m1;
      if (mc == 0) continue;
      const int32_t r = c == 0 ? r0 : r1;
      const int32_t cnt = c == 0 ? n0 : n1;
      if (r < 0) {
        const int32_t f = ~r;
        // leaf-level reject: lanes whose bound to the whole leaf exceeds their threshold drop out,
        // and a leaf no lane keeps is skipped without screening its triangles
        uint64_t ml = mc;
        if (LB && lbnd) {
          const LeafBound *L = lbnd + f;
          const float lx = unif(L->a.x), ly = unif(L->a.y), lz = unif(L->a.z), lr = unif(L->a.w);
          const float nx = unif(L->b.x), ny = unif(L->b.y), nz = unif(L->b.z), lt = unif(L->b.w);
          const float dx = p.x - lx, dy = p.y - ly, dz = p.z - lz;
          const float dc2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
          const float sd = fmaf(dx, nx, fmaf(dy, ny, dz * nz));
          ml &= ballot(!(leaf_lb2(dc2, sd, lt, lr, sl) > rej2));
          if (ml == 0) continue;
        }
        float4 lsb = make_float4(0.0f, 0.0f, 0.0f, 0.0f), lpl = lsb;  // PF: lane j holds triangle f + j's
        if (PF && lane < cnt) {
          const ScreenRec *sj = scr + f + lane;
          lsb = sj->c;
          lpl = sj->n;
        }
        for (int32_t k = f; k < f + cnt; ++k) {
          const ScreenRec *sk = scr + k;
          if (STATS) st_loaded++;
          const int jl = k - f;
          const float4 sbv = PF ? make_float4(lane_readf(lsb.x, jl), lane_readf(lsb.y, jl), lane_readf(lsb.z, jl),
                                              lane_readf(lsb.w, jl))
                                : make_float4(unif(sk->c.x), unif(sk->c.y), unif(sk->c.z), unif(sk->c.w));
          // conservative reject (fused multiply-adds: the slack dwarfs their rounding): the plane
          // and disc bound. Large batches apply it alone (a bounding-sphere test first, to skip the
          // plane's scalar loads, cost more than it saved: C3 18.1 against 19.4 ms per launch);
          // prefetched leaves test the sphere first, which skips the plane's four lane reads when
          // no lane passes (C1: 1.45 against 1.57 ms). sqrt(rej2) widened by 2^-20 covers the
          // hardware square root's rounding.
          const float dx = p.x - sbv.x, dy = p.y - sbv.y, dz = p.z - sbv.z;
          const float dc2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
          if (PF) {
            const float T = fmaf(__builtin_amdgcn_sqrtf(rej2), 1.0f + 1.0f / 1048576.0f, fabsf(sbv.w));
            if ((ballot(!(dc2 > T * T)) & ml) == 0) continue;
          }
          const float nx = PF ? lane_readf(lpl.x, jl) : unif(sk->n.x), ny = PF ? lane_readf(lpl.y, jl) : unif(sk->n.y);
          const float nz = PF ? lane_readf(lpl.z, jl) : unif(sk->n.z);
          const float sd = fmaf(dx, nx, fmaf(dy, ny, dz * nz));
          // every lane computes the bound (no divergence); lanes that did not admit the leaf are
          // masked out of the ballot
          const uint64_t nb = ballot(!(tri_lb2(dc2, sd, sbv.w, sl) > rej2)) & ml;
          if (nb == 0) continue;
          const bool need = (nb >> lane) & 1;
          if (need) {
            const int pos = qn + __popcll(nb & lanemask_lt());
            W.qe[pos] = k;
            W.qo[pos] = (uint8_t)lane;
          }
          qn += __popcll(nb);
          if (STATS) st_pairs += __popcll(nb);
          if (qn >= WAVE) {
            __builtin_amdgcn_wave_barrier();
            cmp_flush(W, tr, lane, WAVE, qn);
            qn -= WAVE;
            if (STATS) st_flush++;
            if (active) {
              best = cmp_key_d(W.key[lane]);
              rej2 = rej_sq(best, sl);
            }
          }
        }
      } else if (next < 0) {
        next = r;
      } else if (sp < W_STACK) {  // never full when depth < 64 (checked by every caller)
        if (lane == sp) {
          s_ref = r;
          if (c == 0) { s_lx = a0; s_ly = a1; s_lz = a2; s_hx = a3; s_hy = b0; s_hz = b1; }
          else { s_lx = b2; s_ly = b3; s_lz = c0; s_hx = c1; s_hy = c2; s_hz = c3; }
        }
        ++sp;
      }
    }
    if (next >= 0) { node = next; continue; }
    bool found = false;
    while (sp > 0) {
      --sp;
      const float lx = lane_readf(s_lx, sp), ly = lane_readf(s_ly, sp), lz = lane_readf(s_lz, sp);
      const float hx = lane_readf(s_hx, sp), hy = lane_readf(s_hy, sp), hz = lane_readf(s_hz, sp);
      if (STATS) st_boxes++;
      if (ballot(box_lb(qlo, qhi, lx, ly, lz, hx, hy, hz) <= best)) {
        node = lane_read(s_ref, sp);
        found = true;
        break;
      }
    }
    if (!found) break;
  }
  if (qn > 0) {
    __builtin_amdgcn_wave_barrier();
    cmp_flush(W, tr, lane, qn, qn);
    if (STATS) st_flush++;
  }
  __builtin_amdgcn_wave_barrier();
  if (STATS && lane == 0 && amask != 0) {
    atomicAdd(&stats[0], 1ull);
    atomicAdd(&stats[1], (unsigned long long)nact);
    atomicAdd(&stats[2], (unsigned long long)st_nodes);
    atomicAdd(&stats[3], (unsigned long long)st_boxes);
    atomicAdd(&stats[4], (unsigned long long)st_loaded);
    atomicAdd(&stats[5], (unsigned long long)st_flush * WAVE);
    atomicAdd(&stats[6], (unsigned long long)st_pairs);
    atomicAdd(&stats[7], 0ull);
  }
  if (!inrange) return;
  const unsigned long long key = W.key[lane];
  const int32_t bidx = (int32_t)(uint32_t)(key & 0xffffffffull);
  const float own = cmp_key_d(key);
  const bool hit = active && bidx != 0x7fffffff;
  out_idx[qi] = hit ? bidx : -1;
  if (out_cp) {
    const float4 c = W.qcp[lane];
    out_cp[3 * qi] = hit ? c.x : 0.0f;
    out_cp[3 * qi + 1] = hit ? c.y : 0.0f;
    out_cp[3 * qi + 2] = hit ? c.z : 0.0f;
  }
  if (out_dist) out_dist[qi] = hit ? sqrtf(own) : -1.0f;
  if (out_d2) out_d2[qi] = hit ? own : -1.0f;
}

// WAVES: resident waves per SIMD to compile for (1: the compiler's choice, 6 here, no spill).
template <bool STATS = false, int WAVES = 1, bool PF = false>
__global__ void __launch_bounds__(W_BLOCK, WAVES) k_closest_cmp(const BNode *__restrict__ nodes, const TriRec *__restrict__ tr,
                                                         const float *__restrict__ q, const uint32_t *__restrict__ order,
                                                         int64_t n, uint32_t nb_real, float best_init, float scale,
                                                         int exclude_self, int32_t *__restrict__ out_idx,
                                                         float *__restrict__ out_cp, float *__restrict__ out_dist,
                                                         float *__restrict__ out_d2,
                                                         unsigned long long *__restrict__ stats = nullptr,
                                                         int child_order = CHILD_MAJORITY,
                                                         const LeafBound *__restrict__ lbnd = nullptr,
                                                         const ScreenRec *__restrict__ scr = nullptr) {
  __shared__ CmpWave s_w[W_WAVES];
  const uint32_t lb_blk = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t wave = (int64_t)lb_blk * W_WAVES + (threadIdx.x / WAVE);
  closest_cmp_body<STATS, PF, true>(s_w[threadIdx.x / WAVE], nodes, tr, q, order, n, wave, lb_blk < nb_real,
                                    best_init, scale, exclude_self, out_idx, out_cp, out_dist, out_d2, stats, 0,
                                    child_order, lbnd, scr);
}

template <bool POINTS, int G, bool PRE, bool STATS = false>
__global__ void __launch_bounds__(W_BLOCK) k_closest_wave(const BNode *__restrict__ nodes, const void *__restrict__ elems,
                                                          const float *__restrict__ q, const uint32_t *__restrict__ order,
                                                          int64_t n, uint32_t nb_real, float best_init, float scale,
                                                          int exclude_self, int32_t *__restrict__ out_idx,
                                                          float *__restrict__ out_cp, float *__restrict__ out_dist,
                                                          float *__restrict__ out_d2,
                                                          unsigned long long *__restrict__ stats = nullptr) {
  const uint32_t lb_blk = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t wave = (int64_t)lb_blk * W_WAVES + (threadIdx.x / WAVE);
  closest_wave_body<POINTS, G, PRE, STATS>(nodes, elems, q, order, n, wave, lb_blk < nb_real, best_init, scale,
                                           exclude_self, out_idx, out_cp, out_dist, out_d2, stats);
}

// Many batches (each against its own BVH) in one launch: the waves of segment s are
// [seg[s].wave_begin, seg[s + 1].wave_begin); a wave finds its segment by binary search.
__device__ __forceinline__ int find_segment(const SegDesc *__restrict__ seg, int nseg, int64_t wave) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if ((int64_t)seg[mid].wave_begin <= wave) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// WAVES: resident waves per SIMD to compile for. The request server, whose four streams run these
// launches concurrently, keeps the spill-free 6 (the 8-wave variant, which spills 32 bytes, cut its
// saturated throughput 7%); the batched plan's single stream takes the 8-wave variant.
// WPB: waves per workgroup (the launch's block shape), one wave's work per 64 lanes either way.
template <bool POINTS, int WAVES, int WPB = W_WAVES>
__global__ void __launch_bounds__(WPB * WAVE, WAVES) k_closest_multi(const SegDesc *__restrict__ seg, int nseg,
                                                                uint64_t total_waves, uint32_t nb_real) {
  const uint32_t lb_blk = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t gw = (int64_t)lb_blk * WPB + (threadIdx.x / WAVE);
  if (lb_blk >= nb_real || gw >= (int64_t)total_waves) return;
  const int si = uni(find_segment(seg, nseg, gw));
  const SegDesc &d = seg[si];
  if constexpr (POINTS) {
    closest_wave_body<true, 1, true, false>(readonly_ptr<BNode>(d.nodes), readonly_ptr<float>(d.elems), readonly_ptr<float>(d.q),
                                            readonly_ptr<uint32_t>(d.order), d.n,
                                            gw - (int64_t)d.wave_begin, true, d.best_init, d.scale, d.exclude_self,
                                            d.idx, d.cp, d.dist, d.d2, nullptr, d.root);
  } else {
    __shared__ CmpWave s_w[WPB];
    closest_cmp_body<false>(s_w[threadIdx.x / WAVE], readonly_ptr<BNode>(d.nodes), readonly_ptr<TriRec>(d.elems),
                            readonly_ptr<float>(d.q), readonly_ptr<uint32_t>(d.order),
                            d.n, gw - (int64_t)d.wave_begin, true, d.best_init, d.scale, d.exclude_self, d.idx, d.cp,
                            d.dist, d.d2, nullptr, d.root, d.child_order, nullptr,
                            readonly_ptr<ScreenRec>(d.screen));
  }
}

//--------------------------------------------------------------------------------------------------
// k nearest (sorted list of K (d^2, id) pairs per lane in registers)
//--------------------------------------------------------------------------------------------------
template <int STACK, bool POINTS, int K>
__device__ __forceinline__ void knn_lane_body(int32_t *__restrict__ s_ref, float *__restrict__ s_lb,
                                              const BNode *__restrict__ nodes, const void *__restrict__ elems,
                                              const float *__restrict__ q, int64_t qi, int k, float best_init,
                                              float scale, int exclude_self, int32_t *__restrict__ out_idx,
                                              float *__restrict__ out_dist, float *__restrict__ out_d2,
                                              int32_t root) {
  const int tid = threadIdx.x;
  float3 p = load_q(q, qi);
  float m = fmaxf(scale, fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
  float sl = m * SLACK_REL;
  float3 qlo = f3(p.x - sl, p.y - sl, p.z - sl), qhi = f3(p.x + sl, p.y + sl, p.z + sl);
  float kd[K];
  int32_t ki[K];
#pragma unroll
  for (int j = 0; j < K; ++j) { kd[j] = best_init; ki[j] = 0x7fffffff; }
  // Only the first k slots are live; the worst live slot is the pruning bound.
  const int32_t self = exclude_self ? (int32_t)qi : -1;
  int sp = 0;
  LaneStack<KN_REG, KN_LB> stk;
  int32_t node = root;
  const TriRec *__restrict__ tr = (const TriRec *)elems;
  const PointRec *__restrict__ pr = (const PointRec *)elems;
  // The k-th best pair (slot k - 1): the pruning bound and the admission threshold.
  auto kth = [&](float &bd, int32_t &bi) {
    bd = kd[K - 1];
    bi = ki[K - 1];
#pragma unroll
    for (int j = 0; j < K - 1; ++j)
      if (j == k - 1) { bd = kd[j]; bi = ki[j]; }
  };
  auto bound = [&]() {
    float bd;
    int32_t bi;
    kth(bd, bi);
    return bd;
  };
  while (true) {
    const BNode nd = nodes[node];
    float lb0 = box_lb(qlo, qhi, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y);
    float lb1 = box_lb(qlo, qhi, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w);
    int32_t r0 = nd.d.x, r1 = nd.d.y, c0 = nd.d.z, c1 = nd.d.w;
    if (lb1 < lb0) {
      float tl = lb0; lb0 = lb1; lb1 = tl;
      int32_t tr_ = r0; r0 = r1; r1 = tr_;
      int32_t tc = c0; c0 = c1; c1 = tc;
    }
    int32_t next = -1;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      float lb = c == 0 ? lb0 : lb1;
      int32_t r = c == 0 ? r0 : r1;
      int32_t cnt = c == 0 ? c0 : c1;
      if (!(lb <= bound())) continue;
      if (r < 0) {
        int32_t f = ~r;
        for (int32_t kk = f; kk < f + cnt; ++kk) {
          float d;
          int32_t id;
          if (POINTS) {
            float4 e = pr[kk].p;
            id = __float_as_int(e.w);
            d = sqn3(sub3(f3(e.x, e.y, e.z), p));
          } else {
            const TriRec *t = tr + kk;
            id = trirec_id(t->pl);
            float3 cp;
            d = tri_closest(t->p0, t->p1, t->p2, t->pl, t->e0, t->e1, t->e2, p, cp);
          }
          float bd;
          int32_t bi;
          kth(bd, bi);
          if (id != self && better(d, id, bd, bi)) knn_insert<K>(kd, ki, d, id);
        }
      } else if (next < 0) {
        next = r;
      } else if (sp < STACK) {  // never taken when depth < STACK (checked by every caller)
        stk.push(sp, r, lb, s_ref, s_lb, tid);
        ++sp;
      }
    }
    if (next >= 0) { node = next; continue; }
    bool found = false;
    while (sp > 0) {
      --sp;
      int32_t pr_ref;
      float pr_lb;
      stk.pop(sp, pr_ref, pr_lb, s_ref, s_lb, tid);
      if (pr_lb <= bound()) { node = pr_ref; found = true; break; }
    }
    if (!found) break;
  }
  for (int j = 0; j < K; ++j) {
    if (j >= k) break;
    bool hit = ki[j] != 0x7fffffff;
    out_idx[qi * k + j] = hit ? ki[j] : -1;
    if (out_dist) out_dist[qi * k + j] = hit ? sqrtf(kd[j]) : -1.0f;
    if (out_d2) out_d2[qi * k + j] = hit ? kd[j] : -1.0f;
  }
}

template <int STACK, bool POINTS, int K>
__global__ void __launch_bounds__(Q_BLOCK) k_knn(const BNode *__restrict__ nodes, const void *__restrict__ elems,
                                                 const float *__restrict__ q, const uint32_t *__restrict__ order,
                                                 int64_t n, uint32_t nb_real, int k, float best_init, float scale,
                                                 int exclude_self, int32_t *__restrict__ out_idx,
                                                 float *__restrict__ out_dist, float *__restrict__ out_d2) {
  __shared__ int32_t s_ref[lane_lds_entries<STACK, KN_REG>() * Q_BLOCK];
  __shared__ float s_lb[KN_LB ? lane_lds_entries<STACK, KN_REG>() * Q_BLOCK : 1];
  bool active;
  const int64_t qi = query_slot(nb_real, n, order, active);
  if (!active) return;
  knn_lane_body<STACK, POINTS, K>(s_ref, s_lb, nodes, elems, q, qi, k, best_init, scale, exclude_self, out_idx,
                                  out_dist, out_d2, 0);
}

// Many closest-point or k-NN batches in one launch, one query per lane: one 64-lane block per
// SEG_WAVE_QUERIES queries of a segment (the segment table of the wave-cooperative multi kernels).
template <int STACK, bool POINTS>
__global__ void __launch_bounds__(Q_BLOCK) k_closest_lane_multi(const SegDesc *__restrict__ seg, int nseg,
                                                                uint64_t total_waves, uint32_t nb_real) {
  __shared__ int32_t s_ref[lane_lds_entries<STACK, CL_REG>() * Q_BLOCK];
  __shared__ float s_lb[CL_LB ? lane_lds_entries<STACK, CL_REG>() * Q_BLOCK : 1];
  const uint32_t lb_blk = xcd_remap(blockIdx.x, gridDim.x);
  if (lb_blk >= nb_real || (uint64_t)lb_blk >= total_waves) return;
  const int si = uni(find_segment(seg, nseg, (int64_t)lb_blk));
  const SegDesc &d = seg[si];
  const int64_t t = ((int64_t)lb_blk - (int64_t)d.wave_begin) * Q_BLOCK + threadIdx.x;
  if (t >= d.n) return;
  const int64_t qi = d.order ? (int64_t)d.order[t] : t;
  closest_lane_body<STACK, POINTS>(s_ref, s_lb, readonly_ptr<BNode>(d.nodes), readonly_ptr<float>(d.elems),
                                   readonly_ptr<float>(d.q), qi, d.best_init, d.scale,
                                   d.exclude_self, d.idx, d.cp, d.dist, d.d2, d.root);
}

template <int STACK, bool POINTS, int K>
__global__ void __launch_bounds__(Q_BLOCK) k_knn_lane_multi(const SegDesc *__restrict__ seg, int nseg,
                                                            uint64_t total_waves, uint32_t nb_real) {
  __shared__ int32_t s_ref[lane_lds_entries<STACK, KN_REG>() * Q_BLOCK];
  __shared__ float s_lb[KN_LB ? lane_lds_entries<STACK, KN_REG>() * Q_BLOCK : 1];
  const uint32_t lb_blk = xcd_remap(blockIdx.x, gridDim.x);
  if (lb_blk >= nb_real || (uint64_t)lb_blk >= total_waves) return;
  const int si = uni(find_segment(seg, nseg, (int64_t)lb_blk));
  const SegDesc &d = seg[si];
  const int64_t t = ((int64_t)lb_blk - (int64_t)d.wave_begin) * Q_BLOCK + threadIdx.x;
  if (t >= d.n) return;
  const int64_t qi = d.order ? (int64_t)d.order[t] : t;
  knn_lane_body<STACK, POINTS, K>(s_ref, s_lb, readonly_ptr<BNode>(d.nodes), readonly_ptr<float>(d.elems),
                                  readonly_ptr<float>(d.q), qi, d.k, d.best_init, d.scale,
                                  d.exclude_self, d.idx, d.dist, d.d2, d.root);
}

//--------------------------------------------------------------------------------------------------
// Sparse batches on a large tree, G lanes per query
//
// A lane walk costs its query's whole search: a query far inside a closed surface meets nearly
// every leaf, and a launch of a few thousand such queries lasts as long as its slowest one. Here G
// lanes (a power of two, 2..64, aligned within the wave) share a query. Lane gl of the group owns
// the subtrees reached from the root by taking both children at depths 0..GRP_FREE-1 and child
// slot bit (depth - GRP_FREE) of gl at the next log2(G) depths, so the lanes of a group partition
// the tree into 4G subtrees spread over it; an element range met above that depth is evaluated by
// the one lane whose remaining route bits are zero. Each lane walks its subtrees one after another
// with its own stack; every step the group's bound (the minimum of the lanes' bounds, a distance some element
// of the group attains) is refreshed with cross-lane shuffles, so the whole wave runs one uniform
// loop. The answer is the group's (d^2, index) minimum, the same pair a single lane finds.
//--------------------------------------------------------------------------------------------------
constexpr int GRP_FREE = 2;

__device__ __forceinline__ float group_min(float v, int g) {
  for (int m = 1; m < g; m <<= 1) v = fminf(v, __shfl_xor(v, m));
  return v;
}

// Minimum over the group of the (d^2, index) pair packed as d^2's bits above the index (d^2 >= 0,
// so the bit order is the value order).
__device__ __forceinline__ uint64_t group_min_key(uint64_t v, int g) {
  for (int m = 1; m < g; m <<= 1) {
    const uint32_t lo = __shfl_xor((uint32_t)v, m), hi = __shfl_xor((uint32_t)(v >> 32), m);
    const uint64_t o = ((uint64_t)hi << 32) | lo;
    v = o < v ? o : v;
  }
  return v;
}

__device__ __forceinline__ uint64_t pair_key(float d, int32_t i) {
  return ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)i;
}

// Route pth (0 .. 2^GRP_FREE - 1) of lane gl through the first GRP_FREE + log2(G) levels: element
// ranges on the way go to `leaf`; returns the subtree root at the route's end (with its lower bound
// in lb_out), or -1 when the route ends early or its child's lower bound exceeds `bound()`. A lane
// walks its routes one after another, so its stack never holds more than one path.
template <typename LB, typename Leaf>
__device__ __forceinline__ int32_t group_route(const BNode *__restrict__ nodes, int32_t root, int lg, int gl, int pth,
                                               float3 qlo, float3 qhi, LB bound, Leaf leaf, float &lb_out) {
  const int depth = GRP_FREE + lg;
  const uint32_t route = (uint32_t)pth | ((uint32_t)gl << GRP_FREE);
  int32_t node = root;
  for (int j = 0; j < depth; ++j) {
    const BNode nd = nodes[node];
    const int slot = (route >> j) & 1;
    const float lb = slot ? box_lb(qlo, qhi, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w)
                          : box_lb(qlo, qhi, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y);
    const int32_t r = slot ? nd.d.y : nd.d.x;
    const int32_t cnt = slot ? nd.d.w : nd.d.z;
    if (!(lb <= bound())) return -1;
    if (r < 0) {
      // every route sharing bits 0..j meets this range; the one whose higher bits are zero takes it
      if ((route >> (j + 1)) == 0) leaf(~r, cnt);
      return -1;
    }
    if (j + 1 == depth) {
      lb_out = lb;
      return r;
    }
    node = r;
  }
  return -1;
}

template <int STACK, bool POINTS>
__device__ __forceinline__ void closest_group_body(int32_t *__restrict__ s_ref, float *__restrict__ s_lb,
                                                   const BNode *__restrict__ nodes, const void *__restrict__ elems,
                                                   const float *__restrict__ q, int64_t qi, bool live, int lg,
                                                   float best_init, float scale, int exclude_self,
                                                   int32_t *__restrict__ out_idx, float *__restrict__ out_cp,
                                                   float *__restrict__ out_dist, float *__restrict__ out_d2,
                                                   int32_t root) {
  const int tid = threadIdx.x;
  const int g = 1 << lg, gl = tid & (g - 1);
  float3 p = live ? load_q(q, qi) : f3(0, 0, 0);
  float m = fmaxf(scale, fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
  float sl = m * SLACK_REL;
  float3 qlo = f3(p.x - sl, p.y - sl, p.z - sl), qhi = f3(p.x + sl, p.y + sl, p.z + sl);
  float best = best_init;
  int32_t bidx = 0x7fffffff;
  float3 bcp = f3(0, 0, 0);
  const int32_t self = exclude_self ? (int32_t)qi : -1;
  const TriRec *__restrict__ tr = (const TriRec *)elems;
  const PointRec *__restrict__ pr = (const PointRec *)elems;
  auto eval = [&](int32_t f, int32_t cnt) {
    for (int32_t k = f; k < f + cnt; ++k) {
      float d;
      float3 cp;
      int32_t id;
      if (POINTS) {
        float4 e = pr[k].p;
        id = __float_as_int(e.w);
        cp = f3(e.x, e.y, e.z);
        d = sqn3(sub3(cp, p));
      } else {
        const TriRec *t = tr + k;
        id = trirec_id(t->pl);
        d = tri_closest(t->p0, t->p1, t->p2, t->pl, t->e0, t->e1, t->e2, p, cp);
      }
      if (id != self && better(d, id, best, bidx)) {
        best = d;
        bidx = id;
        bcp = cp;
      }
    }
  };
  int sp = 0;
  LaneStack<CL_REG, CL_LB> stk;
  int pth = live ? 0 : 1 << GRP_FREE;  // next route
  float gbest = best_init;
  int32_t node = -1;
  while (true) {
    if (node < 0) {
      while (sp > 0) {
        --sp;
        int32_t pr_ref;
        float pr_lb;
        stk.pop(sp, pr_ref, pr_lb, s_ref, s_lb, tid);
        if (pr_lb <= gbest) { node = pr_ref; break; }
      }
      if (node < 0 && pth < (1 << GRP_FREE)) {
        float lb = 0;
        node = group_route(nodes, root, lg, gl, pth++, qlo, qhi, [&]() { return fminf(gbest, best); }, eval, lb);
      }
    }
    if (node >= 0) {
      const BNode nd = nodes[node];
      float lb0 = box_lb(qlo, qhi, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y);
      float lb1 = box_lb(qlo, qhi, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w);
      int32_t r0 = nd.d.x, r1 = nd.d.y, c0 = nd.d.z, c1 = nd.d.w;
      if (lb1 < lb0) {
        float tl = lb0; lb0 = lb1; lb1 = tl;
        int32_t tr_ = r0; r0 = r1; r1 = tr_;
        int32_t tc = c0; c0 = c1; c1 = tc;
      }
      int32_t next = -1;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        float lb = c == 0 ? lb0 : lb1;
        int32_t r = c == 0 ? r0 : r1;
        int32_t cnt = c == 0 ? c0 : c1;
        if (!(lb <= fminf(gbest, best))) continue;
        if (r < 0) {
          eval(~r, cnt);
        } else if (next < 0) {
          next = r;
        } else if (sp < STACK) {  // never taken when depth < STACK (checked by every caller)
          stk.push(sp, r, lb, s_ref, s_lb, tid);
          ++sp;
        }
      }
      node = next;
    }
    gbest = group_min(best, g);
    if (!__any(node >= 0 || sp > 0 || pth < (1 << GRP_FREE))) break;
  }
  const uint64_t key = pair_key(best, bidx);
  const uint64_t gkey = group_min_key(key, g);
  if (!live) return;
  const bool hit = (uint32_t)gkey != 0x7fffffffu;
  if (hit ? key != gkey : gl != 0) return;
  out_idx[qi] = hit ? bidx : -1;
  if (out_cp) {
    out_cp[3 * qi] = hit ? bcp.x : 0.0f;
    out_cp[3 * qi + 1] = hit ? bcp.y : 0.0f;
    out_cp[3 * qi + 2] = hit ? bcp.z : 0.0f;
  }
  if (out_dist) out_dist[qi] = hit ? sqrtf(best) : -1.0f;
  if (out_d2) out_d2[qi] = hit ? best : -1.0f;
}

template <int STACK, bool POINTS, int K>
__device__ __forceinline__ void knn_group_body(int32_t *__restrict__ s_ref, float *__restrict__ s_lb,
                                               const BNode *__restrict__ nodes, const void *__restrict__ elems,
                                               const float *__restrict__ q, int64_t qi, bool live, int lg, int k,
                                               float best_init, float scale, int exclude_self,
                                               int32_t *__restrict__ out_idx, float *__restrict__ out_dist,
                                               float *__restrict__ out_d2, int32_t root) {
  const int tid = threadIdx.x;
  const int g = 1 << lg, gl = tid & (g - 1);
  float3 p = live ? load_q(q, qi) : f3(0, 0, 0);
  float m = fmaxf(scale, fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
  float sl = m * SLACK_REL;
  float3 qlo = f3(p.x - sl, p.y - sl, p.z - sl), qhi = f3(p.x + sl, p.y + sl, p.z + sl);
  float kd[K];
  int32_t ki[K];
#pragma unroll
  for (int j = 0; j < K; ++j) { kd[j] = best_init; ki[j] = 0x7fffffff; }
  const int32_t self = exclude_self ? (int32_t)qi : -1;
  const TriRec *__restrict__ tr = (const TriRec *)elems;
  const PointRec *__restrict__ pr = (const PointRec *)elems;
  // The lane's k-th best pair (slot k - 1): its admission threshold; the group's pruning bound is
  // the minimum of these over the lanes (each lane alone holds k pairs at or below its own).
  auto kth = [&](float &bd, int32_t &bi) {
    bd = kd[K - 1];
    bi = ki[K - 1];
#pragma unroll
    for (int j = 0; j < K - 1; ++j)
      if (j == k - 1) { bd = kd[j]; bi = ki[j]; }
  };
  auto bound = [&]() {
    float bd;
    int32_t bi;
    kth(bd, bi);
    return bd;
  };
  auto eval = [&](int32_t f, int32_t cnt) {
    for (int32_t kk = f; kk < f + cnt; ++kk) {
      float d;
      int32_t id;
      if (POINTS) {
        float4 e = pr[kk].p;
        id = __float_as_int(e.w);
        d = sqn3(sub3(f3(e.x, e.y, e.z), p));
      } else {
        const TriRec *t = tr + kk;
        id = trirec_id(t->pl);
        float3 cp;
        d = tri_closest(t->p0, t->p1, t->p2, t->pl, t->e0, t->e1, t->e2, p, cp);
      }
      float bd;
      int32_t bi;
      kth(bd, bi);
      if (id != self && better(d, id, bd, bi)) knn_insert<K>(kd, ki, d, id);
    }
  };
  int sp = 0;
  LaneStack<KN_REG, KN_LB> stk;
  int pth = live ? 0 : 1 << GRP_FREE;  // next route
  float gbound = best_init;
  int32_t node = -1;
  while (true) {
    if (node < 0 && sp > 0) {
      --sp;
      int32_t pr_ref;
      float pr_lb;
      stk.pop(sp, pr_ref, pr_lb, s_ref, s_lb, tid);
      if (pr_lb <= gbound) node = pr_ref;
    } else if (node < 0 && pth < (1 << GRP_FREE)) {
      float lb = 0;
      node = group_route(nodes, root, lg, gl, pth++, qlo, qhi, [&]() { return fminf(gbound, bound()); }, eval, lb);
    }
    if (node >= 0) {
      const BNode nd = nodes[node];
      float lb0 = box_lb(qlo, qhi, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y);
      float lb1 = box_lb(qlo, qhi, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w);
      int32_t r0 = nd.d.x, r1 = nd.d.y, c0 = nd.d.z, c1 = nd.d.w;
      if (lb1 < lb0) {
        float tl = lb0; lb0 = lb1; lb1 = tl;
        int32_t tr_ = r0; r0 = r1; r1 = tr_;
        int32_t tc = c0; c0 = c1; c1 = tc;
      }
      int32_t next = -1;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        float lb = c == 0 ? lb0 : lb1;
        int32_t r = c == 0 ? r0 : r1;
        int32_t cnt = c == 0 ? c0 : c1;
        if (!(lb <= fminf(gbound, bound()))) continue;
        if (r < 0) {
          eval(~r, cnt);
        } else if (next < 0) {
          next = r;
        } else if (sp < STACK) {  // never taken when depth < STACK (checked by every caller)
          stk.push(sp, r, lb, s_ref, s_lb, tid);
          ++sp;
        }
      }
      node = next;
    }
    gbound = group_min(bound(), g);
    if (!__any(node >= 0 || sp > 0 || pth < (1 << GRP_FREE))) break;
  }
  // Merge: the group's k smallest pairs are among the lanes' first k; each round takes the least
  // head and its lane drops it.
  int left = k;
  for (int j = 0; j < k; ++j) {
    const uint64_t head = left > 0 ? pair_key(kd[0], ki[0]) : ~(uint64_t)0;
    const uint64_t gkey = group_min_key(head, g);
    const bool hit = (uint32_t)gkey != 0x7fffffffu && gkey != ~(uint64_t)0;
    if (live && gl == 0) {
      const float d = __uint_as_float((uint32_t)(gkey >> 32));
      out_idx[qi * k + j] = hit ? (int32_t)(uint32_t)gkey : -1;
      if (out_dist) out_dist[qi * k + j] = hit ? sqrtf(d) : -1.0f;
      if (out_d2) out_d2[qi * k + j] = hit ? d : -1.0f;
    }
    if (hit && head == gkey) {
#pragma unroll
      for (int t = 0; t < K - 1; ++t) { kd[t] = kd[t + 1]; ki[t] = ki[t + 1]; }
      --left;
    }
  }
}

// Lanes per query for a sparse batch of n queries: the largest power of two up to 64 that keeps
// n * G within GRP_LANES lanes (about four waves per SIMD), as log2.
constexpr int64_t GRP_LANES = (int64_t)1 << 18;

template <int STACK, bool POINTS>
__global__ void __launch_bounds__(Q_BLOCK, STACK <= 32 ? 5 : 1) k_closest_group(const BNode *__restrict__ nodes,
                                                           const void *__restrict__ elems, const float *__restrict__ q,
                                                           const uint32_t *__restrict__ order, int64_t n, int lg,
                                                           uint32_t nb_real, float best_init, float scale,
                                                           int exclude_self, int32_t *__restrict__ out_idx,
                                                           float *__restrict__ out_cp, float *__restrict__ out_dist,
                                                           float *__restrict__ out_d2) {
  __shared__ int32_t s_ref[lane_lds_entries<STACK, CL_REG>() * Q_BLOCK];
  __shared__ float s_lb[CL_LB ? lane_lds_entries<STACK, CL_REG>() * Q_BLOCK : 1];
  const uint32_t lb_blk = xcd_remap(blockIdx.x, gridDim.x);
  if (lb_blk >= nb_real) return;
  const int64_t s = ((int64_t)lb_blk * Q_BLOCK + threadIdx.x) >> lg;
  const bool live = s < n;
  const int64_t qi = live ? (order ? (int64_t)order[s] : s) : 0;
  closest_group_body<STACK, POINTS>(s_ref, s_lb, nodes, elems, q, qi, live, lg, best_init, scale, exclude_self,
                                    out_idx, out_cp, out_dist, out_d2, 0);
}

template <int STACK, bool POINTS, int K>
__global__ void __launch_bounds__(Q_BLOCK) k_knn_group(const BNode *__restrict__ nodes, const void *__restrict__ elems,
                                                       const float *__restrict__ q, const uint32_t *__restrict__ order,
                                                       int64_t n, int lg, uint32_t nb_real, int k, float best_init,
                                                       float scale, int exclude_self, int32_t *__restrict__ out_idx,
                                                       float *__restrict__ out_dist, float *__restrict__ out_d2) {
  __shared__ int32_t s_ref[lane_lds_entries<STACK, KN_REG>() * Q_BLOCK];
  __shared__ float s_lb[KN_LB ? lane_lds_entries<STACK, KN_REG>() * Q_BLOCK : 1];
  const uint32_t lb_blk = xcd_remap(blockIdx.x, gridDim.x);
  if (lb_blk >= nb_real) return;
  const int64_t s = ((int64_t)lb_blk * Q_BLOCK + threadIdx.x) >> lg;
  const bool live = s < n;
  const int64_t qi = live ? (order ? (int64_t)order[s] : s) : 0;
  knn_group_body<STACK, POINTS, K>(s_ref, s_lb, nodes, elems, q, qi, live, lg, k, best_init, scale, exclude_self,
                                   out_idx, out_dist, out_d2, 0);
}

// Server form: one 64-lane block per wave of a segment; a segment of n queries at 2^lg lanes each
// spans ceil(n 2^lg / 64) waves (SegDesc::lanes_log2).
template <int STACK, bool POINTS>
__global__ void __launch_bounds__(Q_BLOCK, STACK <= 32 ? 5 : 1) k_closest_group_multi(const SegDesc *__restrict__ seg, int nseg,
                                                                 uint64_t total_waves, uint32_t nb_real) {
  __shared__ int32_t s_ref[lane_lds_entries<STACK, CL_REG>() * Q_BLOCK];
  __shared__ float s_lb[CL_LB ? lane_lds_entries<STACK, CL_REG>() * Q_BLOCK : 1];
  const uint32_t lb_blk = xcd_remap(blockIdx.x, gridDim.x);
  if (lb_blk >= nb_real || (uint64_t)lb_blk >= total_waves) return;
  const int si = uni(find_segment(seg, nseg, (int64_t)lb_blk));
  const SegDesc &d = seg[si];
  const int lg = d.lanes_log2;
  const int64_t s = (((int64_t)lb_blk - (int64_t)d.wave_begin) * Q_BLOCK + threadIdx.x) >> lg;
  const bool live = s < d.n;
  const int64_t qi = live ? (d.order ? (int64_t)d.order[s] : s) : 0;
  closest_group_body<STACK, POINTS>(s_ref, s_lb, readonly_ptr<BNode>(d.nodes), readonly_ptr<float>(d.elems),
                                    readonly_ptr<float>(d.q), qi, live, lg, d.best_init,
                                    d.scale, d.exclude_self, d.idx, d.cp, d.dist, d.d2, d.root);
}

template <int STACK, bool POINTS, int K>
__global__ void __launch_bounds__(Q_BLOCK) k_knn_group_multi(const SegDesc *__restrict__ seg, int nseg,
                                                             uint64_t total_waves, uint32_t nb_real) {
  __shared__ int32_t s_ref[lane_lds_entries<STACK, KN_REG>() * Q_BLOCK];
  __shared__ float s_lb[KN_LB ? lane_lds_entries<STACK, KN_REG>() * Q_BLOCK : 1];
  const uint32_t lb_blk = xcd_remap(blockIdx.x, gridDim.x);
  if (lb_blk >= nb_real || (uint64_t)lb_blk >= total_waves) return;
  const int si = uni(find_segment(seg, nseg, (int64_t)lb_blk));
  const SegDesc &d = seg[si];
  const int lg = d.lanes_log2;
  const int64_t s = (((int64_t)lb_blk - (int64_t)d.wave_begin) * Q_BLOCK + threadIdx.x) >> lg;
  const bool live = s < d.n;
  const int64_t qi = live ? (d.order ? (int64_t)d.order[s] : s) : 0;
  knn_group_body<STACK, POINTS, K>(s_ref, s_lb, readonly_ptr<BNode>(d.nodes), readonly_ptr<float>(d.elems),
                                   readonly_ptr<float>(d.q), qi, live, lg, d.k,
                                   d.best_init, d.scale, d.exclude_self, d.idx, d.dist, d.d2, d.root);
}

//--------------------------------------------------------------------------------------------------
// k nearest, wave-cooperative: the traversal of k_closest_wave with a per-lane sorted list of K
// (d^2, id) pairs in registers; a subtree or element is skipped when every lane's k-th best beats
// its lower bound. Element records are wave-uniform (scalar loads).
//--------------------------------------------------------------------------------------------------
template <bool POINTS, int K, bool LB = false>
__device__ __forceinline__ void knn_wave_body(const BNode *__restrict__ nodes, const void *__restrict__ elems,
                                              const float *__restrict__ q, const uint32_t *__restrict__ order,
                                              int64_t n, int64_t wave, bool wave_ok, int k, float best_init,
                                              float scale, int exclude_self, int32_t *__restrict__ out_idx,
                                              float *__restrict__ out_dist, float *__restrict__ out_d2,
                                              int32_t nelem, int32_t root = 0,
                                              const LeafBound *__restrict__ lbnd = nullptr) {
  const int lane = lane_id();
  const int64_t t = wave * WAVE + lane;
  const bool inrange = wave_ok && t < n;
  bool active = inrange;
  int64_t qi = 0;
  float3 p = f3(0, 0, 0);
  if (active) {
    qi = order ? (int64_t)order[t] : t;
    p = load_q(q, qi);
    active = !(isnan(p.x) || isnan(p.y) || isnan(p.z));
  }
  const uint64_t amask = __ballot(active);
  const int nact = __popcll(amask);
  float m = fmaxf(scale, fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
  float sl = m * SLACK_REL;
  float3 qlo = f3(p.x - sl, p.y - sl, p.z - sl), qhi = f3(p.x + sl, p.y + sl, p.z + sl);
  float kd[K];
  int32_t ki[K];
#pragma unroll
  for (int j = 0; j < K; ++j) { kd[j] = active ? best_init : -1.0f; ki[j] = 0x7fffffff; }
  // k-th best (slot k - 1): the pruning bound and admission threshold
  float kb = active ? best_init : -1.0f;
  int32_t kbi = 0x7fffffff;
  const int32_t self = exclude_self ? (int32_t)qi : -1;
  const TriRec *__restrict__ tr = (const TriRec *)elems;
  const PointRec *__restrict__ pr = (const PointRec *)elems;
  // Seed: the elements around the nearest leaf in leaf (Morton) order, which are spatially close;
  // they bound the k-th distance from the start and are skipped by the walk below. A window of
  // k + 1 can lie across a jump of the Morton curve (see knn_pts_body), so points take 4k + 1 and
  // triangles, whose tests cost more, 2k + 1.
  int32_t s0 = 0, s1 = 0;
  if (active && nelem > 0) {
    int32_t f, cnt;
    seed_leaf(nodes, root, qlo, qhi, f, cnt);
    const int32_t want = POINTS ? 4 * k + 1 : 2 * k + 1;
    const int32_t w = nelem < want ? nelem : want;
    s0 = f + cnt / 2 - w / 2;
    s0 = s0 < 0 ? 0 : (s0 > nelem - w ? nelem - w : s0);
    s1 = s0 + w;
    for (int32_t e = s0; e < s1; ++e) {
      float d;
      int32_t id;
      if (POINTS) {
        const float4 pe = pr[e].p;
        id = __float_as_int(pe.w);
        d = sqn3(sub3(f3(pe.x, pe.y, pe.z), p));
      } else {
        const TriRec *tk = tr + e;
        float3 cp;
        id = trirec_id(tk->pl);
        d = tri_closest(tk->p0, tk->p1, tk->p2, tk->pl, tk->e0, tk->e1, tk->e2, p, cp);
      }
      if (id != self && better(d, id, kb, kbi)) {
        knn_insert<K>(kd, ki, d, id);
        kb = kd[K - 1];
        kbi = ki[K - 1];
#pragma unroll
        for (int j = 0; j < K - 1; ++j)
          if (j == k - 1) { kb = kd[j]; kbi = ki[j]; }
      }
    }
  }
  // triangles: the reject threshold from the k-th distance, refreshed after every insertion
  float rej2 = (!POINTS && active) ? rej_sq(kb, sl) : 0.0f;
  int32_t s_ref = 0;
  float s_lx = 0, s_ly = 0, s_lz = 0, s_hx = 0, s_hy = 0, s_hz = 0;
  int32_t sp = 0;
  int32_t node = root;
  // The node to enter next is known from the ballots before the leaf children are evaluated, so its
  // record is loaded first and arrives while the wave works through the leaf's points.
  BNode nd = nodes[uni(node)];
  while (amask != 0) {
    node = uni(node);
    const float a0 = unif(nd.a.x), a1 = unif(nd.a.y), a2 = unif(nd.a.z), a3 = unif(nd.a.w);
    const float b0 = unif(nd.b.x), b1 = unif(nd.b.y), b2 = unif(nd.b.z), b3 = unif(nd.b.w);
    const float c0 = unif(nd.c.x), c1 = unif(nd.c.y), c2 = unif(nd.c.z), c3 = unif(nd.c.w);
    const int32_t r0 = uni(nd.d.x), r1 = uni(nd.d.y), n0 = uni(nd.d.z), n1 = uni(nd.d.w);
    float lb0 = box_lb(qlo, qhi, a0, a1, a2, a3, b0, b1);
    float lb1 = box_lb(qlo, qhi, b2, b3, c0, c1, c2, c3);
    uint64_t m0 = __ballot(lb0 <= kb), m1 = __ballot(lb1 <= kb);
    bool first1 = 2 * __popcll(__ballot(lb1 < lb0) & amask) > nact;
    // the first admitted internal child in visiting order is the one the walk enters next
    int32_t next = -1;
    {
      const int cf = first1 ? 1 : 0;
      const uint64_t mf = cf == 0 ? m0 : m1, ms = cf == 0 ? m1 : m0;
      const int32_t rf = cf == 0 ? r0 : r1, rs = cf == 0 ? r1 : r0;
      next = (mf != 0 && rf >= 0) ? rf : ((ms != 0 && rs >= 0) ? rs : -1);
    }
    BNode nd_next;
    if (next >= 0) nd_next = nodes[next];
    bool entered = false;
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
      const int c = first1 ? 1 - cc : cc;
      const uint64_t mc = c == 0 ? m0 : m1;
      if (mc == 0) continue;
      const int32_t r = c == 0 ? r0 : r1;
      const int32_t cnt = c == 0 ? n0 : n1;
      if (r < 0) {
        const int32_t f = ~r;
        // triangles: leaf-level reject against each lane's k-th distance (see closest_cmp_body)
        uint64_t ml = mc;
        if (!POINTS && LB && lbnd) {
          const LeafBound *L = lbnd + f;
          const float lx = unif(L->a.x), ly = unif(L->a.y), lz = unif(L->a.z), lr = unif(L->a.w);
          const float nx = unif(L->b.x), ny = unif(L->b.y), nz = unif(L->b.z), lt = unif(L->b.w);
          const float dx = p.x - lx, dy = p.y - ly, dz = p.z - lz;
          const float dc2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
          const float sd = fmaf(dx, nx, fmaf(dy, ny, dz * nz));
          ml &= ballot(!(leaf_lb2(dc2, sd, lt, lr, sl) > rej2));
          if (ml == 0) continue;
        }
        for (int32_t e = f; e < f + cnt; ++e) {
          float d;
          int32_t id;
          if (POINTS) {
            const float4 pe = pr[e].p;
            const float3 pt = f3(unif(pe.x), unif(pe.y), unif(pe.z));
            id = uni(__float_as_int(pe.w));
            d = sqn3(sub3(pt, p));
          } else {
            const TriRec *tk = tr + e;
            // conservative reject of the whole wave before Thea's exact arithmetic: the plane and
            // disc bound against each admitting lane's k-th distance
            {
              const float4 sbv = make_float4(unif(tk->sb.x), unif(tk->sb.y), unif(tk->sb.z), unif(tk->sb.w));
              const float dx = p.x - sbv.x, dy = p.y - sbv.y, dz = p.z - sbv.z;
              const float dc2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
              const float sd = fmaf(dx, unif(tk->p0.w), fmaf(dy, unif(tk->p1.w), dz * unif(tk->p2.w)));
              if ((ballot(!(tri_lb2(dc2, sd, sbv.w, sl) > rej2)) & ml) == 0) continue;
            }
            float3 cp;
            id = uni(trirec_id(tk->pl));
            d = tri_closest(make_float4(unif(tk->p0.x), unif(tk->p0.y), unif(tk->p0.z), unif(tk->p0.w)),
                            make_float4(unif(tk->p1.x), unif(tk->p1.y), unif(tk->p1.z), unif(tk->p1.w)),
                            make_float4(unif(tk->p2.x), unif(tk->p2.y), unif(tk->p2.z), unif(tk->p2.w)),
                            make_float4(unif(tk->pl.x), unif(tk->pl.y), unif(tk->pl.z), unif(tk->pl.w)),
                            make_float4(unif(tk->e0.x), unif(tk->e0.y), unif(tk->e0.z), unif(tk->e0.w)),
                            make_float4(unif(tk->e1.x), unif(tk->e1.y), unif(tk->e1.z), unif(tk->e1.w)),
                            make_float4(unif(tk->e2.x), unif(tk->e2.y), unif(tk->e2.z), unif(tk->e2.w)), p, cp);
          }
          const bool take = active && id != self && (e < s0 || e >= s1) && better(d, id, kb, kbi);
          if (__ballot(take)) {
            if (take) {
              knn_insert<K>(kd, ki, d, id);
              kb = kd[K - 1];
              kbi = ki[K - 1];
#pragma unroll
              for (int j = 0; j < K - 1; ++j)
                if (j == k - 1) { kb = kd[j]; kbi = ki[j]; }
              if (!POINTS) rej2 = rej_sq(kb, sl);
            }
          }
        }
      } else if (!entered && r == next) {
        entered = true;
      } else if (sp < W_STACK) {  // never full when depth < 64 (checked by every caller)
        if (lane == sp) {
          s_ref = r;
          if (c == 0) { s_lx = a0; s_ly = a1; s_lz = a2; s_hx = a3; s_hy = b0; s_hz = b1; }
          else { s_lx = b2; s_ly = b3; s_lz = c0; s_hx = c1; s_hy = c2; s_hz = c3; }
        }
        ++sp;
      }
    }
    if (next >= 0) { node = next; nd = nd_next; continue; }
    bool found = false;
    while (sp > 0) {
      --sp;
      const float lx = lane_readf(s_lx, sp), ly = lane_readf(s_ly, sp), lz = lane_readf(s_lz, sp);
      const float hx = lane_readf(s_hx, sp), hy = lane_readf(s_hy, sp), hz = lane_readf(s_hz, sp);
      if (__ballot(box_lb(qlo, qhi, lx, ly, lz, hx, hy, hz) <= kb)) {
        node = lane_read(s_ref, sp);
        found = true;
        break;
      }
    }
    if (!found) break;
    nd = nodes[uni(node)];
  }
  if (!inrange) return;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    if (j < k) {
      bool hit = active && ki[j] != 0x7fffffff;
      out_idx[qi * k + j] = hit ? ki[j] : -1;
      if (out_dist) out_dist[qi * k + j] = hit ? sqrtf(kd[j]) : -1.0f;
      if (out_d2) out_d2[qi * k + j] = hit ? kd[j] : -1.0f;
    }
  }
}

//--------------------------------------------------------------------------------------------------
// k nearest over a point BVH, wave-cooperative, with batched leaves and deferred insertion.
//
// Candidates are 64-bit keys (float bits of d^2 << 32 | index). d^2 is never negative, so the
// unsigned order of the keys is the lexicographic (d^2, index) order of the results; a lane's list
// is K ascending keys and its admission threshold the k-th. A node's admitted leaves (at most two,
// 16 points each) are read with one vector load, lane j holding point j of the batch, which is
// issued together with the scalar load of the next node. Every lane then tests every point of the
// batch against its k-th key as the batch began (pass 1, a bit per admitted point), and afterwards
// inserts only its own admitted points, gathering their coordinates from the loading lanes (pass 2).
// The wave runs the insertion as often as its busiest lane needs it, instead of once for every
// point some lane admits. A stale threshold admits a superset of what immediate insertion would;
// pass 2 re-tests each against the current k-th key, and the list is the k smallest keys of all
// admitted points in any order, so results are those of knn_wave_body.
//--------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t knn_key(float d, int32_t id) {
  return ((uint64_t)(uint32_t)__float_as_int(d) << 32) | (uint32_t)id;
}

template <int K> __device__ __forceinline__ void key_insert(uint64_t (&kk)[K], uint64_t key) {
  bool b[K];
#pragma unroll
  for (int j = 0; j < K; ++j) b[j] = key < kk[j];
#pragma unroll
  for (int j = K - 1; j > 0; --j) kk[j] = b[j] ? (b[j - 1] ? kk[j - 1] : key) : kk[j];
  kk[0] = b[0] ? key : kk[0];
}

// Key of slot k - 1 (1 <= k <= K); k is uniform.
template <int K> __device__ __forceinline__ uint64_t key_kth(const uint64_t (&kk)[K], int k) {
  if (k == K) return kk[K - 1];
  uint64_t r = kk[K - 1];
#pragma unroll
  for (int j = 0; j < K - 1; ++j)
    if (j == k - 1) r = kk[j];
  return r;
}

// STATS: per wave, wstats[4 w ..] = elapsed wall clock (10 ns ticks), nodes entered, points tested,
// insertion rounds (diagnostics).
template <int K, bool STATS = false>
__device__ __forceinline__ void knn_pts_body(const BNode *__restrict__ nodes, const PointRec *__restrict__ pr,
                                             const float *__restrict__ q, const uint32_t *__restrict__ order,
                                             int64_t n, int64_t wave, bool wave_ok, int k, float best_init, float scale,
                                             int exclude_self, int32_t *__restrict__ out_idx,
                                             float *__restrict__ out_dist, float *__restrict__ out_d2, int32_t nelem,
                                             int32_t root = 0, int64_t *__restrict__ wstats = nullptr) {
  const uint64_t t_start = STATS ? wall_clock64() : 0;
  int64_t st_nodes = 0, st_pts = 0, st_rounds = 0;
  const int lane = lane_id();
  const int64_t t = wave * WAVE + lane;
  const bool inrange = wave_ok && t < n;
  bool active = inrange;
  int64_t qi = 0;
  float3 p = f3(0, 0, 0);
  if (active) {
    qi = order ? (int64_t)order[t] : t;
    p = load_q(q, qi);
    active = !(isnan(p.x) || isnan(p.y) || isnan(p.z));
  }
  uint64_t amask = __ballot(active);
  const int nact = __popcll(amask);
  const float m = fmaxf(scale, fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
  const float sl = m * SLACK_REL;
  const float3 qlo = f3(p.x - sl, p.y - sl, p.z - sl), qhi = f3(p.x + sl, p.y + sl, p.z + sl);
  // Empty slots hold (best_init, INT_MAX): a point is admitted iff d^2 < bound^2, as for the float
  // lists. Inactive lanes and a negative bound admit nothing (every key is >= 0).
  const uint64_t init = (active && best_init >= 0.0f) ? knn_key(best_init, 0x7fffffff) : 0ull;
  uint64_t kk[K];
#pragma unroll
  for (int j = 0; j < K; ++j) kk[j] = init;
  uint64_t kth = init;
  const int32_t self = exclude_self ? (int32_t)qi : -1;
  // Seed: the 4k + 1 points around the nearest leaf in leaf (Morton) order (skipped by the walk).
  // k + 1 points can lie across a jump of the Morton curve for a query near one, and one such loose
  // bound makes the whole wave walk towards the far side: on the 1M C2 samples the slowest waves
  // started from k-th bounds up to 200x the true 16-NN distance (2.5 against 0.0146) and took 15x
  // the median wave's time; 4k + 1 points keep every bound within a few times the true one.
  int32_t s0 = 0, s1 = 0;
  if (active && nelem > 0) {
    int32_t f, cnt;
    seed_leaf(nodes, root, qlo, qhi, f, cnt);
    const int32_t w = nelem < 4 * k + 1 ? nelem : 4 * k + 1;
    s0 = f + cnt / 2 - w / 2;
    s0 = s0 < 0 ? 0 : (s0 > nelem - w ? nelem - w : s0);
    s1 = s0 + w;
    for (int32_t e = s0; e < s1; ++e) {
      const float4 pe = pr[e].p;
      const int32_t id = __float_as_int(pe.w);
      const uint64_t key = knn_key(sqn3(sub3(f3(pe.x, pe.y, pe.z), p)), id);
      if (id != self && key < kth) {
        key_insert<K>(kk, key);
        kth = key_kth<K>(kk, k);
      }
    }
  }
  int32_t s_ref = 0;
  float s_lx = 0, s_ly = 0, s_lz = 0, s_hx = 0, s_hy = 0, s_hz = 0;
  int32_t sp = 0;
  int32_t node = root;
  BNode nd = nodes[uni(node)];
  while (amask != 0) {
    node = uni(node);
    if (STATS) ++st_nodes;
    const float a0 = unif(nd.a.x), a1 = unif(nd.a.y), a2 = unif(nd.a.z), a3 = unif(nd.a.w);
    const float b0 = unif(nd.b.x), b1 = unif(nd.b.y), b2 = unif(nd.b.z), b3 = unif(nd.b.w);
    const float c0 = unif(nd.c.x), c1 = unif(nd.c.y), c2 = unif(nd.c.z), c3 = unif(nd.c.w);
    const int32_t r0 = uni(nd.d.x), r1 = uni(nd.d.y), n0 = uni(nd.d.z), n1 = uni(nd.d.w);
    // the k-th distance bounds the search ball (pruning compares floats, as in knn_wave_body)
    const float kb = active ? __int_as_float((int32_t)(kth >> 32)) : -1.0f;
    const float lb0 = box_lb(qlo, qhi, a0, a1, a2, a3, b0, b1);
    const float lb1 = box_lb(qlo, qhi, b2, b3, c0, c1, c2, c3);
    const uint64_t m0 = __ballot(lb0 <= kb), m1 = __ballot(lb1 <= kb);
    const bool first1 = 2 * __popcll(__ballot(lb1 < lb0) & amask) > nact;
    int32_t next = -1;
    {
      const int cf = first1 ? 1 : 0;
      const uint64_t mf = cf == 0 ? m0 : m1, ms = cf == 0 ? m1 : m0;
      const int32_t rf = cf == 0 ? r0 : r1, rs = cf == 0 ? r1 : r0;
      next = (mf != 0 && rf >= 0) ? rf : ((ms != 0 && rs >= 0) ? rs : -1);
    }
    // admitted leaves of this node: [lf0, lf0 + lc0) then [lf1, lf1 + lc1)
    int32_t lf0 = 0, lc0 = 0, lf1 = 0, lc1 = 0;
    if (m0 != 0 && r0 < 0) { lf0 = ~r0; lc0 = n0; }
    if (m1 != 0 && r1 < 0) {
      if (lc0 == 0) { lf0 = ~r1; lc0 = n1; }
      else { lf1 = ~r1; lc1 = n1; }
    }
    const int32_t total = lc0 + lc1;  // <= 2 x 16 (leaf_size <= 16)
    float4 pe = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (lane < total) pe = pr[lane < lc0 ? lf0 + lane : lf1 + (lane - lc0)].p;
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
      } else if (sp < W_STACK) {  // never full when depth < 64 (checked by every caller)
        if (lane == sp) {
          s_ref = r;
          if (c == 0) { s_lx = a0; s_ly = a1; s_lz = a2; s_hx = a3; s_hy = b0; s_hz = b1; }
          else { s_lx = b2; s_ly = b3; s_lz = c0; s_hx = c1; s_hy = c2; s_hz = c3; }
        }
        ++sp;
      }
    }
    if (total > 0) {
      if (STATS) st_pts += total;
      // pass 1: bit j of `take` = point j of the batch beats this lane's k-th key
      uint32_t take = 0;
      for (int32_t j = 0; j < total; ++j) {
        const float x = lane_readf(pe.x, j), y = lane_readf(pe.y, j), z = lane_readf(pe.z, j);
        const int32_t id = lane_read(__float_as_int(pe.w), j);
        const int32_t e = j < lc0 ? lf0 + j : lf1 + (j - lc0);
        const uint64_t key = knn_key(sqn3(sub3(f3(x, y, z), p)), id);
        const bool tk = id != self && (e < s0 || e >= s1) && key < kth;
        take |= tk ? (1u << j) : 0u;
      }
      // pass 2: each lane inserts its own admitted points
      while (__ballot(take != 0)) {
        if (STATS) ++st_rounds;
        const bool has = take != 0;
        const int j = has ? __builtin_ctz(take) : 0;
        take &= take - 1u;
        const float x = __shfl(pe.x, j), y = __shfl(pe.y, j), z = __shfl(pe.z, j);
        const int32_t id = __shfl(__float_as_int(pe.w), j);
        const uint64_t key = knn_key(sqn3(sub3(f3(x, y, z), p)), id);
        if (has && key < kth) {
          key_insert<K>(kk, key);
          kth = key_kth<K>(kk, k);
        }
      }
    }
    if (next >= 0) { node = next; nd = nd_next; continue; }
    bool found = false;
    const float kbn = active ? __int_as_float((int32_t)(kth >> 32)) : -1.0f;
    while (sp > 0) {
      --sp;
      const float lx = lane_readf(s_lx, sp), ly = lane_readf(s_ly, sp), lz = lane_readf(s_lz, sp);
      const float hx = lane_readf(s_hx, sp), hy = lane_readf(s_hy, sp), hz = lane_readf(s_hz, sp);
      if (__ballot(box_lb(qlo, qhi, lx, ly, lz, hx, hy, hz) <= kbn)) {
        node = lane_read(s_ref, sp);
        found = true;
        break;
      }
    }
    if (!found) break;
    nd = nodes[uni(node)];
  }
  if (STATS && lane == 0 && wave_ok) {
    wstats[4 * wave] = (int64_t)(wall_clock64() - t_start);
    wstats[4 * wave + 1] = st_nodes;
    wstats[4 * wave + 2] = st_pts;
    wstats[4 * wave + 3] = st_rounds;
  }
  if (!inrange) return;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    if (j < k) {
      const int32_t id = (int32_t)(uint32_t)kk[j];
      const float d = __int_as_float((int32_t)(kk[j] >> 32));
      const bool hit = active && id != 0x7fffffff;
      out_idx[qi * k + j] = hit ? id : -1;
      if (out_dist) out_dist[qi * k + j] = hit ? sqrtf(d) : -1.0f;
      if (out_d2) out_d2[qi * k + j] = hit ? d : -1.0f;
    }
  }
}

template <bool POINTS, int K>
__global__ void __launch_bounds__(W_BLOCK) k_knn_wave(const BNode *__restrict__ nodes, const void *__restrict__ elems,
                                                      const float *__restrict__ q, const uint32_t *__restrict__ order,
                                                      int64_t n, uint32_t nb_real, int k, float best_init, float scale,
                                                      int exclude_self, int32_t *__restrict__ out_idx,
                                                      float *__restrict__ out_dist, float *__restrict__ out_d2,
                                                      int32_t nelem, const LeafBound *__restrict__ lbnd = nullptr) {
  const uint32_t lb_blk = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t wave = (int64_t)lb_blk * W_WAVES + (threadIdx.x / WAVE);
  knn_wave_body<POINTS, K, true>(nodes, elems, q, order, n, wave, lb_blk < nb_real, k, best_init, scale,
                                 exclude_self, out_idx, out_dist, out_d2, nelem, 0, lbnd);
}

template <int K, bool STATS = false>
__global__ void __launch_bounds__(W_BLOCK) k_knn_pts(const BNode *__restrict__ nodes, const PointRec *__restrict__ pr,
                                                     const float *__restrict__ q, const uint32_t *__restrict__ order,
                                                     int64_t n, uint32_t nb_real, int k, float best_init, float scale,
                                                     int exclude_self, int32_t *__restrict__ out_idx,
                                                     float *__restrict__ out_dist, float *__restrict__ out_d2,
                                                     int32_t nelem, int64_t *__restrict__ wstats = nullptr) {
  const uint32_t lb_blk = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t wave = (int64_t)lb_blk * W_WAVES + (threadIdx.x / WAVE);
  knn_pts_body<K, STATS>(nodes, pr, q, order, n, wave, lb_blk < nb_real, k, best_init, scale, exclude_self, out_idx,
                         out_dist, out_d2, nelem, 0, wstats);
}

template <bool POINTS, int K>
__global__ void __launch_bounds__(W_BLOCK) k_knn_multi(const SegDesc *__restrict__ seg, int nseg, uint64_t total_waves,
                                                       uint32_t nb_real) {
  const uint32_t lb_blk = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t gw = (int64_t)lb_blk * W_WAVES + (threadIdx.x / WAVE);
  if (lb_blk >= nb_real || gw >= (int64_t)total_waves) return;
  const int si = uni(find_segment(seg, nseg, gw));
  const SegDesc &d = seg[si];
  knn_wave_body<POINTS, K>(readonly_ptr<BNode>(d.nodes), readonly_ptr<float>(d.elems), readonly_ptr<float>(d.q),
                           readonly_ptr<uint32_t>(d.order), d.n, gw - (int64_t)d.wave_begin, true, d.k,
                           d.best_init, d.scale, d.exclude_self, d.idx, d.dist, d.d2, d.nelem, d.root);
}

//--------------------------------------------------------------------------------------------------
// Rays
//--------------------------------------------------------------------------------------------------
// Slab entry/exit of the box widened by `w` (per axis), using the reciprocal direction.
__device__ __forceinline__ float2 slab(float3 o, float3 inv, float w, float lx, float ly, float lz, float hx, float hy,
                                       float hz) {
  float tx0 = (lx - w - o.x) * inv.x, tx1 = (hx + w - o.x) * inv.x;
  float ty0 = (ly - w - o.y) * inv.y, ty1 = (hy + w - o.y) * inv.y;
  float tz0 = (lz - w - o.z) * inv.z, tz1 = (hz + w - o.z) * inv.z;
  float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
  float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
  return make_float2(tn, tf);
}

// ANY = true: stop at the first accepted hit (any-hit query).
// REG > 0 (with TS = false): the top REG stack entries live in registers as a shift register
// (reg[0] on top) and only deeper entries in LDS (STACK - REG per lane), so a block needs less LDS
// and more waves fit on a CU.
template <int STACK, bool ANY, bool STATS = false, bool TS = true, int REG = 0>
__device__ __forceinline__ void ray_body(int32_t *__restrict__ s_ref, float *__restrict__ s_t,
                                         const BNode *__restrict__ nodes, const RayTri *__restrict__ rt,
                                         const TriRec *__restrict__ tr,
                                         const float *__restrict__ orig, const float *__restrict__ dirs, int64_t ri,
                                         float best_init, float max_time, float scale, int32_t *__restrict__ out_idx,
                                         float *__restrict__ out_t, float *__restrict__ out_n,
                                         uint8_t *__restrict__ out_any, int32_t root = 0,
                                         int32_t *__restrict__ counts = nullptr) {
  int32_t c_nodes = 0, c_tris = 0, c_maxsp = 0, c_pops = 0;
  const int tid = threadIdx.x;
  float3 o = load_q(orig, ri), d = load_q(dirs, ri);
  float3 inv = f3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  float m = fmaxf(scale, fmaxf(fabsf(o.x), fmaxf(fabsf(o.y), fabsf(o.z))));
  float w = m * SLACK_REL;
  float best = best_init;
  int32_t bidx = 0x7fffffff;
  int32_t bslot = -1;
  int sp = 0;
  int32_t reg[REG > 0 ? REG : 1];
  int32_t node = root;
  // A box is entered when its (widened) entry time is within this relative margin of the best.
  const float REL = 1.0f + 1.0f / 4096.0f;
  bool done = false;
  while (!done) {
    if (STATS) c_nodes++;
    const BNode nd = nodes[node];
    float2 s0 = slab(o, inv, w, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y);
    float2 s1 = slab(o, inv, w, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w);
    // A child is visited only if the ray meets its widened box (h) and enters it no later than the
    // best hit so far; an unbounded best must not admit a box the ray misses.
    bool h0 = s0.x <= s0.y * REL, h1 = s1.x <= s1.y * REL;
    float t0 = s0.x, t1 = s1.x;
    int32_t r0 = nd.d.x, r1 = nd.d.y, c0 = nd.d.z, c1 = nd.d.w;
    if (h1 && (!h0 || t1 < t0)) {
      float tt = t0; t0 = t1; t1 = tt;
      int32_t a = r0; r0 = r1; r1 = a;
      int32_t b = c0; c0 = c1; c1 = b;
      bool hh = h0; h0 = h1; h1 = hh;
    }
    int32_t next = -1;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      float tn = c == 0 ? t0 : t1;
      int32_t r = c == 0 ? r0 : r1;
      int32_t cnt = c == 0 ? c0 : c1;
      if (!((c == 0 ? h0 : h1) && tn <= best * REL)) continue;
      if (r < 0) {
        int32_t f = ~r;
        for (int32_t k = f; k < f + cnt; ++k) {
          const RayTri *t = rt + k;
          if (STATS) c_tris++;
          const float4 a = t->a, b = t->b, cc = t->c;
          float th = tri_ray_e(f3(a.x, a.y, a.z), f3(b.x, b.y, b.z), f3(cc.x, cc.y, cc.z), o, d, max_time);
          if (th >= 0) {
            int32_t id = __float_as_int(a.w);
            if (better(th, id, best, bidx)) {
              best = th;
              bidx = id;
              bslot = k;
              if (ANY) { done = true; break; }
            }
          }
        }
        if (ANY && done) break;
      } else if (next < 0) {
        next = r;
      } else if (sp < STACK) {  // never taken when depth < STACK (checked by every caller)
        if (REG > 0 && !TS) {
          if (sp >= REG) s_ref[(sp - REG) * Q_BLOCK + tid] = reg[REG - 1];
#pragma unroll
          for (int j = REG - 1; j > 0; --j) reg[j] = reg[j - 1];
          reg[0] = r;
        } else {
          s_ref[sp * Q_BLOCK + tid] = r;
          if (TS) s_t[sp * Q_BLOCK + tid] = tn;
        }
        ++sp;
        if (STATS) c_maxsp = max(c_maxsp, sp);
      }
    }
    if (done) break;
    if (next >= 0) { node = next; continue; }
    bool found = false;
    while (sp > 0) {
      --sp;
      if (STATS) c_pops++;
      // TS = false keeps only node refs (half the LDS, more waves per CU); a popped node is then
      // entered and its children are tested against the current best instead.
      if (REG > 0 && !TS) {
        node = reg[0];
#pragma unroll
        for (int j = 0; j < REG - 1; ++j) reg[j] = reg[j + 1];
        if (sp >= REG) reg[REG - 1] = s_ref[(sp - REG) * Q_BLOCK + tid];
        found = true;
        break;
      }
      if (!TS || s_t[sp * Q_BLOCK + tid] <= best * REL) { node = s_ref[sp * Q_BLOCK + tid]; found = true; break; }
    }
    if (!found) break;
  }
  if (STATS) {
    counts[4 * ri] = c_nodes;
    counts[4 * ri + 1] = c_tris;
    counts[4 * ri + 2] = c_maxsp;
    counts[4 * ri + 3] = c_pops;
  }
  bool hit = bidx != 0x7fffffff;
  if (ANY) {
    out_any[ri] = hit ? 1 : 0;
    return;
  }
  out_idx[ri] = hit ? bidx : -1;
  if (out_t) out_t[ri] = hit ? best : -1.0f;
  if (out_n) {
    float nx = 0, ny = 0, nz = 0;
    if (hit) { nx = tr[bslot].p0.w; ny = tr[bslot].p1.w; nz = tr[bslot].p2.w; }
    out_n[3 * ri] = nx; out_n[3 * ri + 1] = ny; out_n[3 * ri + 2] = nz;
  }
}

template <int STACK, bool ANY, bool TS = true, int REG = 0>
__global__ void __launch_bounds__(Q_BLOCK) k_ray(const BNode *__restrict__ nodes, const RayTri *__restrict__ rt,
                                                 const TriRec *__restrict__ tr,
                                                 const float *__restrict__ orig, const float *__restrict__ dirs,
                                                 const uint32_t *__restrict__ order, int64_t n, uint32_t nb_real,
                                                 float best_init, float max_time, float scale,
                                                 int32_t *__restrict__ out_idx, float *__restrict__ out_t,
                                                 float *__restrict__ out_n, uint8_t *__restrict__ out_any) {
  __shared__ int32_t s_ref[(STACK - (TS ? 0 : REG)) * Q_BLOCK];
  __shared__ float s_t[TS ? STACK * Q_BLOCK : 1];
  bool active;
  int64_t ri = query_slot(nb_real, n, order, active);
  if (!active) return;
  ray_body<STACK, ANY, false, TS, TS ? 0 : REG>(s_ref, s_t, nodes, rt, tr, orig, dirs, ri, best_init, max_time, scale,
                                                out_idx, out_t, out_n, out_any);
}

// Per-ray traversal counts (diagnostics): nodes visited, triangles tested, deepest stack, pops.
template <int STACK>
__global__ void __launch_bounds__(Q_BLOCK) k_ray_stats(const BNode *__restrict__ nodes, const RayTri *__restrict__ rt,
                                                       const TriRec *__restrict__ tr,
                                                       const float *__restrict__ orig, const float *__restrict__ dirs,
                                                       int64_t n, float scale, int32_t *__restrict__ out_idx,
                                                       int32_t *__restrict__ counts) {
  __shared__ int32_t s_ref[STACK * Q_BLOCK];
  __shared__ float s_t[STACK * Q_BLOCK];
  const int64_t ri = (int64_t)blockIdx.x * Q_BLOCK + threadIdx.x;
  if (ri >= n) return;
  ray_body<STACK, false, true>(s_ref, s_t, nodes, rt, tr, orig, dirs, ri, INFINITY, -1.0f, scale, out_idx, nullptr,
                               nullptr, nullptr, 0, counts);
}

// Many ray batches in one launch; one 64-lane block per wave of a segment (SEG_WAVE_QUERIES rays).
template <int STACK>
__global__ void __launch_bounds__(Q_BLOCK) k_ray_multi(const SegDesc *__restrict__ seg, int nseg, uint64_t total_waves,
                                                       uint32_t nb_real) {
  __shared__ int32_t s_ref[lane_lds_entries<STACK, LANE_REG>() * Q_BLOCK];
  __shared__ float s_t[1];
  const uint32_t lb_blk = xcd_remap(blockIdx.x, gridDim.x);
  if (lb_blk >= nb_real || (uint64_t)lb_blk >= total_waves) return;
  const int si = uni(find_segment(seg, nseg, (int64_t)lb_blk));
  const SegDesc &d = seg[si];
  const int64_t t = ((int64_t)lb_blk - (int64_t)d.wave_begin) * Q_BLOCK + threadIdx.x;
  if (t >= d.n) return;
  const int64_t ri = d.order ? (int64_t)d.order[t] : t;
  ray_body<STACK, false, false, false, LANE_REG>(s_ref, s_t, readonly_ptr<BNode>(d.nodes), readonly_ptr<RayTri>(d.rays),
                                                 readonly_ptr<TriRec>(d.elems), readonly_ptr<float>(d.q),
                                                 readonly_ptr<float>(d.q2), ri, d.best_init, d.max_time, d.scale, d.idx, d.dist, d.cp,
                                                 nullptr, d.root);
}

// Per-lane state of a ray (one ray per lane) and its set-up and output.
struct RayLane {
  float3 o, d, inv;
  float3 olo, ohi;  // o + w and o - w: the widened slab planes are (lo - olo) * inv, (hi - ohi) * inv
  float best;
  int32_t bidx, bslot;
  int64_t ri;
};

__device__ __forceinline__ void ray_lane_init(RayLane &L, const float *__restrict__ orig,
                                              const float *__restrict__ dirs, int64_t ri, float best_init,
                                              float scale) {
  L.ri = ri;
  L.o = load_q(orig, ri);
  L.d = load_q(dirs, ri);
  L.inv = f3(1.0f / L.d.x, 1.0f / L.d.y, 1.0f / L.d.z);
  const float m = fmaxf(scale, fmaxf(fabsf(L.o.x), fmaxf(fabsf(L.o.y), fabsf(L.o.z))));
  const float w = m * SLACK_REL;
  L.olo = f3(L.o.x + w, L.o.y + w, L.o.z + w);
  L.ohi = f3(L.o.x - w, L.o.y - w, L.o.z - w);
  L.best = best_init;
  L.bidx = 0x7fffffff;
  L.bslot = -1;
}

template <bool ANY>
__device__ __forceinline__ void ray_lane_finish(const RayLane &L, const TriRec *__restrict__ tr,
                                                int32_t *__restrict__ out_idx, float *__restrict__ out_t,
                                                float *__restrict__ out_n, uint8_t *__restrict__ out_any) {
  const bool hit = L.bidx != 0x7fffffff;
  if (ANY) {
    out_any[L.ri] = hit ? 1 : 0;
    return;
  }
  out_idx[L.ri] = hit ? L.bidx : -1;
  if (out_t) out_t[L.ri] = hit ? L.best : -1.0f;
  if (out_n) {
    float nx = 0, ny = 0, nz = 0;
    if (hit) { nx = tr[L.bslot].p0.w; ny = tr[L.bslot].p1.w; nz = tr[L.bslot].p2.w; }
    out_n[3 * L.ri] = nx; out_n[3 * L.ri + 1] = ny; out_n[3 * L.ri + 2] = nz;
  }
}

// Leaf end flag of a RayTri (int bits of b.w).
__device__ __forceinline__ bool raytri_last(const float4 &b) { return __float_as_int(b.w) != 0; }

// Node reference that ends a lane's traversal (~RAY_DONE would be leaf INT32_MAX: never a real one).
constexpr int32_t RAY_DONE = INT32_MIN;

// Rays over the 4-wide quantized tree (QNode4), one per lane, in while-while form (Aila and Laine,
// "Understanding the efficiency of ray traversal on GPUs", HPG 2009): a lane walks inner nodes,
// postponing the first leaf it meets, and the wave runs its leaf tests together once every lane
// still walking holds one. Leaves are node references like inner nodes (on the stack too), and a
// leaf's triangles run up to the record flagged as its last (RayTri::b.w). Results do not depend on
// the visiting order: a ray's answer is the least (t, index) over the triangles it tests, and a box
// is skipped only when it cannot hold that.
// A node's four child boxes are decoded only on the ray's near and far planes per axis (the
// direction's sign picks qlo or qhi). Fast form, when every direction component of the wave's rays
// has |1/d| <= 2^40 and the scene is below 2^60: t = fmaf(q, 2^e / d, (origin - o') / d), whose
// error is below 2^-19 of max(scale, |o|), inside the 2^-16 widening. Otherwise the wave decodes
// each plane with the build's own fmaf and intersects it as k_ray does, so rays parallel to an axis
// keep IEEE infinities. The traversal stack holds LST entries per lane in LDS and OVF more in
// private memory; the host picks OVF so that LST + OVF covers three entries per QNode4 level.
__device__ __forceinline__ float q_byte(uint32_t w, int k) { return (float)((w >> (8 * k)) & 0xffu); }

// Scenes at or beyond 2^60 take the decoding form (the fast form's terms could overflow).
__host__ __device__ __forceinline__ int q4_exact_only(float scale) { return scale < 1.152921504606846976e18f ? 0 : 1; }

template <bool ANY, int LST, int OVF, bool POST>
__device__ __forceinline__ void ray_q4_body(int32_t *__restrict__ s_ref, const QNode4 *__restrict__ qn,
                                            const RayTri *__restrict__ rt, const TriRec *__restrict__ tr,
                                            const float *__restrict__ orig, const float *__restrict__ dirs,
                                            int64_t ri, float best_init, float max_time, float scale,
                                            int exact_only, int32_t *__restrict__ out_idx,
                                            float *__restrict__ out_t, float *__restrict__ out_n,
                                            uint8_t *__restrict__ out_any) {
  int32_t ovf[OVF];
  const int tid = threadIdx.x;
  const float REL = 1.0f + 1.0f / 4096.0f;
  RayLane L;
  ray_lane_init(L, orig, dirs, ri, best_init, scale);
  const bool px = L.inv.x >= 0.0f, py = L.inv.y >= 0.0f, pz = L.inv.z >= 0.0f;
  const float3 on = f3(px ? L.olo.x : L.ohi.x, py ? L.olo.y : L.ohi.y, pz ? L.olo.z : L.ohi.z);
  const float3 of = f3(px ? L.ohi.x : L.olo.x, py ? L.ohi.y : L.olo.y, pz ? L.ohi.z : L.olo.z);
  const float INV_MAX = 1099511627776.0f;  // 2^40
  const bool lane_fast = fabsf(L.inv.x) <= INV_MAX && fabsf(L.inv.y) <= INV_MAX && fabsf(L.inv.z) <= INV_MAX;
  const bool fast = !exact_only && __ballot(!lane_fast) == 0;
  int sp = 0;
  auto push = [&](int32_t v) {
    if (sp < LST) s_ref[sp * Q_BLOCK + tid] = v;
    else if (sp < LST + OVF) ovf[sp - LST] = v;
    else return;  // never taken: LST + OVF covers the deepest path
    ++sp;
  };
  auto pop = [&]() -> int32_t {
    if (sp == 0) return RAY_DONE;
    --sp;
    return sp < LST ? s_ref[sp * Q_BLOCK + tid] : ovf[sp - LST];
  };
  // The triangles of leaf lf; true when an any-hit query has its answer.
  auto run_leaf = [&](int32_t lf) -> bool {
    int32_t k = ~lf;
    bool last;
    do {
      const RayTri *t = rt + k;
      const float4 a = t->a, b = t->b, c = t->c;
      last = raytri_last(b);
      const float th = tri_ray_e(f3(a.x, a.y, a.z), f3(b.x, b.y, b.z), f3(c.x, c.y, c.z), L.o, L.d, max_time);
      if (th >= 0) {
        const int32_t id = __float_as_int(a.w);
        if (better(th, id, L.best, L.bidx)) {
          L.best = th;
          L.bidx = id;
          L.bslot = k;
          if (ANY) return true;
        }
      }
      ++k;
    } while (!last);
    return false;
  };
  int32_t node = 0;
  int32_t leaf = 0;
  while (node != RAY_DONE || leaf < 0) {
    while (node >= 0) {
      const QNode4 *np = qn + node;
      const float4 na = np->a;
      const uint4 nb4 = np->b;
      const int4 nc4 = np->c;
      const int2 nd2 = *reinterpret_cast<const int2 *>(&np->d);
      const uint32_t meta = __float_as_uint(na.w);
      const int nc = (int)(meta >> 24);
      const float3 sc = f3(__uint_as_float((meta & 0xffu) << 23), __uint_as_float(((meta >> 8) & 0xffu) << 23),
                           __uint_as_float(((meta >> 16) & 0xffu) << 23));
      const uint32_t qhy = (uint32_t)nc4.x, qhz = (uint32_t)nc4.y;
      const uint32_t qnx = px ? nb4.x : nb4.w, qfx = px ? nb4.w : nb4.x;
      const uint32_t qny = py ? nb4.y : qhy, qfy = py ? qhy : nb4.y;
      const uint32_t qnz = pz ? nb4.z : qhz, qfz = pz ? qhz : nb4.z;
      const float lim = L.best * REL;
      float key[4];
      int32_t r[4] = {nc4.z, nc4.w, nd2.x, nd2.y};
      int h = 0;
      if (fast) {
        const float3 A = f3(sc.x * L.inv.x, sc.y * L.inv.y, sc.z * L.inv.z);
        const float3 Bn = f3((na.x - on.x) * L.inv.x, (na.y - on.y) * L.inv.y, (na.z - on.z) * L.inv.z);
        const float3 Bf = f3((na.x - of.x) * L.inv.x, (na.y - of.y) * L.inv.y, (na.z - of.z) * L.inv.z);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const float t0 = fmaxf(fmaxf(fmaf(q_byte(qnx, k), A.x, Bn.x), fmaf(q_byte(qny, k), A.y, Bn.y)),
                                 fmaxf(fmaf(q_byte(qnz, k), A.z, Bn.z), 0.0f));
          const float t1 = fminf(fminf(fmaf(q_byte(qfx, k), A.x, Bf.x), fmaf(q_byte(qfy, k), A.y, Bf.y)),
                                 fmaf(q_byte(qfz, k), A.z, Bf.z));
          const bool hit = k < nc && t0 <= t1 * REL && t0 <= lim;
          key[k] = hit ? fminf(t0, 3.40282347e+38f) : INFINITY;
          h += hit ? 1 : 0;
        }
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const float t0 =
              fmaxf(fmaxf((fmaf(q_byte(qnx, k), sc.x, na.x) - on.x) * L.inv.x,
                          (fmaf(q_byte(qny, k), sc.y, na.y) - on.y) * L.inv.y),
                    fmaxf((fmaf(q_byte(qnz, k), sc.z, na.z) - on.z) * L.inv.z, 0.0f));
          const float t1 = fminf(fminf((fmaf(q_byte(qfx, k), sc.x, na.x) - of.x) * L.inv.x,
                                       (fmaf(q_byte(qfy, k), sc.y, na.y) - of.y) * L.inv.y),
                                 (fmaf(q_byte(qfz, k), sc.z, na.z) - of.z) * L.inv.z);
          const bool hit = k < nc && t0 <= t1 * REL && t0 <= lim;
          key[k] = hit ? fminf(t0, 3.40282347e+38f) : INFINITY;
          h += hit ? 1 : 0;
        }
      }
      // Nearest first (misses, keyed +inf, sort last).
#define THG_CSWAP(i, j)                                              \
  if (key[j] < key[i]) {                                             \
    const float tk = key[i]; key[i] = key[j]; key[j] = tk;           \
    const int32_t tr_ = r[i]; r[i] = r[j]; r[j] = tr_;               \
  }
      THG_CSWAP(0, 1) THG_CSWAP(2, 3) THG_CSWAP(0, 2) THG_CSWAP(1, 3) THG_CSWAP(1, 2)
#undef THG_CSWAP
      if (h == 0) {
        node = pop();
      } else {
        node = r[0];
        if (h > 3) push(r[3]);
        if (h > 2) push(r[2]);
        if (h > 1) push(r[1]);
      }
      if (!POST) {
        // Leaves are tested as soon as they are met (any hit: the first hit ends the ray).
        while (node < 0 && node != RAY_DONE) node = run_leaf(node) ? RAY_DONE : pop();
        continue;
      }
      if (node < 0 && node != RAY_DONE && leaf == 0) {
        leaf = node;
        node = pop();
      }
      if (__ballot(leaf == 0) == 0) break;
    }
    while (leaf < 0) {
      if (run_leaf(leaf)) {
        node = RAY_DONE;
        leaf = 0;
        break;
      }
      leaf = 0;
      if (node < 0 && node != RAY_DONE) {
        leaf = node;
        node = pop();
      }
    }
  }
  ray_lane_finish<ANY>(L, tr, out_idx, out_t, out_n, out_any);
}

template <bool ANY, int LST, int OVF, bool POST>
__global__ void __launch_bounds__(Q_BLOCK) k_ray_q4(const QNode4 *__restrict__ qn, const RayTri *__restrict__ rt,
                                                    const TriRec *__restrict__ tr, const float *__restrict__ orig,
                                                    const float *__restrict__ dirs, const uint32_t *__restrict__ order,
                                                    int64_t n, uint32_t nb_real, float best_init, float max_time,
                                                    float scale, int exact_only, int32_t *__restrict__ out_idx,
                                                    float *__restrict__ out_t, float *__restrict__ out_n,
                                                    uint8_t *__restrict__ out_any) {
  __shared__ int32_t s_ref[LST * Q_BLOCK];
  bool active;
  const int64_t ri = query_slot(nb_real, n, order, active);
  if (!active) return;
  ray_q4_body<ANY, LST, OVF, POST>(s_ref, qn, rt, tr, orig, dirs, ri, best_init, max_time, scale, exact_only,
                                   out_idx, out_t, out_n, out_any);
}

// Many first-hit ray batches in one launch over their meshes' QNode4 trees; one 64-lane block per
// wave of a segment (SEG_WAVE_QUERIES rays).
template <int LST, int OVF>
__global__ void __launch_bounds__(Q_BLOCK) k_ray_q4_multi(const SegDesc *__restrict__ seg, int nseg,
                                                          uint64_t total_waves, uint32_t nb_real) {
  __shared__ int32_t s_ref[LST * Q_BLOCK];
  const uint32_t lb_blk = xcd_remap(blockIdx.x, gridDim.x);
  if (lb_blk >= nb_real || (uint64_t)lb_blk >= total_waves) return;
  const int si = uni(find_segment(seg, nseg, (int64_t)lb_blk));
  const SegDesc &d = seg[si];
  const int64_t t = ((int64_t)lb_blk - (int64_t)d.wave_begin) * Q_BLOCK + threadIdx.x;
  if (t >= d.n) return;
  const int64_t ri = d.order ? (int64_t)d.order[t] : t;
  ray_q4_body<false, LST, OVF, true>(s_ref, readonly_ptr<QNode4>(d.qnodes), readonly_ptr<RayTri>(d.rays),
                                     readonly_ptr<TriRec>(d.elems), readonly_ptr<float>(d.q), readonly_ptr<float>(d.q2), ri, d.best_init, d.max_time, d.scale, q4_exact_only(d.scale), d.idx, d.dist,
                               d.cp, nullptr);
}

//--------------------------------------------------------------------------------------------------
// Spatial ordering of a query batch
//--------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t spread10q(uint32_t v) {
  v &= 0x3FF;
  v = (v | (v << 16)) & 0x030000FF;
  v = (v | (v << 8)) & 0x0300F00F;
  v = (v | (v << 4)) & 0x030C30C3;
  v = (v | (v << 2)) & 0x09249249;
  return v;
}

__global__ void __launch_bounds__(256) k_query_keys(const float *__restrict__ q, int64_t n, float3 lo, float3 inv,
                                                    uint32_t *__restrict__ keys, uint32_t *__restrict__ vals) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float3 p = load_q(q, i);
  float x = (p.x - lo.x) * inv.x, y = (p.y - lo.y) * inv.y, z = (p.z - lo.z) * inv.z;
  // NaN coordinates map to cell 0 (fmaxf drops NaN).
  uint32_t ix = (uint32_t)fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
  uint32_t iy = (uint32_t)fminf(fmaxf(y * 1024.0f, 0.0f), 1023.0f);
  uint32_t iz = (uint32_t)fminf(fmaxf(z * 1024.0f, 0.0f), 1023.0f);
  keys[i] = (spread10q(ix) << 2) | (spread10q(iy) << 1) | spread10q(iz);
  vals[i] = (uint32_t)i;
}

// Ray tracing order key: the direction's octant (3 bits) above the 27-bit Morton code of the
// origin (9 bits per axis), so that a wave takes rays of one octant from nearby origins.
__global__ void __launch_bounds__(256) k_ray_keys(const float *__restrict__ o, const float *__restrict__ d, int64_t n,
                                                  float3 lo, float3 inv, uint32_t *__restrict__ keys,
                                                  uint32_t *__restrict__ vals) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float3 p = load_q(o, i), v = load_q(d, i);
  float x = (p.x - lo.x) * inv.x, y = (p.y - lo.y) * inv.y, z = (p.z - lo.z) * inv.z;
  uint32_t ix = (uint32_t)fminf(fmaxf(x * 512.0f, 0.0f), 511.0f);
  uint32_t iy = (uint32_t)fminf(fmaxf(y * 512.0f, 0.0f), 511.0f);
  uint32_t iz = (uint32_t)fminf(fmaxf(z * 512.0f, 0.0f), 511.0f);
  const uint32_t oct = (v.x < 0 ? 4u : 0u) | (v.y < 0 ? 2u : 0u) | (v.z < 0 ? 1u : 0u);
  keys[i] = (oct << 27) | (spread10q(ix) << 2) | (spread10q(iy) << 1) | spread10q(iz);
  vals[i] = (uint32_t)i;
}

//--------------------------------------------------------------------------------------------------
// Host launchers
//--------------------------------------------------------------------------------------------------
static int pick_stack(int depth) {
  if (depth <= 16) return 16;
  if (depth <= 32) return 32;
  if (depth <= 64) return 64;
  if (depth <= 128) return 128;
  fail(THG_E_UNSUPPORTED, "BVH deeper than 128 levels");
}

// log2 of the lanes per query of a sparse batch of n queries on a tree of the given depth (0: one
// lane each): the largest power of two up to 64 with n * G <= GRP_LANES, and at least 2 (a batch
// too large for the budget still halves its slowest query). Trees deeper than 63 levels keep one
// lane per query. THG_GROUP_LANES overrides the lane budget (0 turns grouping off).
int group_lanes_log2(int64_t n, int depth) {
  static const int64_t budget = getenv("THG_GROUP_LANES") ? atoll(getenv("THG_GROUP_LANES")) : GRP_LANES;
  if (depth + 1 > 64 || budget <= 0) return 0;
  int lg = 1;
  while (lg < 6 && n * ((int64_t)2 << lg) <= budget) ++lg;
  return lg;
}

// Lowest Morton bit the query sort uses (THG_ORDER_BITS = number of top bits of the 30, default 24).
static int order_bits_lo() {
  static const char *env = getenv("THG_ORDER_BITS");
  int bits = env ? atoi(env) : 24;  // 3 radix passes; 30 bits (4 passes) is no faster on C3
  bits = bits < 8 ? 8 : (bits > 30 ? 30 : bits);
  return 30 - bits;
}

size_t sort_scratch_words(int64_t cap) { return (size_t)4 * (size_t)std::max<int64_t>(cap, 1) + radix_hist_words(cap); }

QueryScratch::Sort QueryScratch::sort_memory(int64_t n, hipStream_t s) {
  if (ext) {
    require(n <= ext_cap, "batch larger than the caller-provided sort memory");
    const size_t c = (size_t)std::max<int64_t>(ext_cap, 1);
    return Sort{ext, ext + c, ext + 2 * c, ext + 3 * c, ext + 4 * c};
  }
  keys.ensure(n * 4, s);
  vals.ensure(n * 4, s);
  tk.ensure(n * 4, s);
  tv.ensure(n * 4, s);
  hist.ensure(radix_hist_words(n) * 4, s);
  return Sort{keys.as<uint32_t>(), vals.as<uint32_t>(), tk.as<uint32_t>(), tv.as<uint32_t>(), hist.as<uint32_t>()};
}

void order_queries(const Bvh &b, const float *q, int64_t n, QueryScratch &ws, hipStream_t s) {
  ws.order = nullptr;
  if (n < ORDER_MIN_QUERIES) return;
  KernelTimer kt(ws.timer_ctx(b.ctx), s, "order");
  const QueryScratch::Sort m = ws.sort_memory(n, s);
  float3 ext = f3(b.hi.x - b.lo.x, b.hi.y - b.lo.y, b.hi.z - b.lo.z);
  float3 inv = f3(ext.x > 0 ? 1.0f / ext.x : 0.0f, ext.y > 0 ? 1.0f / ext.y : 0.0f, ext.z > 0 ? 1.0f / ext.z : 0.0f);
  hipLaunchKernelGGL(k_query_keys, dim3(ceil_div(n, 256)), dim3(256), 0, s, q, n, b.lo, inv, m.keys, m.vals);
  launch_check("k_query_keys");
  // The top 24 of the 30 code bits give cells of 1/256 of the scene per axis: enough coherence.
  radix_sort_pairs(m.keys, m.vals, m.tk, m.tv, m.hist, n, order_bits_lo(), 30, s);
  ws.order = m.vals;
  kt.stop();
}

// Tracing order of a ray batch: direction octant, then origin Morton code (THG_RAY_ORDER=origin:
// the origin's Morton code alone, as for point queries). On C3 rays (random origins and
// directions) the octant-first order takes the first-hit kernel from 1.39 to 1.25 ms.
void order_rays(const Bvh &b, const float *o, const float *d, int64_t n, QueryScratch &ws, hipStream_t s) {
  static const bool by_origin = [] {
    const char *e = getenv("THG_RAY_ORDER");
    return e && std::string(e) == "origin";
  }();
  if (by_origin) {
    order_queries(b, o, n, ws, s);
    return;
  }
  ws.order = nullptr;
  if (n < ORDER_MIN_QUERIES) return;
  KernelTimer kt(ws.timer_ctx(b.ctx), s, "order");
  const QueryScratch::Sort m = ws.sort_memory(n, s);
  float3 ext = f3(b.hi.x - b.lo.x, b.hi.y - b.lo.y, b.hi.z - b.lo.z);
  float3 inv = f3(ext.x > 0 ? 1.0f / ext.x : 0.0f, ext.y > 0 ? 1.0f / ext.y : 0.0f, ext.z > 0 ? 1.0f / ext.z : 0.0f);
  hipLaunchKernelGGL(k_ray_keys, dim3(ceil_div(n, 256)), dim3(256), 0, s, o, d, n, b.lo, inv, m.keys, m.vals);
  launch_check("k_ray_keys");
  radix_sort_pairs(m.keys, m.vals, m.tk, m.tv, m.hist, n, order_bits_lo(), 30, s);
  ws.order = m.vals;
  kt.stop();
}

static float prev_float(float x) { return std::nextafter(x, -INFINITY); }

// Traversal strategy for a batch of n point-proximity queries. One wave walking the tree for 64
// sorted queries pays for the union of their search regions, which stays small only when the batch
// is dense: at least about one query per 8 leaves (measured on spheres and tori of 0.1M-4M
// triangles with queries uniform in the inflated box, where the cheaper of the two strategies
// changes between 1/4 and 1/16 of a query per leaf). Sparser batches, and batches too small to be
// spatially sorted, walk one query per lane. THG_TRAVERSAL=lane or =wave forces a strategy; trees
// deeper than the lane-distributed stack always use lanes.
bool use_wave_traversal(const Bvh &b, int64_t n) {
  static const char *env = getenv("THG_TRAVERSAL");
  if (b.max_depth >= W_STACK) return false;
  if (env && std::string(env) == "lane") return false;
  if (env && std::string(env) == "wave") return true;
  return n >= ORDER_MIN_QUERIES && n >= b.num_leaves / 8;
}

// Lanes per query and per-element prefilter of the wave kernel (THG_WAVE_G, THG_PREFILTER override).
struct WaveConfig {
  int g;
  bool pre;
  bool cmp;
};
static WaveConfig wave_config(const Bvh &b) {
  WaveConfig c{b.wave_g, b.wave_pre, b.wave_cmp};
  if (const char *e = getenv("THG_WAVE_G")) c.g = atoi(e);
  if (const char *e = getenv("THG_PREFILTER")) c.pre = atoi(e) != 0;
  if (const char *e = getenv("THG_COMPACT")) c.cmp = atoi(e) != 0;
  if (c.g != 1 && c.g != 2 && c.g != 4) c.g = 1;
  return c;
}

// Largest float strictly below the squared bound, or +inf when unbounded: with the lowest-index tie
// rule, initialising the best to (this, INT_MAX) accepts exactly d^2 < dist_bound^2, as Thea's
// double comparison does (BvhN.hpp:446-447, :459-466).
float bound_to_best(double dist_bound) {
  if (dist_bound < 0) return INFINITY;
  double b2 = dist_bound * dist_bound;
  float f = (float)b2;  // round to nearest
  if ((double)f < b2) f = std::nextafter(f, INFINITY);  // smallest float >= b2
  return prev_float(f);
}

void launch_closest(const Bvh &b, const float *q, int64_t n, double dist_bound, int exclude_self, int32_t *idx,
                    float *cp, float *dist, float *d2, QueryScratch &ws, hipStream_t s) {
  if (n <= 0) return;
  order_queries(b, q, n, ws, s);
  uint32_t nb = ceil_div(n, Q_BLOCK);
  uint32_t grid = round_up8(nb);
  float best = bound_to_best(dist_bound);
  if (use_wave_traversal(b, n)) {
    const WaveConfig wc = wave_config(b);
    uint32_t wnb = ceil_div(n, (uint64_t)W_WAVES * (WAVE / wc.g));
    uint32_t wgrid = round_up8(wnb);
    KernelTimer kt(ws.timer_ctx(b.ctx), s, "closest");
    if (wc.cmp && !b.points) {
      // A batch whose waves all fit at once at 6 per SIMD gains nothing from more resident waves and
      // would only pay the 8-wave build's spill (C1, 100k queries: 1.56 against 1.64 ms).
      const bool one_round = ceil_div(n, (uint64_t)WAVE) <= (uint64_t)b.ctx->num_cus * 4 * 6;
      const uint32_t cgrid = round_up8(ceil_div(n, (uint64_t)W_BLOCK)), cnb = (uint32_t)ceil_div(n, (uint64_t)W_BLOCK);
      // When every wave is resident at once, the walk's load latency is exposed: prefetch leaves
      // (PF; C1, 100k queries: 1.55 -> 1.34 ms; 300k: 1.61 -> 1.36 ms). THG_LEAF_PREFETCH=0 or 1
      // overrides, for either build.
      static const int pf_env = env_int("THG_LEAF_PREFETCH", -1);
      const bool pf = pf_env >= 0 ? pf_env != 0 : one_round;
      if (!one_round && pf)
        hipLaunchKernelGGL((k_closest_cmp<false, CMP_WAVES, true>), dim3(cgrid), dim3(W_BLOCK), 0, s,
                           b.nodes.as<BNode>(), b.elems.as<TriRec>(), q, ws.order, n, cnb, best, b.scale, exclude_self,
                           idx, cp, dist, d2, nullptr, b.child_order, b.lbounds.as<LeafBound>(),
                           b.screen.as<ScreenRec>());
      else if (one_round && pf)
        hipLaunchKernelGGL((k_closest_cmp<false, 1, true>), dim3(cgrid), dim3(W_BLOCK), 0, s, b.nodes.as<BNode>(),
                           b.elems.as<TriRec>(), q, ws.order, n, cnb, best, b.scale, exclude_self, idx, cp, dist, d2,
                           nullptr, b.child_order, b.lbounds.as<LeafBound>(),
                           b.screen.as<ScreenRec>());
      else if (one_round)
        hipLaunchKernelGGL((k_closest_cmp<false, 1>), dim3(cgrid), dim3(W_BLOCK), 0, s, b.nodes.as<BNode>(),
                           b.elems.as<TriRec>(), q, ws.order, n, cnb, best, b.scale, exclude_self, idx, cp, dist, d2,
                           nullptr, b.child_order, b.lbounds.as<LeafBound>(),
                           b.screen.as<ScreenRec>());
      else
        hipLaunchKernelGGL((k_closest_cmp<false, CMP_WAVES>), dim3(cgrid), dim3(W_BLOCK), 0, s, b.nodes.as<BNode>(),
                           b.elems.as<TriRec>(), q, ws.order, n, cnb, best, b.scale, exclude_self, idx, cp, dist, d2,
                           nullptr, b.child_order, b.lbounds.as<LeafBound>(),
                           b.screen.as<ScreenRec>());
      launch_check("k_closest_cmp");
      kt.stop();
      return;
    }
#define THG_CW(P, G, PRE)                                                                                              \
  hipLaunchKernelGGL((k_closest_wave<P, G, PRE>), dim3(wgrid), dim3(W_BLOCK), 0, s, b.nodes.as<BNode>(), b.elems.p, q, \
                     ws.order, n, wnb, best, b.scale, exclude_self, idx, cp, dist, d2)
    if (b.points) {
      if (wc.g == 1) THG_CW(true, 1, false); else if (wc.g == 2) THG_CW(true, 2, false); else THG_CW(true, 4, false);
    } else if (wc.pre) {
      if (wc.g == 1) THG_CW(false, 1, true); else if (wc.g == 2) THG_CW(false, 2, true); else THG_CW(false, 4, true);
    } else {
      if (wc.g == 1) THG_CW(false, 1, false); else if (wc.g == 2) THG_CW(false, 2, false); else THG_CW(false, 4, false);
    }
#undef THG_CW
    launch_check("k_closest_wave");
    kt.stop();
    return;
  }
  if (const int lg = group_lanes_log2(n, b.max_depth)) {
    const uint32_t gnb = (uint32_t)ceil_div(n << lg, Q_BLOCK);
    const bool deep = b.max_depth + 1 > 32;
    KernelTimer kt(ws.timer_ctx(b.ctx), s, "closest");
#define THG_CG(S, P)                                                                                                  \
  hipLaunchKernelGGL((k_closest_group<S, P>), dim3(round_up8(gnb)), dim3(Q_BLOCK), 0, s, b.nodes.as<BNode>(), b.elems.p, \
                     q, ws.order, n, lg, gnb, best, b.scale, exclude_self, idx, cp, dist, d2)
    if (b.points) { if (deep) THG_CG(64, true); else THG_CG(32, true); }
    else { if (deep) THG_CG(64, false); else THG_CG(32, false); }
#undef THG_CG
    launch_check("k_closest_group");
    kt.stop();
    return;
  }
  int st = pick_stack(b.max_depth);
  KernelTimer kt(ws.timer_ctx(b.ctx), s, "closest");
#define THG_CLOSEST(S, P)                                                                                            \
  hipLaunchKernelGGL((k_closest<S, P>), dim3(grid), dim3(Q_BLOCK), 0, s, b.nodes.as<BNode>(), b.elems.p, q, ws.order, n, \
                     nb, best, b.scale, exclude_self, idx, cp, dist, d2)
  if (b.points) {
    if (st == 16) THG_CLOSEST(16, true); else if (st == 32) THG_CLOSEST(32, true);
    else if (st == 64) THG_CLOSEST(64, true); else THG_CLOSEST(128, true);
  } else {
    if (st == 16) THG_CLOSEST(16, false); else if (st == 32) THG_CLOSEST(32, false);
    else if (st == 64) THG_CLOSEST(64, false); else THG_CLOSEST(128, false);
  }
#undef THG_CLOSEST
  launch_check("k_closest");
  kt.stop();
}

template <int K>
static void launch_knn_k(const Bvh &b, int st, uint32_t grid, uint32_t nb, const float *q, int64_t n, int k, float best,
                         int exclude_self, int32_t *idx, float *dist, float *d2, const uint32_t *order, hipStream_t s) {
#define THG_KNN(S, P)                                                                                                \
  hipLaunchKernelGGL((k_knn<S, P, K>), dim3(grid), dim3(Q_BLOCK), 0, s, b.nodes.as<BNode>(), b.elems.p, q, order, n, nb, \
                     k, best, b.scale, exclude_self, idx, dist, d2)
  if (b.points) {
    if (st <= 32) THG_KNN(32, true); else if (st == 64) THG_KNN(64, true); else THG_KNN(128, true);
  } else {
    if (st <= 32) THG_KNN(32, false); else if (st == 64) THG_KNN(64, false); else THG_KNN(128, false);
  }
#undef THG_KNN
}

template <int K>
static void launch_knn_group_k(const Bvh &b, bool deep, uint32_t gnb, int lg, const float *q, int64_t n, int k,
                               float best, int exclude_self, int32_t *idx, float *dist, float *d2,
                               const uint32_t *order, hipStream_t s) {
#define THG_KG(S, P)                                                                                                \
  hipLaunchKernelGGL((k_knn_group<S, P, K>), dim3(round_up8(gnb)), dim3(Q_BLOCK), 0, s, b.nodes.as<BNode>(), b.elems.p, \
                     q, order, n, lg, gnb, k, best, b.scale, exclude_self, idx, dist, d2)
  if (b.points) { if (deep) THG_KG(64, true); else THG_KG(32, true); }
  else { if (deep) THG_KG(64, false); else THG_KG(32, false); }
#undef THG_KG
}

void launch_closest_multi(const SegDesc *seg_dev, int nseg, uint64_t total_waves, bool points, hipStream_t s,
                          bool concurrent, int block_waves, bool spill_build) {
  if (nseg <= 0 || total_waves == 0) return;
  if (points) {
    const uint32_t nb = ceil_div(total_waves, W_WAVES);
    hipLaunchKernelGGL((k_closest_multi<true, 1>), dim3(round_up8(nb)), dim3(W_BLOCK), 0, s, seg_dev, nseg, total_waves,
                       nb);
    launch_check("k_closest_multi");
    return;
  }
  // block_waves: waves per workgroup (1, 2, 4 or 8); spill_build: the build for 8 resident waves per
  // SIMD (a 44-byte spill) instead of the compiler's register allocation. Concurrent launches (the
  // request server's streams) keep the spill-free build.
  const bool w8 = spill_build && !concurrent;
  const uint32_t nb = ceil_div(total_waves, (uint64_t)block_waves);
  const uint32_t grid = round_up8(nb);
#define THG_CM(W, B)                                                                                                   \
  hipLaunchKernelGGL((k_closest_multi<false, W, B>), dim3(grid), dim3(B * WAVE), 0, s, seg_dev, nseg, total_waves, nb)
  if (w8) {
    if (block_waves == 1) THG_CM(CMP_WAVES, 1); else if (block_waves == 2) THG_CM(CMP_WAVES, 2);
    else if (block_waves == 8) THG_CM(CMP_WAVES, 8); else THG_CM(CMP_WAVES, 4);
  } else {
    if (block_waves == 1) THG_CM(1, 1); else if (block_waves == 2) THG_CM(1, 2);
    else if (block_waves == 8) THG_CM(1, 8); else THG_CM(1, 4);
  }
#undef THG_CM
  launch_check("k_closest_multi");
}

void launch_knn_multi(const SegDesc *seg_dev, int nseg, uint64_t total_waves, int k, bool points, hipStream_t s) {
  if (nseg <= 0 || total_waves == 0) return;
  uint32_t nb = ceil_div(total_waves, W_WAVES);
  uint32_t grid = round_up8(nb);
#define THG_KM(P, K) hipLaunchKernelGGL((k_knn_multi<P, K>), dim3(grid), dim3(W_BLOCK), 0, s, seg_dev, nseg, total_waves, nb)
  if (points) {
    if (k <= 4) THG_KM(true, 4); else if (k <= 8) THG_KM(true, 8); else if (k <= 16) THG_KM(true, 16); else THG_KM(true, 32);
  } else {
    if (k <= 4) THG_KM(false, 4); else if (k <= 8) THG_KM(false, 8); else if (k <= 16) THG_KM(false, 16); else THG_KM(false, 32);
  }
#undef THG_KM
  launch_check("k_knn_multi");
}

void launch_closest_lane_multi(const SegDesc *seg_dev, int nseg, uint64_t total_waves, bool points, int stack,
                               hipStream_t s) {
  if (nseg <= 0 || total_waves == 0) return;
  const uint32_t nb = (uint32_t)total_waves;
  const uint32_t grid = round_up8(nb);
#define THG_CLM(S, P) hipLaunchKernelGGL((k_closest_lane_multi<S, P>), dim3(grid), dim3(Q_BLOCK), 0, s, seg_dev, nseg, \
                                         total_waves, nb)
  if (points) { if (stack <= 32) THG_CLM(32, true); else THG_CLM(64, true); }
  else { if (stack <= 32) THG_CLM(32, false); else THG_CLM(64, false); }
#undef THG_CLM
  launch_check("k_closest_lane_multi");
}

void launch_knn_lane_multi(const SegDesc *seg_dev, int nseg, uint64_t total_waves, int k, bool points, int stack,
                           hipStream_t s) {
  if (nseg <= 0 || total_waves == 0) return;
  const uint32_t nb = (uint32_t)total_waves;
  const uint32_t grid = round_up8(nb);
#define THG_KLM(S, P, K) hipLaunchKernelGGL((k_knn_lane_multi<S, P, K>), dim3(grid), dim3(Q_BLOCK), 0, s, seg_dev, nseg, \
                                            total_waves, nb)
#define THG_KLM_K(S, P) \
  if (k <= 4) THG_KLM(S, P, 4); else if (k <= 8) THG_KLM(S, P, 8); else if (k <= 16) THG_KLM(S, P, 16); else THG_KLM(S, P, 32)
  if (points) { if (stack <= 32) { THG_KLM_K(32, true); } else { THG_KLM_K(64, true); } }
  else { if (stack <= 32) { THG_KLM_K(32, false); } else { THG_KLM_K(64, false); } }
#undef THG_KLM_K
#undef THG_KLM
  launch_check("k_knn_lane_multi");
}

void launch_closest_group_multi(const SegDesc *seg_dev, int nseg, uint64_t total_waves, bool points, int stack,
                                hipStream_t s) {
  if (nseg <= 0 || total_waves == 0) return;
  const uint32_t nb = (uint32_t)total_waves;
  const uint32_t grid = round_up8(nb);
#define THG_CGM(S, P) hipLaunchKernelGGL((k_closest_group_multi<S, P>), dim3(grid), dim3(Q_BLOCK), 0, s, seg_dev, nseg, \
                                         total_waves, nb)
  if (points) { if (stack <= 32) THG_CGM(32, true); else THG_CGM(64, true); }
  else { if (stack <= 32) THG_CGM(32, false); else THG_CGM(64, false); }
#undef THG_CGM
  launch_check("k_closest_group_multi");
}

void launch_knn_group_multi(const SegDesc *seg_dev, int nseg, uint64_t total_waves, int k, bool points, int stack,
                            hipStream_t s) {
  if (nseg <= 0 || total_waves == 0) return;
  const uint32_t nb = (uint32_t)total_waves;
  const uint32_t grid = round_up8(nb);
#define THG_KGM(S, P, K) hipLaunchKernelGGL((k_knn_group_multi<S, P, K>), dim3(grid), dim3(Q_BLOCK), 0, s, seg_dev, nseg, \
                                            total_waves, nb)
#define THG_KGM_K(S, P) \
  if (k <= 4) THG_KGM(S, P, 4); else if (k <= 8) THG_KGM(S, P, 8); else if (k <= 16) THG_KGM(S, P, 16); else THG_KGM(S, P, 32)
  if (points) { if (stack <= 32) { THG_KGM_K(32, true); } else { THG_KGM_K(64, true); } }
  else { if (stack <= 32) { THG_KGM_K(32, false); } else { THG_KGM_K(64, false); } }
#undef THG_KGM_K
#undef THG_KGM
  launch_check("k_knn_group_multi");
}

void launch_ray_multi(const SegDesc *seg_dev, int nseg, uint64_t total_waves, int stack, hipStream_t s,
                      int q_stack) {
  if (nseg <= 0 || total_waves == 0) return;
  uint32_t nb = (uint32_t)total_waves;
  uint32_t grid = round_up8(nb);
  if (q_stack > 0) {
    require(q_stack <= Q4_MAX_STACK, "ray tree too deep for the 4-wide traversal stack");
    if (q_stack <= 16 + 32)
      hipLaunchKernelGGL((k_ray_q4_multi<16, 32>), dim3(grid), dim3(Q_BLOCK), 0, s, seg_dev, nseg, total_waves, nb);
    else
      hipLaunchKernelGGL((k_ray_q4_multi<16, 96>), dim3(grid), dim3(Q_BLOCK), 0, s, seg_dev, nseg, total_waves, nb);
    launch_check("k_ray_q4_multi");
    return;
  }
  if (stack <= 32) hipLaunchKernelGGL(k_ray_multi<32>, dim3(grid), dim3(Q_BLOCK), 0, s, seg_dev, nseg, total_waves, nb);
  else hipLaunchKernelGGL(k_ray_multi<64>, dim3(grid), dim3(Q_BLOCK), 0, s, seg_dev, nseg, total_waves, nb);
  launch_check("k_ray_multi");
}

void ray_stats(const Bvh &b, const float *o, const float *d, int64_t n, int32_t *idx, int32_t *counts, hipStream_t s) {
  if (n <= 0 || b.n == 0) return;
  const uint32_t nb = ceil_div(n, Q_BLOCK);
  const int st = pick_stack(b.max_depth);
#define THG_RS(S) hipLaunchKernelGGL(k_ray_stats<S>, dim3(nb), dim3(Q_BLOCK), 0, s, b.nodes.as<BNode>(), \
                                     b.rays.as<RayTri>(), b.elems.as<TriRec>(), o, d, n, b.scale, idx, counts)
  if (st <= 32) THG_RS(32); else if (st == 64) THG_RS(64); else THG_RS(128);
#undef THG_RS
  launch_check("k_ray_stats");
}

void closest_wave_stats(const Bvh &b, const float *q, int64_t n, unsigned long long *stats, QueryScratch &ws,
                        hipStream_t s) {
  if (n <= 0 || b.n == 0) return;
  order_queries(b, q, n, ws, s);
  AsyncBuf idx(n * 4, s);
  uint32_t wnb = ceil_div(n, W_BLOCK);
  uint32_t wgrid = round_up8(wnb);
  float best = INFINITY;
  if (!b.points && wave_config(b).cmp)
    hipLaunchKernelGGL(k_closest_cmp<true>, dim3(wgrid), dim3(W_BLOCK), 0, s, b.nodes.as<BNode>(), b.elems.as<TriRec>(),
                       q, ws.order, n, wnb, best, b.scale, 0, idx.as<int32_t>(), nullptr, nullptr, nullptr, stats,
                       (int)b.child_order, b.lbounds.as<LeafBound>(),
                           b.screen.as<ScreenRec>());
  else if (b.points)
    hipLaunchKernelGGL((k_closest_wave<true, 1, false, true>), dim3(wgrid), dim3(W_BLOCK), 0, s, b.nodes.as<BNode>(),
                       b.elems.p, q, ws.order, n, wnb, best, b.scale, 0, idx.as<int32_t>(), nullptr, nullptr, nullptr, stats);
  else
    hipLaunchKernelGGL((k_closest_wave<false, 1, true, true>), dim3(wgrid), dim3(W_BLOCK), 0, s, b.nodes.as<BNode>(),
                       b.elems.p, q, ws.order, n, wnb, best, b.scale, 0, idx.as<int32_t>(), nullptr, nullptr, nullptr, stats);
  launch_check("k_closest_wave(stats)");
}

void knn_wave_stats(const Bvh &b, const float *q, int64_t n, int exclude_self, int64_t *wstats, QueryScratch &ws,
                    hipStream_t s) {
  require(b.points && b.leaf_size <= 16, "k-NN wave statistics need a point BVH with leaves of at most 16");
  if (n <= 0 || b.n == 0) return;
  order_queries(b, q, n, ws, s);
  AsyncBuf idx(n * 16 * 4, s);
  const uint32_t wnb = ceil_div(n, W_BLOCK);
  hipLaunchKernelGGL((k_knn_pts<16, true>), dim3(round_up8(wnb)), dim3(W_BLOCK), 0, s, b.nodes.as<BNode>(),
                     b.elems.as<PointRec>(), q, ws.order, n, wnb, 16, INFINITY, b.scale, exclude_self, idx.as<int32_t>(),
                     nullptr, nullptr, (int32_t)b.n, wstats);
  launch_check("k_knn_pts(stats)");
}

void launch_knn(const Bvh &b, const float *q, int64_t n, int k, double dist_bound, int exclude_self, int32_t *idx,
                float *dist, float *d2, QueryScratch &ws, hipStream_t s) {
  if (n <= 0) return;
  order_queries(b, q, n, ws, s);
  uint32_t nb = ceil_div(n, Q_BLOCK);
  uint32_t grid = round_up8(nb);
  float best = bound_to_best(dist_bound);
  if (use_wave_traversal(b, n)) {
    uint32_t wnb = ceil_div(n, W_BLOCK);
    uint32_t wgrid = round_up8(wnb);
    KernelTimer kt(b.ctx, s, "knn");
    // Point BVHs take the batched-leaf kernel (THG_KNN_BATCH=0 selects the one-point-at-a-time walk).
    static const bool batch = !(getenv("THG_KNN_BATCH") && atoi(getenv("THG_KNN_BATCH")) == 0);
#define THG_KP(K)                                                                                                     \
  hipLaunchKernelGGL((k_knn_pts<K>), dim3(wgrid), dim3(W_BLOCK), 0, s, b.nodes.as<BNode>(), b.elems.as<PointRec>(), q, \
                     ws.order, n, wnb, k, best, b.scale, exclude_self, idx, dist, d2, (int32_t)b.n)
#define THG_KW(P, K)                                                                                                 \
  hipLaunchKernelGGL((k_knn_wave<P, K>), dim3(wgrid), dim3(W_BLOCK), 0, s, b.nodes.as<BNode>(), b.elems.p, q, ws.order, \
                     n, wnb, k, best, b.scale, exclude_self, idx, dist, d2, (int32_t)b.n, b.lbounds.as<LeafBound>())
    if (b.points && batch && b.leaf_size <= 16 && k <= 16) {  // K = 32 would spill its 64-bit keys
      if (k <= 4) THG_KP(4); else if (k <= 8) THG_KP(8); else THG_KP(16);
    } else if (b.points) {
      if (k <= 4) THG_KW(true, 4); else if (k <= 8) THG_KW(true, 8); else if (k <= 16) THG_KW(true, 16); else THG_KW(true, 32);
    } else {
      if (k <= 4) THG_KW(false, 4); else if (k <= 8) THG_KW(false, 8); else if (k <= 16) THG_KW(false, 16); else THG_KW(false, 32);
    }
#undef THG_KW
#undef THG_KP
    launch_check("k_knn_wave");
    kt.stop();
    return;
  }
  if (const int lg = group_lanes_log2(n, b.max_depth)) {
    const uint32_t gnb = (uint32_t)ceil_div(n << lg, Q_BLOCK);
    const bool deep = b.max_depth + 1 > 32;
    KernelTimer kt(b.ctx, s, "knn");
    if (k <= 4) launch_knn_group_k<4>(b, deep, gnb, lg, q, n, k, best, exclude_self, idx, dist, d2, ws.order, s);
    else if (k <= 8) launch_knn_group_k<8>(b, deep, gnb, lg, q, n, k, best, exclude_self, idx, dist, d2, ws.order, s);
    else if (k <= 16) launch_knn_group_k<16>(b, deep, gnb, lg, q, n, k, best, exclude_self, idx, dist, d2, ws.order, s);
    else launch_knn_group_k<32>(b, deep, gnb, lg, q, n, k, best, exclude_self, idx, dist, d2, ws.order, s);
    launch_check("k_knn_group");
    kt.stop();
    return;
  }
  int st = pick_stack(b.max_depth);
  KernelTimer kt(b.ctx, s, "knn");
  if (k <= 4) launch_knn_k<4>(b, st, grid, nb, q, n, k, best, exclude_self, idx, dist, d2, ws.order, s);
  else if (k <= 8) launch_knn_k<8>(b, st, grid, nb, q, n, k, best, exclude_self, idx, dist, d2, ws.order, s);
  else if (k <= 16) launch_knn_k<16>(b, st, grid, nb, q, n, k, best, exclude_self, idx, dist, d2, ws.order, s);
  else launch_knn_k<32>(b, st, grid, nb, q, n, k, best, exclude_self, idx, dist, d2, ws.order, s);
  launch_check("k_knn");
  kt.stop();
}

void launch_ray(const Bvh &b, const float *o, const float *d, int64_t n, float max_time, bool any, int32_t *idx,
                float *t, float *nrm, uint8_t *hit, QueryScratch &ws, hipStream_t s, bool sort_rays,
                const uint32_t *given_order) {
  if (n <= 0) return;
  require(!b.points && (b.n == 0 || b.rays.p != nullptr), "ray queries need a triangle BVH");
  if (sort_rays) order_rays(b, o, d, n, ws, s);
  else ws.order = given_order;
  uint32_t nb = ceil_div(n, Q_BLOCK);
  uint32_t grid = round_up8(nb);
  // Thea accepts only hits strictly before max_time at the structure level (BvhN.hpp:2247-2258).
  float best = max_time >= 0 ? prev_float(max_time) : INFINITY;
  int st = pick_stack(b.max_depth);
  KernelTimer kt(b.ctx, s, any ? "any_hit" : "ray");
  // Quantized 4-wide tree (k_ray_q4; THG_RAY_Q4=0 selects the binary kernels below) when its stack
  // bound (three entries per level) fits 16 LDS + 96 private entries.
  static const bool q4 = env_int("THG_RAY_Q4", 1) != 0;
  if (q4 && b.qnodes.p && q4_stack(b) <= Q4_MAX_STACK) {
    const int exact_only = q4_exact_only(b.scale);
    const bool small = q4_stack(b) <= 16 + 32;
    // Postponed leaves (while-while) for first hits, leaves tested when met for any hit
    // (THG_RAY_POST / THG_RAY_ANY_POST override).
    static const bool post = env_int("THG_RAY_POST", 1) != 0, any_post = env_int("THG_RAY_ANY_POST", 0) != 0;
#define THG_RQ(A, O, P)                                                                                               \
  hipLaunchKernelGGL((k_ray_q4<A, 16, O, P>), dim3(grid), dim3(Q_BLOCK), 0, s, b.qnodes.as<QNode4>(),               \
                     b.rays.as<RayTri>(), b.elems.as<TriRec>(), o, d, ws.order, n, nb, best, max_time, b.scale,     \
                     exact_only, idx, t, nrm, hit)
#define THG_RQP(A, O, P) \
  if (P) { THG_RQ(A, O, true); } else { THG_RQ(A, O, false); }
    if (any) {
      if (small) { THG_RQP(true, 32, any_post); } else { THG_RQP(true, 96, any_post); }
    } else {
      if (small) { THG_RQP(false, 32, post); } else { THG_RQP(false, 96, post); }
    }
#undef THG_RQP
#undef THG_RQ
    launch_check("k_ray_q4");
    kt.stop();
    return;
  }
  // Entry times on the stack (THG_RAY_TSTACK=1) cost half the LDS and with it half the resident
  // waves; on C3 rays the ref-only stack is 1.7x faster.
  static const bool tstack = [] {
    const char *e = getenv("THG_RAY_TSTACK");
    return e && atoi(e) == 1;
  }();
#define THG_RAY(S, A)                                                                                                \
  if (tstack)                                                                                                        \
    hipLaunchKernelGGL((k_ray<S, A, true>), dim3(grid), dim3(Q_BLOCK), 0, s, b.nodes.as<BNode>(), b.rays.as<RayTri>(), \
                       b.elems.as<TriRec>(), o, d, ws.order, n, nb, best, max_time, b.scale, idx, t, nrm, hit);      \
  else                                                                                                               \
    hipLaunchKernelGGL((k_ray<S, A, false, LANE_REG>), dim3(grid), dim3(Q_BLOCK), 0, s, b.nodes.as<BNode>(),         \
                       b.rays.as<RayTri>(), b.elems.as<TriRec>(), o, d, ws.order, n, nb, best, max_time, b.scale, idx, \
                       t, nrm, hit)
  if (any) {
    if (st <= 32) THG_RAY(32, true); else if (st == 64) THG_RAY(64, true); else THG_RAY(128, true);
  } else {
    if (st <= 32) THG_RAY(32, false); else if (st == 64) THG_RAY(64, false); else THG_RAY(128, false);
  }
#undef THG_RAY
  launch_check("k_ray");
  kt.stop();
}

}  // namespace thg

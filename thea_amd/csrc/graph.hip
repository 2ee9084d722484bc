// Sample proximity graph: Thea's PointSet3::updateGraph (Algorithms/PointSet3.hpp:324-386) for a set
// of samples, with or without oversampling (extractOriginalAdjacencies below), without a surface
// for the reachability test.
//
//  1. Separation estimate: for T samples picked at random, the squared distance to the nearest other
//     sample, summed in pick order in float; avg = sqrt(sum / T) (PointSet3.hpp:362-376). Thea picks
//     with the time-seeded Random::common(); here pick t is the sample (w * n) >> 32 for the first
//     word w of Philox4x32-10 at counter (t, 0, GRAPH_PICK_STREAM, 0) with the caller's seed as key.
//     The nearest other sample comes from a 2-NN query on the point BVH (the first neighbour that is
//     not the pick itself).
//  2. Neighbours of each sample (findSampleNeighbors, PointSet3.hpp:660-678): the max_degree smallest
//     (separation, index) pairs among the other samples inside the ball of radius
//     sqrt(max_degree) * avg (|p - c|^2 <= r^2 in float, BallN.hpp:83) whose normals are not opposed
//     (n_i . n_j >= -0.5, PointSet3.hpp:623-628), separation = Thea's fastSqrt of the squared distance
//     (Math.hpp:194-219); if fewer than min_degree = clamp(max_degree / 4, 4, max_degree) qualify, the
//     radius doubles, at most 3 tries.
//  3. Final average separation: the float sum of every sample's separations in order, over the number
//     of edges (updateAverageSeparation, PointSet3.hpp:683-697), on the host.
// One lane per sample walks the point BVH with the ball as the pruning bound (no k-th-distance
// pruning: fastSqrt is not monotone to the last bit, so candidates are ordered by their actual keys).
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "device_types.hpp"
#include "kernels_common.hpp"
#include "philox.hpp"
#include "query.hpp"
#include "traverse.hpp"
#include "tri_math.hpp"

namespace thg {

constexpr uint32_t GRAPH_PICK_STREAM = 0x47524150u;
constexpr int G_BLOCK = 64;

// Math::fastRsq(float32) (Math.hpp:194-211): Quake III's reciprocal square root with one Newton step.
__device__ __forceinline__ float fast_rsq(float x) {
  const float x2 = x * 0.5f;
  int32_t i = __float_as_int(x);
  i = 0x5f3759df - (i >> 1);
  float y = __int_as_float(i);
  y = y * (1.5f - (x2 * y * y));
  return y;
}
__device__ __forceinline__ float fast_sqrt(float x) { return 1.0f / fast_rsq(x); }

__global__ void __launch_bounds__(256) k_scatter_points(const PointRec *__restrict__ pr, int64_t n, float *__restrict__ pos) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const float4 p = pr[k].p;
  const int64_t id = __float_as_int(p.w);
  pos[3 * id] = p.x;
  pos[3 * id + 1] = p.y;
  pos[3 * id + 2] = p.z;
}

// The separation estimate's queries: the positions of the picked samples.
__global__ void __launch_bounds__(256) k_pick_points(const float *__restrict__ pos, const int64_t *__restrict__ picks,
                                                     int64_t trials, float *__restrict__ q) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= trials) return;
  const int64_t a = picks[t];
  q[3 * t] = pos[3 * a];
  q[3 * t + 1] = pos[3 * a + 1];
  q[3 * t + 2] = pos[3 * a + 2];
}

// The smallest squared distance from each pick to any other sample: of its two nearest samples
// (sorted by (d^2, index)), the first that is not the pick itself. A coincident sample of lower
// index comes before the pick and is the answer (d^2 = 0), as a scan over all other samples gives.
__global__ void __launch_bounds__(256) k_pick_select(const int64_t *__restrict__ picks, int64_t trials,
                                                     const int32_t *__restrict__ idx2, const float *__restrict__ d2,
                                                     float *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= trials) return;
  out[t] = (int64_t)idx2[2 * t] != picks[t] ? d2[2 * t] : d2[2 * t + 1];
}

template <int STACK, int K>
__global__ void __launch_bounds__(G_BLOCK) k_graph(const BNode *__restrict__ nodes, const PointRec *__restrict__ pr,
                                                   const float *__restrict__ pos, const float *__restrict__ nrm,
                                                   int64_t n, float r0, int max_degree, int min_degree, float scale,
                                                   int32_t *__restrict__ out_idx, float *__restrict__ out_sep,
                                                   int32_t *__restrict__ out_deg) {
  __shared__ int32_t s_ref[STACK * G_BLOCK];
  const int tid = threadIdx.x;
  const int64_t i = (int64_t)blockIdx.x * G_BLOCK + tid;
  if (i >= n) return;
  const float3 p = f3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  const float3 ni = nrm ? f3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]) : f3(0, 0, 0);
  const float m = fmaxf(scale, fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
  const float sl = m * SLACK_REL;
  const float3 qlo = f3(p.x - sl, p.y - sl, p.z - sl), qhi = f3(p.x + sl, p.y + sl, p.z + sl);
  float kd[K];
  int32_t ki[K];
  int count = 0;
  float radius = r0;
  for (int it = 0; it < 3; ++it) {
#pragma unroll
    for (int j = 0; j < K; ++j) { kd[j] = INFINITY; ki[j] = 0x7fffffff; }
    count = 0;
    const float r2 = radius * radius;
    int sp = 0;
    int32_t node = 0;
    while (true) {
      const BNode nd = nodes[node];
      const float lb0 = box_lb(qlo, qhi, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y);
      const float lb1 = box_lb(qlo, qhi, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w);
      int32_t next = -1;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const float lb = c == 0 ? lb0 : lb1;
        const int32_t r = c == 0 ? nd.d.x : nd.d.y;
        const int32_t cnt = c == 0 ? nd.d.z : nd.d.w;
        if (!(lb <= r2)) continue;
        if (r < 0) {
          const int32_t f = ~r;
          for (int32_t e = f; e < f + cnt; ++e) {
            const float4 pe = pr[e].p;
            const int32_t id = __float_as_int(pe.w);
            if (id == (int32_t)i) continue;
            const float d2 = sqn3(sub3(f3(pe.x, pe.y, pe.z), p));
            if (!(d2 <= r2)) continue;
            if (nrm && dot3(ni, f3(nrm[3 * (int64_t)id], nrm[3 * (int64_t)id + 1], nrm[3 * (int64_t)id + 2])) < -0.5f)
              continue;
            ++count;
            const float sep = fast_sqrt(d2);
            if (better(sep, id, kd[K - 1], ki[K - 1])) knn_insert<K>(kd, ki, sep, id);
          }
        } else if (next < 0) {
          next = r;
        } else if (sp < STACK) {  // never full when depth < STACK (checked by the caller)
          s_ref[sp * G_BLOCK + tid] = r;
          ++sp;
        }
      }
      if (next >= 0) { node = next; continue; }
      if (sp == 0) break;
      --sp;
      node = s_ref[sp * G_BLOCK + tid];
    }
    if (count >= min_degree) break;  // numNeighbors() = min(count, max_degree) >= min_degree
    radius *= 2.0f;
  }
  const int deg = count < max_degree ? count : max_degree;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    if (j < max_degree) {
      out_idx[i * max_degree + j] = j < deg ? ki[j] : -1;
      if (out_sep) out_sep[i * max_degree + j] = j < deg ? kd[j] : -1.0f;
    }
  }
  if (out_deg) out_deg[i] = deg;
}

//--------------------------------------------------------------------------------------------------
// Graphs with a surface (PointSet3::updateGraph(surface), the SampleGraph tool's -r): a candidate in
// the ball that passes the identity and normal tests must also be reachable: the ray from the
// sample lifted by 5 * separation along normalize(n_i + n_j), in direction p_j - p_i, must not hit
// the surface before time 1 (NeighborFunctor::isValidNeighbor, PointSet3.hpp:633-641). The rays of
// every candidate are cast together with the any-hit kernel, so a pass over the samples that still
// need it (the radius doubling of findSampleNeighbors) is: count the ball's candidates, fill them,
// cast their rays, keep the max_degree smallest (separation, index) reachable ones.
//--------------------------------------------------------------------------------------------------
// FILL = false: cnt[a] = candidates of sample act[a] (identity and normal tests passed) in its ball;
// FILL = true: writes them at off[a] as (neighbour, separation) with the ray of the reachability test.
template <int STACK, bool FILL>
__global__ void __launch_bounds__(G_BLOCK) k_reach_cand(const BNode *__restrict__ nodes, const PointRec *__restrict__ pr,
                                                        const float *__restrict__ pos, const float *__restrict__ nrm,
                                                        const int32_t *__restrict__ act, int64_t na, float radius,
                                                        float scale, int32_t *__restrict__ cnt,
                                                        const int64_t *__restrict__ off, int32_t *__restrict__ c_nbr,
                                                        float *__restrict__ c_sep, float *__restrict__ c_org,
                                                        float *__restrict__ c_dir) {
  __shared__ int32_t s_ref[STACK * G_BLOCK];
  const int tid = threadIdx.x;
  const int64_t a = (int64_t)blockIdx.x * G_BLOCK + tid;
  if (a >= na) return;
  const int64_t i = act[a];
  const float3 p = f3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  const float3 ni = f3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
  const float m = fmaxf(scale, fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
  const float sl = m * SLACK_REL;
  const float3 qlo = f3(p.x - sl, p.y - sl, p.z - sl), qhi = f3(p.x + sl, p.y + sl, p.z + sl);
  const float r2 = radius * radius;
  int64_t w = FILL ? off[a] : 0;
  int32_t count = 0;
  int sp = 0;
  int32_t node = 0;
  while (true) {
    const BNode nd = nodes[node];
    const float lb0 = box_lb(qlo, qhi, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y);
    const float lb1 = box_lb(qlo, qhi, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w);
    int32_t next = -1;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const float lb = c == 0 ? lb0 : lb1;
      const int32_t r = c == 0 ? nd.d.x : nd.d.y;
      const int32_t cn = c == 0 ? nd.d.z : nd.d.w;
      if (!(lb <= r2)) continue;
      if (r < 0) {
        for (int32_t e = ~r; e < ~r + cn; ++e) {
          const float4 pe = pr[e].p;
          const int32_t id = __float_as_int(pe.w);
          if (id == (int32_t)i) continue;
          const float3 d = sub3(f3(pe.x, pe.y, pe.z), p);
          const float d2 = sqn3(d);
          if (!(d2 <= r2)) continue;
          const float3 nj = f3(nrm[3 * (int64_t)id], nrm[3 * (int64_t)id + 1], nrm[3 * (int64_t)id + 2]);
          if (dot3(ni, nj) < -0.5f) continue;
          if (FILL) {
            const float sep = fast_sqrt(d2);
            // lift_dir = (n_i + n_j).normalized(); origin = p + (LIFT_FACTOR * sep) * lift_dir
            const float3 lift = normalize3(add3(ni, nj));
            const float h = 5.0f * sep;
            c_nbr[w] = id;
            c_sep[w] = sep;
            c_org[3 * w] = p.x + h * lift.x;
            c_org[3 * w + 1] = p.y + h * lift.y;
            c_org[3 * w + 2] = p.z + h * lift.z;
            c_dir[3 * w] = d.x;
            c_dir[3 * w + 1] = d.y;
            c_dir[3 * w + 2] = d.z;
            ++w;
          }
          ++count;
        }
      } else if (next < 0) {
        next = r;
      } else if (sp < STACK) {  // never full when depth < STACK (checked by the caller)
        s_ref[sp * G_BLOCK + tid] = r;
        ++sp;
      }
    }
    if (next >= 0) { node = next; continue; }
    if (sp == 0) break;
    --sp;
    node = s_ref[sp * G_BLOCK + tid];
  }
  if (!FILL) cnt[a] = count;
}

// For each active sample: the max_degree smallest (separation, index) candidates whose rays hit
// nothing; done[a] = the neighbour count reaches min_degree or this is the last radius.
template <int K>
__global__ void __launch_bounds__(256) k_reach_select(const int32_t *__restrict__ act, int64_t na,
                                                      const int64_t *__restrict__ off, const int32_t *__restrict__ c_nbr,
                                                      const float *__restrict__ c_sep, const uint8_t *__restrict__ hit,
                                                      int md, int min_degree, bool last, int32_t *__restrict__ out_idx,
                                                      float *__restrict__ out_sep, int32_t *__restrict__ out_deg,
                                                      uint8_t *__restrict__ again) {
  const int64_t a = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (a >= na) return;
  float kd[K];
  int32_t ki[K];
#pragma unroll
  for (int j = 0; j < K; ++j) { kd[j] = INFINITY; ki[j] = 0x7fffffff; }
  int count = 0;
  for (int64_t c = off[a]; c < off[a + 1]; ++c) {
    if (hit[c]) continue;
    ++count;
    if (better(c_sep[c], c_nbr[c], kd[K - 1], ki[K - 1])) knn_insert<K>(kd, ki, c_sep[c], c_nbr[c]);
  }
  const bool fin = count >= min_degree || last;
  again[a] = fin ? 0 : 1;
  if (!fin) return;
  const int64_t i = act[a];
  const int deg = count < md ? count : md;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    if (j < md) {
      out_idx[i * md + j] = j < deg ? ki[j] : -1;
      if (out_sep) out_sep[i * md + j] = j < deg ? kd[j] : -1.0f;
    }
  }
  if (out_deg) out_deg[i] = deg;
}

// The neighbour sets of every sample with the reachability test against `surface` (a mesh BVH),
// radius r0 doubling up to twice for samples with fewer than min_degree neighbours.
static void reach_neighbours(Bvh *b, const float *pos, const float *d_nrm, const Bvh &surface, int md, int min_degree,
                             float r0, hipStream_t s, int32_t *d_idx, float *d_sep, int32_t *d_deg) {
  const int64_t n = b->n;
  std::vector<int32_t> act(n);
  for (int64_t i = 0; i < n; ++i) act[i] = (int32_t)i;
  const int stack = b->max_depth < 32 ? 32 : 64;
  float radius = r0;
  for (int it = 0; it < 3 && !act.empty(); ++it, radius *= 2.0f) {
    const int64_t na = (int64_t)act.size();
    AsyncBuf d_act(na * 4, s), cnt(na * 4, s), off((na + 1) * 8, s), again(na, s);
    THG_HIP(hipMemcpyAsync(d_act.p, act.data(), na * 4, hipMemcpyHostToDevice, s));
    const uint32_t grid = ceil_div(na, G_BLOCK);
#define THG_RC(S, F, O, N, SP, OR, DI)                                                                               \
  hipLaunchKernelGGL((k_reach_cand<S, F>), dim3(grid), dim3(G_BLOCK), 0, s, b->nodes.as<BNode>(),                    \
                     b->elems.as<PointRec>(), pos, d_nrm, d_act.as<int32_t>(), na, radius, b->scale, cnt.as<int32_t>(), \
                     O, N, SP, OR, DI)
    if (stack == 32) THG_RC(32, false, nullptr, nullptr, nullptr, nullptr, nullptr);
    else THG_RC(64, false, nullptr, nullptr, nullptr, nullptr, nullptr);
    launch_check("k_reach_cand");
    std::vector<int32_t> hc(na);
    THG_HIP(hipMemcpyAsync(hc.data(), cnt.p, na * 4, hipMemcpyDeviceToHost, s));
    THG_HIP(hipStreamSynchronize(s));
    std::vector<int64_t> ho(na + 1, 0);
    for (int64_t a = 0; a < na; ++a) ho[a + 1] = ho[a] + hc[a];
    const int64_t C = ho[na];
    THG_HIP(hipMemcpyAsync(off.p, ho.data(), (na + 1) * 8, hipMemcpyHostToDevice, s));
    AsyncBuf c_nbr(std::max<int64_t>(C, 1) * 4, s), c_sep(std::max<int64_t>(C, 1) * 4, s),
        c_org(std::max<int64_t>(C, 1) * 12, s), c_dir(std::max<int64_t>(C, 1) * 12, s), hit(std::max<int64_t>(C, 1), s);
    if (C > 0) {
      if (stack == 32) THG_RC(32, true, off.as<int64_t>(), c_nbr.as<int32_t>(), c_sep.as<float>(), c_org.as<float>(),
                              c_dir.as<float>());
      else THG_RC(64, true, off.as<int64_t>(), c_nbr.as<int32_t>(), c_sep.as<float>(), c_org.as<float>(),
                  c_dir.as<float>());
      launch_check("k_reach_cand");
      QueryScratch ws;
      launch_ray(surface, c_org.as<float>(), c_dir.as<float>(), C, 1.0f, true, nullptr, nullptr, nullptr,
                 hit.as<uint8_t>(), ws, s);
    }
#undef THG_RC
    const bool last = it == 2;
#define THG_RS(K)                                                                                                      \
  hipLaunchKernelGGL((k_reach_select<K>), dim3(ceil_div(na, 256)), dim3(256), 0, s, d_act.as<int32_t>(), na,         \
                     off.as<int64_t>(), c_nbr.as<int32_t>(), c_sep.as<float>(), hit.as<uint8_t>(), md, min_degree, last, \
                     d_idx, d_sep, d_deg, again.as<uint8_t>())
    if (md <= 4) THG_RS(4); else if (md <= 8) THG_RS(8); else if (md <= 16) THG_RS(16); else THG_RS(32);
#undef THG_RS
    launch_check("k_reach_select");
    std::vector<uint8_t> ha(na);
    THG_HIP(hipMemcpyAsync(ha.data(), again.p, na, hipMemcpyDeviceToHost, s));
    THG_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> next;
    for (int64_t a = 0; a < na; ++a)
      if (ha[a]) next.push_back(act[a]);
    act.swap(next);
  }
}

//--------------------------------------------------------------------------------------------------
// Oversampled graphs: PointSet3::extractOriginalAdjacencies (PointSet3.cpp:158-178). The graph is
// built over the S original samples followed by the dense samples; each original sample's
// neighbours are then the first max_degree other original samples that Dijkstra (from that sample,
// over out-edges = neighbour lists, weights = separations summed in double, ShortestPaths.hpp:
// 287-345) finalises, kept as (float distance, index) pairs in ascending order (DijkstraCallback,
// PointSet3.cpp:140-156; SamplePoint3.hpp:53-57).
//
// One wave per original sample runs the search with its touched vertices in a hash table (vertex,
// distance, frontier slot) and its frontier in an unsorted list, both in LDS; a pop takes the
// frontier's minimum (distance, vertex) with a wave reduction, an improved distance is written in
// place, and the popped vertex's out-edges are relaxed one per lane (they are distinct vertices).
// Finalisation order is (distance, index): Thea's Fibonacci heap decides equal distances, which no
// test can pin. A search that outgrows the LDS tables stops and is run again with tables in global
// memory sized for the whole graph.
//--------------------------------------------------------------------------------------------------
constexpr int EX_HCAP = 2048;  // LDS hash slots (power of two)
constexpr int EX_CCAP = 1024;  // LDS frontier entries

__device__ __forceinline__ uint32_t ex_hash(int32_t v, int hcap) { return ((uint32_t)v * 2654435761u) & (uint32_t)(hcap - 1); }

__device__ __forceinline__ double shfl_d(double x, int src) {
  const int64_t b = __double_as_longlong(x);
  const int32_t lo = __shfl((int32_t)(uint32_t)b, src), hi = __shfl((int32_t)(b >> 32), src);
  return __longlong_as_double(((int64_t)hi << 32) | (uint32_t)lo);
}

// Returns false if the tables overflowed (the caller reruns the source with larger ones).
__device__ bool extract_one(int64_t src, const int32_t *__restrict__ nbr_idx, const float *__restrict__ nbr_sep,
                            const int32_t *__restrict__ deg, int md, int64_t S, int hcap, int ccap, int32_t *hkey,
                            double *hdist, int32_t *hslot, double *cd, int32_t *cv, int32_t *ch, int32_t *s_cnt,
                            int32_t *s_hcnt, int32_t *s_ovf, int32_t *__restrict__ out_idx, float *__restrict__ out_sep,
                            int32_t *__restrict__ out_deg) {
  const int lane = threadIdx.x;
  for (int i = lane; i < hcap; i += WAVE) hkey[i] = -1;
  if (lane == 0) {
    const uint32_t h = ex_hash((int32_t)src, hcap);
    hkey[h] = (int32_t)src;
    hdist[h] = 0.0;
    hslot[h] = 0;
    cd[0] = 0.0;
    cv[0] = (int32_t)src;
    ch[0] = (int32_t)h;
    *s_cnt = 1;
    *s_hcnt = 1;
    *s_ovf = 0;
  }
  __syncthreads();
  int found = 0;
  int32_t res_v = 0x7fffffff;
  double res_d = 0;
  int cnt = 1;
  bool ovf = false;
  while (cnt > 0) {
    // the frontier's minimum (distance, vertex)
    double bd = INFINITY;
    int32_t bv = 0x7fffffff, bs = -1;
    for (int i = lane; i < cnt; i += WAVE) {
      const double d = cd[i];
      const int32_t v = cv[i];
      if (d < bd || (d == bd && v < bv)) { bd = d; bv = v; bs = i; }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const double od = shfl_d(bd, lane ^ off);
      const int32_t ov = __shfl(bv, lane ^ off), os = __shfl(bs, lane ^ off);
      if (od < bd || (od == bd && ov < bv)) { bd = od; bv = ov; bs = os; }
    }
    __syncthreads();  // every lane has read the frontier
    if (lane == 0) {
      const int last = cnt - 1;
      hslot[ch[bs]] = -1;  // finalised
      if (bs != last) {
        cd[bs] = cd[last];
        cv[bs] = cv[last];
        ch[bs] = ch[last];
        hslot[ch[bs]] = bs;
      }
      *s_cnt = last;
    }
    __syncthreads();
    if (bv != (int32_t)src && (int64_t)bv < S) {
      if (lane == found) { res_v = bv; res_d = bd; }
      if (++found >= md) break;
    }
    // relax the out-edges of bv, one per lane (a neighbour list holds distinct samples)
    const int dg = deg[bv];
    if (lane < dg) {
      const int32_t u = nbr_idx[(int64_t)bv * md + lane];
      const double nd = (double)nbr_sep[(int64_t)bv * md + lane] + bd;
      uint32_t h = ex_hash(u, hcap);
      bool placed = false;
      for (int probe = 0; probe < hcap; ++probe) {
        const int32_t k = hkey[h];
        if (k == u) { placed = true; break; }
        if (k == -1) {
          const int32_t prev = atomicCAS(&hkey[h], -1, u);
          if (prev == -1) {  // new vertex: onto the frontier
            if (atomicAdd(s_hcnt, 1) + 1 > (hcap * 3) / 4) *s_ovf = 1;
            const int slot = atomicAdd(s_cnt, 1);
            hdist[h] = nd;
            if (slot < ccap) {
              hslot[h] = slot;
              cd[slot] = nd;
              cv[slot] = u;
              ch[slot] = (int32_t)h;
            } else {
              hslot[h] = -1;
              *s_ovf = 1;
            }
            placed = false;
            break;
          }
          if (prev == u) { placed = true; break; }
        }
        h = (h + 1) & (uint32_t)(hcap - 1);
        if (probe == hcap - 1) *s_ovf = 1;
      }
      if (placed) {  // already touched: improve it unless finalised
        const int32_t sl = hslot[h];
        if (sl >= 0 && nd < hdist[h]) {
          hdist[h] = nd;
          cd[sl] = nd;
        }
      }
    }
    __syncthreads();
    cnt = *s_cnt;
    ovf = *s_ovf != 0;
    if (ovf) break;
    cnt = cnt < ccap ? cnt : ccap;
  }
  if (ovf) return false;
  // (float distance, index) order of the kept pairs: rank each lane's pair among the others
  const float fd = (float)res_d;
  int rank = 0;
  for (int j = 0; j < found; ++j) {
    const float od = __shfl(fd, j);
    const int32_t ov = __shfl(res_v, j);
    if (od < fd || (od == fd && ov < res_v)) ++rank;
  }
  if (lane < md) {
    if (lane < found) {
      out_idx[src * md + rank] = res_v;
      out_sep[src * md + rank] = fd;
    } else {
      out_idx[src * md + lane] = -1;
      out_sep[src * md + lane] = -1.0f;
    }
  }
  if (lane == 0) out_deg[src] = found;
  return true;
}

__global__ void __launch_bounds__(WAVE) k_extract_lds(const int32_t *__restrict__ nbr_idx,
                                                      const float *__restrict__ nbr_sep, const int32_t *__restrict__ deg,
                                                      int md, int64_t S, int32_t *__restrict__ out_idx,
                                                      float *__restrict__ out_sep, int32_t *__restrict__ out_deg,
                                                      int32_t *__restrict__ redo, uint32_t *__restrict__ nredo) {
  __shared__ int32_t hkey[EX_HCAP], hslot[EX_HCAP], cv[EX_CCAP], ch[EX_CCAP];
  __shared__ double hdist[EX_HCAP], cd[EX_CCAP];
  __shared__ int32_t s_cnt, s_hcnt, s_ovf;
  const int64_t src = blockIdx.x;
  if (src >= S) return;
  if (!extract_one(src, nbr_idx, nbr_sep, deg, md, S, EX_HCAP, EX_CCAP, hkey, hdist, hslot, cd, cv, ch, &s_cnt,
                   &s_hcnt, &s_ovf, out_idx, out_sep, out_deg) &&
      threadIdx.x == 0)
    redo[atomicAdd(nredo, 1u)] = (int32_t)src;
}

// Sources whose search outgrew the LDS tables, with tables for the whole graph (hcap >= 2 n slots,
// ccap = n): wave w takes sources w, w + gridDim.x, ... of the list.
__global__ void __launch_bounds__(WAVE) k_extract_global(const int32_t *__restrict__ nbr_idx,
                                                         const float *__restrict__ nbr_sep,
                                                         const int32_t *__restrict__ deg, int md, int64_t S,
                                                         const int32_t *__restrict__ redo, int nredo, int hcap, int ccap,
                                                         int32_t *g_hkey, double *g_hdist, int32_t *g_hslot,
                                                         double *g_cd, int32_t *g_cv, int32_t *g_ch,
                                                         int32_t *__restrict__ out_idx, float *__restrict__ out_sep,
                                                         int32_t *__restrict__ out_deg, int32_t *__restrict__ failed) {
  __shared__ int32_t s_cnt, s_hcnt, s_ovf;
  const int64_t w = blockIdx.x;
  int32_t *hkey = g_hkey + w * hcap, *hslot = g_hslot + w * hcap, *cv = g_cv + w * ccap, *ch = g_ch + w * ccap;
  double *hdist = g_hdist + w * hcap, *cd = g_cd + w * ccap;
  for (int r = (int)w; r < nredo; r += gridDim.x) {
    if (!extract_one(redo[r], nbr_idx, nbr_sep, deg, md, S, hcap, ccap, hkey, hdist, hslot, cd, cv, ch, &s_cnt,
                     &s_hcnt, &s_ovf, out_idx, out_sep, out_deg) &&
        threadIdx.x == 0)
      *failed = 1;
    __syncthreads();
  }
}

// The original samples' neighbours from the graph of all samples (device arrays; S x md outputs).
void extract_original(const int32_t *nbr_idx, const float *nbr_sep, const int32_t *deg, int64_t n, int md, int64_t S,
                      int32_t *out_idx, float *out_sep, int32_t *out_deg, hipStream_t s) {
  if (S <= 0) return;
  AsyncBuf redo(S * 4, s), nredo(4, s);
  THG_HIP(hipMemsetAsync(nredo.p, 0, 4, s));
  hipLaunchKernelGGL(k_extract_lds, dim3((uint32_t)S), dim3(WAVE), 0, s, nbr_idx, nbr_sep, deg, md, S, out_idx,
                     out_sep, out_deg, redo.as<int32_t>(), nredo.as<uint32_t>());
  launch_check("k_extract_lds");
  uint32_t h_nredo = 0;
  THG_HIP(hipMemcpyAsync(&h_nredo, nredo.p, 4, hipMemcpyDeviceToHost, s));
  THG_HIP(hipStreamSynchronize(s));
  if (h_nredo == 0) return;
  // tables for the whole graph: every vertex fits (3/4 load at most) and the frontier cannot exceed n
  int hcap = 1;
  while ((int64_t)hcap * 3 < 4 * (n + 1)) hcap <<= 1;
  const int ccap = (int)(n + 1);
  const size_t per_wave = (size_t)hcap * 16 + (size_t)ccap * 16;
  const int waves = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)h_nredo, 256, (int64_t)((1ull << 31) / per_wave)}));
  AsyncBuf hk((size_t)waves * hcap * 4, s), hd((size_t)waves * hcap * 8, s), hs((size_t)waves * hcap * 4, s),
      c1((size_t)waves * ccap * 8, s), c2((size_t)waves * ccap * 4, s), c3((size_t)waves * ccap * 4, s), failed(4, s);
  THG_HIP(hipMemsetAsync(failed.p, 0, 4, s));
  hipLaunchKernelGGL(k_extract_global, dim3(waves), dim3(WAVE), 0, s, nbr_idx, nbr_sep, deg, md, S,
                     redo.as<int32_t>(), (int)h_nredo, hcap, ccap, hk.as<int32_t>(), hd.as<double>(), hs.as<int32_t>(),
                     c1.as<double>(), c2.as<int32_t>(), c3.as<int32_t>(), out_idx, out_sep, out_deg,
                     failed.as<int32_t>());
  launch_check("k_extract_global");
  int32_t h_failed = 0;
  THG_HIP(hipMemcpyAsync(&h_failed, failed.p, 4, hipMemcpyDeviceToHost, s));
  THG_HIP(hipStreamSynchronize(s));
  require(h_failed == 0, "shortest-path search outgrew its tables");
}

}  // namespace thg

using namespace thg;

template <typename F> static int graph_guarded(F f) {
  try {
    f();
    return THG_OK;
  } catch (const Status &e) {
    set_last_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return THG_E_OOM;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return THG_E_DEVICE;
  }
}

namespace thg {

// The proximity graph of a point BVH's samples into device arrays (nbr_idx, nbr_sep: n x md; degree:
// n); returns the initial separation estimate. normals: device, nullable.
float graph_build(Bvh *b, const float *d_nrm, int md, int64_t trials, uint64_t seed, hipStream_t s, int32_t *d_idx,
                  float *d_sep, int32_t *d_deg, const Bvh *surface) {
  const int64_t n = b->n;
  require(n >= 2, "the proximity graph needs at least two samples (PointSet3.hpp:350-351)");
  require(md >= 1 && md <= 32, "max_degree must be 1..32");
  require(b->max_depth < 64, "point BVH deeper than the 64-entry traversal stack");
  trials = std::min<int64_t>(trials, n);
  // scatter the samples back to input order
  AsyncBuf pos(n * 12, s), picks_d(trials * 8, s), near_d(trials * 4, s);
  hipLaunchKernelGGL(k_scatter_points, dim3(ceil_div(n, 256)), dim3(256), 0, s, b->elems.as<PointRec>(), n,
                     pos.as<float>());
  launch_check("k_scatter_points");
  // separation estimate
  std::vector<int64_t> picks(trials);
  for (int64_t t = 0; t < trials; ++t) {
    uint32_t w[4];
    philox4x32_10((uint32_t)t, (uint32_t)((uint64_t)t >> 32), GRAPH_PICK_STREAM, 0, (uint32_t)seed,
                  (uint32_t)(seed >> 32), w);
    picks[t] = (int64_t)(((uint64_t)w[0] * (uint64_t)n) >> 32);
  }
  THG_HIP(hipMemcpyAsync(picks_d.p, picks.data(), trials * 8, hipMemcpyHostToDevice, s));
  {
    AsyncBuf qp(trials * 12, s), kidx(trials * 8, s), kd2(trials * 8, s);
    hipLaunchKernelGGL(k_pick_points, dim3(ceil_div(trials, 256)), dim3(256), 0, s, pos.as<float>(),
                       picks_d.as<int64_t>(), trials, qp.as<float>());
    launch_check("k_pick_points");
    QueryScratch ws;
    launch_knn(*b, qp.as<float>(), trials, 2, -1.0, 0, kidx.as<int32_t>(), nullptr, kd2.as<float>(), ws, s);
    hipLaunchKernelGGL(k_pick_select, dim3(ceil_div(trials, 256)), dim3(256), 0, s, picks_d.as<int64_t>(), trials,
                       kidx.as<int32_t>(), kd2.as<float>(), near_d.as<float>());
    launch_check("k_pick_select");
  }
  std::vector<float> near(trials);
  THG_HIP(hipMemcpyAsync(near.data(), near_d.p, trials * 4, hipMemcpyDeviceToHost, s));
  THG_HIP(hipStreamSynchronize(s));
  float acc = 0;
  for (int64_t t = 0; t < trials; ++t) acc += near[t];
  const float avg = std::sqrt(acc / (float)trials);
  const float r0 = std::sqrt((float)md) * avg;
  const int min_degree = std::min(md, std::max(4, (int)(0.25 * md)));
  if (surface && d_nrm) {
    reach_neighbours(b, pos.as<float>(), d_nrm, *surface, md, min_degree, r0, s, d_idx, d_sep, d_deg);
    return avg;
  }
  const int stack = b->max_depth < 32 ? 32 : 64;
  const uint32_t grid = ceil_div(n, G_BLOCK);
#define THG_G(S, K)                                                                                                   \
  hipLaunchKernelGGL((k_graph<S, K>), dim3(grid), dim3(G_BLOCK), 0, s, b->nodes.as<BNode>(), b->elems.as<PointRec>(), \
                     pos.as<float>(), d_nrm, n, r0, md, min_degree, b->scale, d_idx, d_sep, d_deg)
#define THG_GK(S) \
  if (md <= 4) THG_G(S, 4); else if (md <= 8) THG_G(S, 8); else if (md <= 16) THG_G(S, 16); else THG_G(S, 32)
  if (stack == 32) { THG_GK(32); } else { THG_GK(64); }
#undef THG_GK
#undef THG_G
  launch_check("k_graph");
  return avg;
}

}  // namespace thg

extern "C" int thg_proximity_graph(thg_bvh points, const float *normals, const thg_graph_opts *opts, int32_t mem,
                                   void *stream, int32_t *nbr_idx, float *nbr_sep, int32_t *degree,
                                   float separation[2]) {
  return graph_guarded([&] {
    require(points != nullptr && nbr_idx != nullptr, "points/nbr_idx is NULL");
    require(mem == THG_MEM_DEVICE || mem == THG_MEM_HOST, "mem must be THG_MEM_DEVICE or THG_MEM_HOST");
    Bvh *b = reinterpret_cast<Bvh *>(points);
    require(b->points, "the proximity graph needs a point BVH");
    THG_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int md = opts && opts->max_degree > 0 ? opts->max_degree : 8;  // PointSet3::Options (PointSet3.hpp:57)
    require(md <= 32, "max_degree must be at most 32");
    const int64_t n = b->n;
    require(n >= 2, "the proximity graph needs at least two samples (PointSet3.hpp:350-351)");
    const int64_t trials = opts && opts->separation_trials > 0 ? opts->separation_trials : 100;
    const uint64_t seed = opts ? opts->seed : 0;
    // samples [0, S) are the original ones, [S, n) the oversampling (PointSet3::addOversampling)
    const int64_t S = opts && opts->num_original > 0 ? opts->num_original : n;
    const Bvh *surface = opts && opts->surface ? reinterpret_cast<const Bvh *>(opts->surface) : nullptr;
    require(!surface || !surface->points, "the reachability surface must be a triangle BVH");
    require(!surface || normals, "the reachability test needs normals (SampleGraph -r requires -n)");
    require(S <= n, "num_original exceeds the number of points");
    const bool dense = S < n;
    AsyncBuf nrm_d;
    const float *d_nrm = normals;
    if (normals && mem == THG_MEM_HOST) {
      nrm_d.ensure(n * 12, s);
      THG_HIP(hipMemcpyAsync(nrm_d.p, normals, n * 12, hipMemcpyHostToDevice, s));
      d_nrm = nrm_d.as<float>();
    }
    // outputs (S rows) in device memory; host outputs and the dense graph go through temporaries
    const bool host = mem == THG_MEM_HOST;
    const bool want_avg = separation != nullptr;
    AsyncBuf idx_tmp, sep_tmp, deg_tmp, g_idx, g_sep, g_deg;
    int32_t *d_idx = nbr_idx;
    float *d_sep = nbr_sep;
    int32_t *d_deg = degree;
    if (host) {
      idx_tmp.ensure((size_t)S * md * 4, s);
      d_idx = idx_tmp.as<int32_t>();
    }
    if ((host && nbr_sep) || ((want_avg || dense) && !nbr_sep)) {
      sep_tmp.ensure((size_t)S * md * 4, s);
      d_sep = sep_tmp.as<float>();
    }
    if ((host && degree) || ((want_avg || dense) && !degree)) {
      deg_tmp.ensure((size_t)S * 4, s);
      d_deg = deg_tmp.as<int32_t>();
    }
    float avg;
    if (dense) {
      g_idx.ensure((size_t)n * md * 4, s);
      g_sep.ensure((size_t)n * md * 4, s);
      g_deg.ensure((size_t)n * 4, s);
      avg = graph_build(b, d_nrm, md, trials, seed, s, g_idx.as<int32_t>(), g_sep.as<float>(), g_deg.as<int32_t>(),
                        surface);
      extract_original(g_idx.as<int32_t>(), g_sep.as<float>(), g_deg.as<int32_t>(), n, md, S, d_idx, d_sep, d_deg, s);
    } else {
      avg = graph_build(b, d_nrm, md, trials, seed, s, d_idx, d_sep, d_deg, surface);
    }
    if (want_avg) {  // updateAverageSeparation over the original samples (PointSet3.hpp:683-697)
      std::vector<float> hs((size_t)S * md);
      std::vector<int32_t> hd(S);
      THG_HIP(hipMemcpyAsync(hs.data(), d_sep, hs.size() * 4, hipMemcpyDeviceToHost, s));
      THG_HIP(hipMemcpyAsync(hd.data(), d_deg, hd.size() * 4, hipMemcpyDeviceToHost, s));
      THG_HIP(hipStreamSynchronize(s));
      float sum = 0;
      int64_t edges = 0;
      for (int64_t i = 0; i < S; ++i) {
        for (int j = 0; j < hd[i]; ++j) sum += hs[(size_t)i * md + j];
        edges += hd[i];
      }
      separation[0] = avg;
      separation[1] = edges > 0 ? sum / (float)edges : 0.0f;
    }
    if (host) {
      THG_HIP(hipMemcpyAsync(nbr_idx, d_idx, (size_t)S * md * 4, hipMemcpyDeviceToHost, s));
      if (nbr_sep) THG_HIP(hipMemcpyAsync(nbr_sep, d_sep, (size_t)S * md * 4, hipMemcpyDeviceToHost, s));
      if (degree) THG_HIP(hipMemcpyAsync(degree, d_deg, (size_t)S * 4, hipMemcpyDeviceToHost, s));
    }
    THG_HIP(hipStreamSynchronize(s));
  });
}

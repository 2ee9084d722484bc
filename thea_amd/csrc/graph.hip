// Sample proximity graph: Thea's PointSet3::updateGraph (Algorithms/PointSet3.hpp:324-386) for a set
// of samples without oversampling and without a surface for the reachability test.
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
                  float *d_sep, int32_t *d_deg) {
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
    AsyncBuf nrm_d;
    const float *d_nrm = normals;
    if (normals && mem == THG_MEM_HOST) {
      nrm_d.ensure(n * 12, s);
      THG_HIP(hipMemcpyAsync(nrm_d.p, normals, n * 12, hipMemcpyHostToDevice, s));
      d_nrm = nrm_d.as<float>();
    }
    // neighbours (host outputs go through device buffers)
    const bool host = mem == THG_MEM_HOST;
    const bool want_avg = separation != nullptr;
    AsyncBuf idx_tmp, sep_tmp, deg_tmp;
    int32_t *d_idx = nbr_idx;
    float *d_sep = nbr_sep;
    int32_t *d_deg = degree;
    if (host) {
      idx_tmp.ensure((size_t)n * md * 4, s);
      d_idx = idx_tmp.as<int32_t>();
    }
    if ((host && nbr_sep) || (want_avg && !nbr_sep)) {
      sep_tmp.ensure((size_t)n * md * 4, s);
      d_sep = sep_tmp.as<float>();
    }
    if ((host && degree) || (want_avg && !degree)) {
      deg_tmp.ensure((size_t)n * 4, s);
      d_deg = deg_tmp.as<int32_t>();
    }
    const float avg = graph_build(b, d_nrm, md, trials, seed, s, d_idx, d_sep, d_deg);
    if (want_avg) {
      std::vector<float> hs((size_t)n * md);
      std::vector<int32_t> hd(n);
      THG_HIP(hipMemcpyAsync(hs.data(), d_sep, hs.size() * 4, hipMemcpyDeviceToHost, s));
      THG_HIP(hipMemcpyAsync(hd.data(), d_deg, hd.size() * 4, hipMemcpyDeviceToHost, s));
      THG_HIP(hipStreamSynchronize(s));
      float sum = 0;
      int64_t edges = 0;
      for (int64_t i = 0; i < n; ++i) {
        for (int j = 0; j < hd[i]; ++j) sum += hs[(size_t)i * md + j];
        edges += hd[i];
      }
      separation[0] = avg;
      separation[1] = edges > 0 ? sum / (float)edges : 0.0f;
    }
    if (host) {
      THG_HIP(hipMemcpyAsync(nbr_idx, d_idx, (size_t)n * md * 4, hipMemcpyDeviceToHost, s));
      if (nbr_sep) THG_HIP(hipMemcpyAsync(nbr_sep, d_sep, (size_t)n * md * 4, hipMemcpyDeviceToHost, s));
      if (degree) THG_HIP(hipMemcpyAsync(degree, d_deg, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    }
    THG_HIP(hipStreamSynchronize(s));
  });
}

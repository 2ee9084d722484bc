// Separation-based sampling: Thea's FurthestPointSampling::subsample with geodesic distances
// (Algorithms/FurthestPointSampling.cpp:47-125) and MeshSampler::sampleEvenlyBySeparation
// (Algorithms/MeshSampler.hpp:270-329).
//
// Thea builds the proximity graph of the points (PointSet3, default options), picks point 0, and
// then repeatedly runs a multi-source Dijkstra from every point picked so far (edge length: the
// graph's float separation, path lengths summed in double as d(u) + w) and picks the vertex its
// callback sees last: an unreachable vertex if any (the last in index order, ShortestPaths.hpp:
// 350-359), else a vertex of maximum distance. It stops early when that vertex is already picked.
//
// Here the graph comes from the proximity-graph kernel and the selection runs in one workgroup
// that keeps, across picks, each vertex's distance to the picked set. Adding a pick relaxes only
// the vertices whose distance drops (frontier rounds of Bellman-Ford with atomicMin on the double's
// bits, which order like the values for non-negative doubles). Since d(u) + w >= d(u) in floating
// point, the distances are the least fixed point of d(v) = min over in-edges of d(u) + w, the
// values Dijkstra computes, whatever the relaxation order. A max-tournament tree over (distance,
// index) then names the furthest vertex; only the ancestors of vertices whose distance changed
// are recomputed. Among vertices at exactly the same maximum distance the highest index wins:
// Thea's choice there depends on its Fibonacci heap's order for equal keys.
#include <cmath>

#include "common.hpp"
#include "device_types.hpp"
#include "kernels_common.hpp"
#include "query.hpp"

namespace thg {

constexpr int FPS_THREADS = 1024;
constexpr unsigned long long FPS_INF_BITS = 0x7FF0000000000000ull;  // +infinity: not reached

// Shared mutable state lives in global memory and is accessed with device-coherent atomics, so
// the workgroup barriers order it for every thread of the block.
__device__ __forceinline__ unsigned long long ld_u64(const unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_u64(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t ld_i32(const int32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_i32(int32_t *p, int32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct FpsWork {
  unsigned long long *dist;  // n: bits of the distance to the picked set
  int32_t *stamp;            // n: last round that queued the vertex
  int32_t *cstamp;           // n: last pick whose relaxation listed the vertex as changed
  int32_t *qa, *qb;          // n each: frontiers
  int32_t *changed;          // n
  int32_t *tree;             // 2P: winners of the max tournament; leaf v at P + v
  int32_t *status;           // [0] = 1 if a relaxation exceeded its round bound
};

// The larger of (dist, index) of vertices a and b (-1: none).
__device__ __forceinline__ int32_t fps_winner(const unsigned long long *dist, int32_t a, int32_t b) {
  if (a < 0) return b;
  if (b < 0) return a;
  const unsigned long long da = ld_u64(dist + a), db = ld_u64(dist + b);
  return (da > db || (da == db && a > b)) ? a : b;
}

// One workgroup. nbr/sep: n x md neighbour lists (-1 padded) of the directed proximity graph.
__global__ void __launch_bounds__(FPS_THREADS) k_fps(int32_t n, int32_t md, const int32_t *__restrict__ nbr,
                                                     const float *__restrict__ sep, int32_t k, int32_t P,
                                                     int32_t levels, FpsWork w, int64_t *__restrict__ selected,
                                                     int64_t *__restrict__ count) {
  __shared__ int32_t s_qn, s_nn, s_cn, s_src;
  const int t = threadIdx.x;
  for (int32_t v = t; v < n; v += FPS_THREADS) {
    st_u64(w.dist + v, FPS_INF_BITS);
    st_i32(w.stamp + v, -1);
    st_i32(w.cstamp + v, -1);
  }
  for (int32_t i = t; i < P; i += FPS_THREADS) st_i32(w.tree + P + i, i < n ? i : -1);
  if (t == 0) s_src = 0;  // Thea's first pick (FurthestPointSampling.cpp:79-83)
  __syncthreads();
  // Shortest paths are simple (a cycle adds >= 0), so a relaxation settles within n rounds.
  const int32_t round_limit = n + 8;
  int32_t round = 0, got = 0;
  for (int32_t it = 0; it < k; ++it) {
    const int32_t src = s_src;
    __syncthreads();  // everyone has read s_src
    if (t == 0) {
      selected[it] = src;
      st_u64(w.dist + src, 0ull);
      st_i32(w.qa, src);
      st_i32(w.changed, src);
      st_i32(w.cstamp + src, it);
      s_qn = 1;
      s_cn = 1;
    }
    got = it + 1;
    __syncthreads();
    int32_t *cur = w.qa, *nxt = w.qb;
    int32_t rounds = 0;
    bool over = false;
    for (;;) {
      const int32_t qn = s_qn;
      if (qn == 0) break;
      if (++rounds > round_limit) {
        over = true;
        break;
      }
      ++round;
      __syncthreads();  // everyone has read s_qn
      if (t == 0) s_nn = 0;
      __syncthreads();
      for (int32_t j = t; j < qn; j += FPS_THREADS) {
        const int32_t u = ld_i32(cur + j);
        const double du = __longlong_as_double((long long)ld_u64(w.dist + u));
        const int32_t *nu = nbr + (int64_t)u * md;
        const float *su = sep + (int64_t)u * md;
        for (int e = 0; e < md; ++e) {
          const int32_t v = nu[e];
          if (v < 0) break;
          const double cand = (double)su[e] + du;  // ShortestPaths.hpp:306 (nbr_dist + data->dist)
          const unsigned long long cb = (unsigned long long)__double_as_longlong(cand);
          if (cb >= ld_u64(w.dist + v)) continue;
          const unsigned long long old = atomicMin(w.dist + v, cb);
          if (old <= cb) continue;
          if (atomicExch(w.stamp + v, round) != round) st_i32(nxt + atomicAdd(&s_nn, 1), v);
          if (it > 0 && atomicExch(w.cstamp + v, it) != it) st_i32(w.changed + atomicAdd(&s_cn, 1), v);
        }
      }
      __syncthreads();
      if (t == 0) s_qn = s_nn;
      int32_t *tmp = cur;
      cur = nxt;
      nxt = tmp;
      __syncthreads();
    }
    if (over) {
      if (t == 0) st_i32(w.status, 1);
      break;
    }
    if (it + 1 == k) break;
    if (it == 0) {  // every vertex's distance is new: build the whole tree
      for (int32_t lvl = P >> 1; lvl >= 1; lvl >>= 1) {
        for (int32_t i = lvl + t; i < 2 * lvl; i += FPS_THREADS)
          st_i32(w.tree + i, fps_winner(w.dist, ld_i32(w.tree + 2 * i), ld_i32(w.tree + 2 * i + 1)));
        __syncthreads();
      }
    } else {  // replay the ancestors of the changed vertices, level by level
      const int32_t cn = s_cn;
      for (int32_t l = 1; l <= levels; ++l) {
        for (int32_t c = t; c < cn; c += FPS_THREADS) {
          const int32_t node = (P + ld_i32(w.changed + c)) >> l;
          st_i32(w.tree + node, fps_winner(w.dist, ld_i32(w.tree + 2 * node), ld_i32(w.tree + 2 * node + 1)));
        }
        __syncthreads();
      }
    }
    const int32_t best = ld_i32(w.tree + 1);
    // The furthest vertex is at distance 0, so it is already picked (or coincides with a pick):
    // Thea stops with the picks so far (FurthestPointSampling.cpp:87-91).
    if (ld_u64(w.dist + best) == 0ull) break;
    if (t == 0) s_src = best;
    __syncthreads();
  }
  if (t == 0) *count = got;
}

__global__ void __launch_bounds__(256) k_gather_samples(int64_t n, const int64_t *__restrict__ sel,
                                                        const float *__restrict__ pos, const float *__restrict__ nrm,
                                                        const int32_t *__restrict__ tri, float *__restrict__ out_pos,
                                                        float *__restrict__ out_nrm, int32_t *__restrict__ out_tri) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int64_t j = sel[i];
  for (int c = 0; c < 3; ++c) {
    if (out_pos) out_pos[3 * i + c] = pos[3 * j + c];
    if (out_nrm) out_nrm[3 * i + c] = nrm[3 * j + c];
  }
  if (out_tri) out_tri[i] = tri[j];
}

// Furthest-point selection over device points: builds the point BVH and the proximity graph, runs
// k_fps, and returns the number of picks (<= k) written to d_sel.
static int64_t fps_select(Ctx *c, const float *d_pts, int64_t n, int64_t k, const thg_graph_opts *go, hipStream_t s,
                          int64_t *d_sel) {
  if (k <= 0) return 0;
  if (n == 1) {
    const int64_t zero = 0;
    THG_HIP(hipMemcpyAsync(d_sel, &zero, 8, hipMemcpyHostToDevice, s));
    THG_HIP(hipStreamSynchronize(s));
    return 1;
  }
  const int md = go && go->max_degree > 0 ? go->max_degree : 8;  // PointSet3::Options (PointSet3.hpp:57)
  require(md <= 32, "max_degree must be at most 32");
  const int64_t trials = go && go->separation_trials > 0 ? go->separation_trials : 100;
  const uint64_t seed = go ? go->seed : 0;
  Bvh b;
  b.ctx = c;
  build_bvh(b, d_pts, nullptr, n, true, 12, s);
  AsyncBuf idx((size_t)n * md * 4, s), sepb((size_t)n * md * 4, s), deg((size_t)n * 4, s);
  graph_build(&b, nullptr, md, trials, seed, s, idx.as<int32_t>(), sepb.as<float>(), deg.as<int32_t>());
  int32_t P = 1, levels = 0;
  while (P < n) {
    P <<= 1;
    ++levels;
  }
  AsyncBuf dist((size_t)n * 8, s), stamp((size_t)n * 4, s), cstamp((size_t)n * 4, s), qa((size_t)n * 4, s),
      qb((size_t)n * 4, s), changed((size_t)n * 4, s), tree((size_t)P * 8, s), small(16, s);
  THG_HIP(hipMemsetAsync(small.p, 0, 16, s));
  FpsWork w;
  w.dist = dist.as<unsigned long long>();
  w.stamp = stamp.as<int32_t>();
  w.cstamp = cstamp.as<int32_t>();
  w.qa = qa.as<int32_t>();
  w.qb = qb.as<int32_t>();
  w.changed = changed.as<int32_t>();
  w.tree = tree.as<int32_t>();
  w.status = small.as<int32_t>();
  int64_t *d_count = reinterpret_cast<int64_t *>(small.as<char>() + 8);
  KernelTimer kt(c, s, "fps");
  hipLaunchKernelGGL(k_fps, dim3(1), dim3(FPS_THREADS), 0, s, (int32_t)n, (int32_t)md, idx.as<int32_t>(),
                     sepb.as<float>(), (int32_t)k, P, levels, w, d_sel, d_count);
  launch_check("k_fps");
  kt.stop();
  int64_t h[2];
  THG_HIP(hipMemcpyAsync(h, small.p, 16, hipMemcpyDeviceToHost, s));
  THG_HIP(hipStreamSynchronize(s));
  require((int32_t)h[0] == 0, "furthest-point relaxation exceeded its round bound");
  return h[1];
}

}  // namespace thg

using namespace thg;

template <typename F> static int fps_guarded(F f) {
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

extern "C" int thg_furthest_point_sampling(thg_ctx ctx, const float *points, int64_t n, int64_t k,
                                           const thg_graph_opts *graph_opts, int32_t mem, void *stream,
                                           int64_t *selected, int64_t *count) {
  return fps_guarded([&] {
    require(ctx != nullptr && count != nullptr, "ctx/count is NULL");
    require(mem == THG_MEM_DEVICE || mem == THG_MEM_HOST, "mem must be THG_MEM_DEVICE or THG_MEM_HOST");
    require(k >= 0, "can't sample a negative number of points");
    require(n >= k, "can't subsample more points than there are (FurthestPointSampling.cpp:51-54)");
    require(n < (int64_t)1 << 30, "at most 2^30 points");
    require(k == 0 || (points && selected), "points/selected is NULL");
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    THG_HIP(hipSetDevice(c->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    *count = 0;
    if (k == 0) return;
    Stage st(s, mem == THG_MEM_HOST);
    const float *d_pts = st.in(points, (size_t)n * 3);
    int64_t *d_sel = st.out(selected, (size_t)k);
    const int64_t got = fps_select(c, d_pts, n, k, graph_opts, s, d_sel);
    st.finish();
    *count = got;
  });
}

extern "C" int thg_sample_separation(thg_mesh mesh, int64_t n, float oversampling_factor, uint64_t seed,
                                     uint64_t stream_id, const thg_graph_opts *graph_opts, float *pos, float *normal,
                                     int32_t *tri, int64_t *n_out, int32_t mem, void *stream) {
  return fps_guarded([&] {
    require(mesh != nullptr && n_out != nullptr, "mesh/n_out is NULL");
    require(mem == THG_MEM_DEVICE || mem == THG_MEM_HOST, "mem must be THG_MEM_DEVICE or THG_MEM_HOST");
    require(n >= 0, "negative sample count");
    require(n == 0 || pos, "pos is NULL");
    Mesh *m = reinterpret_cast<Mesh *>(mesh);
    THG_HIP(hipSetDevice(m->ctx->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    *n_out = 0;
    if (n == 0) return;
    // MeshSampler.hpp:295-298: a negative factor selects the default 3; the count is computed in
    // the sampler's Real (float) arithmetic.
    const float factor = oversampling_factor < 0 ? 3.0f : oversampling_factor;
    const int64_t over = (int64_t)std::ceil(factor * (float)n);
    require(over >= n, "the oversampling must hold at least the requested number of samples");
    require(over < (int64_t)1 << 30, "at most 2^30 oversamples");
    mesh_cdf(*m, s);
    if (!(m->nt > 0 && m->total_area > 0)) return;  // no oversampling: Thea returns 0 (MeshSampler.hpp:303-307)
    AsyncBuf opos((size_t)over * 12, s), onrm, otri, sel((size_t)n * 8, s);
    if (normal) onrm.ensure((size_t)over * 12, s);
    if (tri) otri.ensure((size_t)over * 4, s);
    launch_sample(*m, over, seed, stream_id, opos.as<float>(), normal ? onrm.as<float>() : nullptr,
                  tri ? otri.as<int32_t>() : nullptr, s);
    const int64_t got = fps_select(m->ctx, opos.as<float>(), over, n, graph_opts, s, sel.as<int64_t>());
    if (got < n) return;  // Thea returns 0 samples (MeshSampler.hpp:314-318)
    Stage st(s, mem == THG_MEM_HOST);
    float *dp = st.out(pos, (size_t)n * 3), *dn = st.out(normal, (size_t)n * 3);
    int32_t *dt = st.out(tri, (size_t)n);
    hipLaunchKernelGGL(k_gather_samples, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, sel.as<int64_t>(),
                       opos.as<float>(), normal ? onrm.as<float>() : nullptr, tri ? otri.as<int32_t>() : nullptr, dp,
                       dn, dt);
    launch_check("k_gather_samples");
    st.finish();
    if (mem == THG_MEM_DEVICE) THG_HIP(hipStreamSynchronize(s));
    *n_out = n;
  });
}

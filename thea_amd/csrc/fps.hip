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
// bits, which order like the values for non-negative doubles; one thread per frontier edge).
// Since d(u) + w >= d(u) in floating point, the distances are the least fixed point of
// d(v) = min over in-edges of d(u) + w, the values Dijkstra computes, whatever the relaxation
// order. A 64-ary max-tournament over (distance, index), each node reduced by one wave, then names
// the furthest vertex; only the ancestors of vertices whose distance changed are recomputed. Among
// vertices at exactly the same maximum distance the highest index wins: Thea's choice there
// depends on its Fibonacci heap's order for equal keys.
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

constexpr int FPS_FAN = 64;  // children per tournament node: one wave reduces a node

struct FpsWork {
  unsigned long long *dist;  // n: bits of the distance to the picked set
  int32_t *stamp;            // n: last relaxation round that improved the vertex
  int32_t *qa, *qb;          // n each: frontiers
  int32_t *changed;          // n: vertices improved by the current pick
  int32_t *win;              // tournament winners, level l at win + woff[l] (level 0: the vertices)
  int32_t *wstamp;           // same layout: last pick that marked the node dirty
  int32_t *da, *db;          // dirty node lists (level 1 size)
  int32_t *status;           // [0] = 1 if a relaxation exceeded its round bound
};

struct FpsLevels {
  int32_t count;                  // levels above the vertices; the root is level `count`
  int32_t size[8];                // nodes per level (size[0] = n)
  int32_t off[8];                 // offsets into win / wstamp (level >= 1)
};

// Max of (distance bits, vertex) over the wave, -1 vertices losing to any real one.
__device__ __forceinline__ void wave_argmax(unsigned long long &key, int32_t &idx) {
  for (int o = 32; o >= 1; o >>= 1) {
    const unsigned long long k2 = __shfl_xor(key, o, WAVE);
    const int32_t i2 = __shfl_xor(idx, o, WAVE);
    if (k2 > key || (k2 == key && i2 > idx)) {
      key = k2;
      idx = i2;
    }
  }
}

// Winner of node `node` at level l >= 1 from its FPS_FAN children (one wave, all lanes).
__device__ __forceinline__ int32_t fps_node(const FpsWork &w, const FpsLevels &L, int l, int32_t node) {
  const int32_t c = node * FPS_FAN + lane_id();
  int32_t idx = -1;
  if (c < L.size[l - 1]) idx = l == 1 ? c : ld_i32(w.win + L.off[l - 1] + c);
  unsigned long long key = idx >= 0 ? ld_u64(w.dist + idx) : 0ull;
  wave_argmax(key, idx);
  return idx;
}

// One workgroup. nbr/sep: n x md neighbour lists (-1 padded) of the directed proximity graph.
__global__ void __launch_bounds__(FPS_THREADS) k_fps(int32_t n, int32_t md, const int32_t *__restrict__ nbr,
                                                     const float *__restrict__ sep, int32_t k, FpsLevels L, FpsWork w,
                                                     int64_t *__restrict__ selected, int64_t *__restrict__ count) {
  __shared__ int32_t s_qn, s_nn, s_cn, s_dn, s_src;
  const int t = threadIdx.x, wave = t / WAVE, nwaves = FPS_THREADS / WAVE;
  for (int32_t v = t; v < n; v += FPS_THREADS) {
    st_u64(w.dist + v, FPS_INF_BITS);
    st_i32(w.stamp + v, -1);
  }
  for (int l = 1; l <= L.count; ++l)
    for (int32_t i = t; i < L.size[l]; i += FPS_THREADS) st_i32(w.wstamp + L.off[l] + i, -1);
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
      s_qn = 1;
      s_cn = 1;
    }
    got = it + 1;
    const int32_t first_round = round + 1;  // rounds of this pick are >= first_round
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
      // one thread per (frontier vertex, edge)
      const int64_t work = (int64_t)qn * md;
      for (int64_t x = t; x < work; x += FPS_THREADS) {
        const int32_t j = (int32_t)(x / md), e = (int32_t)(x - (int64_t)j * md);
        const int32_t u = ld_i32(cur + j);
        const int32_t v = nbr[(int64_t)u * md + e];
        if (v < 0) continue;
        const double du = __longlong_as_double((long long)ld_u64(w.dist + u));
        const double cand = (double)sep[(int64_t)u * md + e] + du;  // ShortestPaths.hpp:306
        const unsigned long long cb = (unsigned long long)__double_as_longlong(cand);
        const unsigned long long old = atomicMin(w.dist + v, cb);
        if (old <= cb) continue;
        const int32_t prev = atomicExch(w.stamp + v, round);
        if (prev != round) st_i32(nxt + atomicAdd(&s_nn, 1), v);
        if (it > 0 && prev < first_round) st_i32(w.changed + atomicAdd(&s_cn, 1), v);
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
    if (it == 0) {  // every distance is new: reduce every node, level by level
      for (int l = 1; l <= L.count; ++l) {
        for (int32_t node = wave; node < L.size[l]; node += nwaves) {
          const int32_t win = fps_node(w, L, l, node);
          if (lane_id() == 0) st_i32(w.win + L.off[l] + node, win);
        }
        __syncthreads();
      }
    } else {  // only the ancestors of changed vertices: dirty lists, deduplicated per pick
      if (t == 0) s_dn = 0;
      __syncthreads();
      const int32_t cn = s_cn;
      for (int32_t c = t; c < cn; c += FPS_THREADS) {
        const int32_t node = ld_i32(w.changed + c) / FPS_FAN;
        if (atomicExch(w.wstamp + L.off[1] + node, it) != it) st_i32(w.da + atomicAdd(&s_dn, 1), node);
      }
      __syncthreads();
      int32_t *dcur = w.da, *dnext = w.db;
      for (int l = 1; l <= L.count; ++l) {
        const int32_t dn = s_dn;
        __syncthreads();  // everyone has read s_dn
        if (t == 0) s_dn = 0;
        __syncthreads();
        for (int32_t q = wave; q < dn; q += nwaves) {
          const int32_t node = ld_i32(dcur + q);
          const int32_t win = fps_node(w, L, l, node);
          if (lane_id() == 0) {
            st_i32(w.win + L.off[l] + node, win);
            if (l < L.count) {
              const int32_t parent = node / FPS_FAN;
              if (atomicExch(w.wstamp + L.off[l + 1] + parent, it) != it) st_i32(dnext + atomicAdd(&s_dn, 1), parent);
            }
          }
        }
        __syncthreads();
        int32_t *tmp = dcur;
        dcur = dnext;
        dnext = tmp;
      }
    }
    const int32_t best = ld_i32(w.win + L.off[L.count]);
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
  FpsLevels L{};
  L.size[0] = (int32_t)n;
  int32_t total = 0;
  while (L.size[L.count] > 1 || L.count == 0) {
    require(L.count < 7, "too many points for the furthest-point tournament");
    ++L.count;
    L.size[L.count] = (L.size[L.count - 1] + FPS_FAN - 1) / FPS_FAN;
    L.off[L.count] = total;
    total += L.size[L.count];
  }
  const size_t n1 = (size_t)L.size[1];
  AsyncBuf dist((size_t)n * 8, s), stamp((size_t)n * 4, s), qa((size_t)n * 4, s), qb((size_t)n * 4, s),
      changed((size_t)n * 4, s), win((size_t)total * 4, s), wstamp((size_t)total * 4, s), da(n1 * 4, s),
      db(n1 * 4, s), small(16, s);
  THG_HIP(hipMemsetAsync(small.p, 0, 16, s));
  FpsWork w;
  w.dist = dist.as<unsigned long long>();
  w.stamp = stamp.as<int32_t>();
  w.qa = qa.as<int32_t>();
  w.qb = qb.as<int32_t>();
  w.changed = changed.as<int32_t>();
  w.win = win.as<int32_t>();
  w.wstamp = wstamp.as<int32_t>();
  w.da = da.as<int32_t>();
  w.db = db.as<int32_t>();
  w.status = small.as<int32_t>();
  int64_t *d_count = reinterpret_cast<int64_t *>(small.as<char>() + 8);
  KernelTimer kt(c, s, "fps");
  hipLaunchKernelGGL(k_fps, dim3(1), dim3(FPS_THREADS), 0, s, (int32_t)n, (int32_t)md, idx.as<int32_t>(),
                     sepb.as<float>(), (int32_t)k, L, w, d_sel, d_count);
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

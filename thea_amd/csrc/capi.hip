// C ABI of libthea_gpu (declared in include/thea_gpu.h): argument checking, host/device staging,
// status codes. Every entry point catches failures and reports them through thg_last_error().
#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "kernels_common.hpp"
#include "query.hpp"
#include "query64.hpp"

namespace thg {

static thread_local std::string g_last_error;

void set_last_error(const std::string &s) { g_last_error = s; }

int device_cus(int device) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 256;
  return v > 0 ? v : 256;
}

template <typename F> static int guarded(F f) {
  try {
    f();
    return THG_OK;
  } catch (const Status &s) {
    set_last_error(s.msg);
    return s.code;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return THG_E_OOM;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return THG_E_DEVICE;
  }
}

static hipStream_t as_stream(void *s) { return static_cast<hipStream_t>(s); }

static void set_device(Ctx *c) { THG_HIP(hipSetDevice(c->device)); }

static void check_mem(int32_t mem) { require(mem == THG_MEM_DEVICE || mem == THG_MEM_HOST, "mem must be THG_MEM_DEVICE or THG_MEM_HOST"); }

__global__ void k_check_tris(const int32_t *__restrict__ tris, int64_t n3, int64_t nv, int *__restrict__ bad) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n3) {
    int32_t v = tris[i];
    if (v < 0 || v >= nv) atomicOr(bad, 1);
  }
}

// Fails with THG_E_INVALID unless every index of the nt triangles (device memory) is in [0, nv).
void check_tri_indices(const int32_t *tris, int64_t nt, int64_t nv, hipStream_t s) {
  if (nt <= 0) return;
  AsyncBuf bad(sizeof(int), s);
  THG_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_check_tris, dim3(ceil_div(nt * 3, 256)), dim3(256), 0, s, tris, nt * 3, nv, bad.as<int>());
  launch_check("k_check_tris");
  int hb = 0;
  THG_HIP(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  THG_HIP(hipStreamSynchronize(s));
  require(hb == 0, "triangle vertex index out of range");
}

// Sort memory of thg_closest_point_noalloc: the spatial order of a batch of up to
// THG_NOALLOC_MAX_QUERIES queries (keys, values and their ping-pong copies, 4 words per query,
// then the radix histograms: 256 words per 4096-key tile plus 256 totals), as
// sort_scratch_words(THG_NOALLOC_MAX_QUERIES) counts them. Every word the sort reads is written
// earlier in the same call, so nothing needs clearing between calls. 68.2 MB per device, held from
// the library's load on that device until the process ends.
constexpr size_t NOALLOC_WORDS = 4 * (size_t)THG_NOALLOC_MAX_QUERIES +
                                 256 * (((size_t)THG_NOALLOC_MAX_QUERIES + 4095) / 4096) + 256;
static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

constexpr int NOALLOC_DEVICES = 64;
static std::mutex g_noalloc_mu;
static uint32_t *g_noalloc_ptr[NOALLOC_DEVICES];  // each device's own address of `scratch`

// Looks up the current device's copy of `scratch` once per device (thg_ctx_create calls it with
// `device` current). Host code uses only the address found here, never the array's name.
static void noalloc_init(int device) {
  require(device >= 0 && device < NOALLOC_DEVICES, "device index beyond the no-allocation table");
  std::lock_guard<std::mutex> g(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return;
  void *p = nullptr;
  THG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  require(p != nullptr, "no device address for the no-allocation sort memory");
  g_noalloc_ptr[device] = static_cast<uint32_t *>(p);
}

static uint32_t *noalloc_ptr(int device) {
  std::lock_guard<std::mutex> g(g_noalloc_mu);
  uint32_t *p = device >= 0 && device < NOALLOC_DEVICES ? g_noalloc_ptr[device] : nullptr;
  require(p != nullptr, "no-allocation sort memory not set up on this device (create a context first)");
  return p;
}

}  // namespace thg

using namespace thg;

extern "C" {

const char *thg_last_error(void) { return g_last_error.c_str(); }

const char *thg_version(void) { return "thea_gpu 0.1 (gfx950)"; }

int thg_ctx_create(int32_t device, thg_ctx *out) {
  return guarded([&] {
    require(out != nullptr, "out is NULL");
    int count = 0;
    THG_HIP(hipGetDeviceCount(&count));
    require(device >= 0 && device < count, "device index out of range");
    THG_HIP(hipSetDevice(device));
    // Stream-ordered scratch (hipMallocAsync) comes from the device's default pool; keep freed
    // blocks in the pool instead of returning them to the driver at every synchronisation, so
    // repeated calls reuse memory rather than allocate it again.
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
      uint64_t keep = UINT64_MAX;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    (void)hipGetLastError();
    noalloc_init(device);
    auto *c = new Ctx();
    c->device = device;
    c->num_cus = device_cus(device);
    *out = reinterpret_cast<thg_ctx>(c);
  });
}

int thg_ctx_destroy(thg_ctx ctx) {
  return guarded([&] { delete reinterpret_cast<Ctx *>(ctx); });
}

int thg_ctx_set_timing(thg_ctx ctx, int32_t enable) {
  return guarded([&] {
    require(ctx != nullptr, "ctx is NULL");
    reinterpret_cast<Ctx *>(ctx)->timing = enable != 0;
  });
}

int thg_ctx_kernel_time(thg_ctx ctx, const char *kernel, double *total_ms, int64_t *launches, int32_t reset) {
  return guarded([&] {
    require(ctx && kernel && total_ms && launches, "NULL argument");
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    set_device(c);
    std::lock_guard<std::mutex> g(c->timing_lock);
    double tot = 0;
    int64_t cnt = 0;
    for (auto &t : c->timed) {
      if (t.kernel != kernel) continue;
      THG_HIP(hipEventSynchronize(t.stop));
      float ms = 0;
      THG_HIP(hipEventElapsedTime(&ms, t.start, t.stop));
      tot += ms;
      ++cnt;
    }
    *total_ms = tot;
    *launches = cnt;
    if (reset) {
      std::vector<TimedLaunch> keep;
      for (auto &t : c->timed) {
        if (t.kernel != kernel) { keep.push_back(t); continue; }
        (void)hipEventSynchronize(t.stop);
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
      }
      c->timed.swap(keep);
    }
  });
}

int thg_mesh_create(thg_ctx ctx, const float *verts, int64_t nv, const int32_t *tris, int64_t nt, int32_t mem,
                    thg_mesh *out) {
  return guarded([&] {
    require(ctx && out, "ctx/out is NULL");
    require(nv >= 0 && nt >= 0, "negative counts");
    require((nv == 0 || verts) && (nt == 0 || tris), "NULL array with nonzero count");
    require(nt <= (int64_t)0x7ffffff0, "too many triangles for int32 indices");
    check_mem(mem);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    set_device(c);
    std::unique_ptr<Mesh> m(new Mesh());
    m->ctx = c;
    m->nv = nv;
    m->nt = nt;
    hipMemcpyKind kind = mem == THG_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    m->verts.alloc(nv * 3 * sizeof(float));
    m->tris.alloc(nt * 3 * sizeof(int32_t));
    if (nv) THG_HIP(hipMemcpy(m->verts.p, verts, nv * 3 * sizeof(float), kind));
    if (nt) {
      THG_HIP(hipMemcpy(m->tris.p, tris, nt * 3 * sizeof(int32_t), kind));
      check_tri_indices(m->tris.as<int32_t>(), nt, nv, 0);
    }
    *out = reinterpret_cast<thg_mesh>(m.release());
  });
}

int thg_mesh_destroy(thg_mesh mesh) {
  return guarded([&] { delete reinterpret_cast<Mesh *>(mesh); });
}

int64_t thg_mesh_num_triangles(thg_mesh mesh) { return mesh ? reinterpret_cast<Mesh *>(mesh)->nt : -1; }

int thg_mesh_total_area(thg_mesh mesh, double *area) {
  return guarded([&] {
    require(mesh && area, "mesh/area is NULL");
    Mesh *m = reinterpret_cast<Mesh *>(mesh);
    set_device(m->ctx);
    mesh_cdf(*m, 0);
    *area = m->total_area;
  });
}

static int leaf_from(const thg_build_opts *o) {
  int l = (o && o->leaf_size > 0) ? o->leaf_size : 4;
  require(l >= 1 && l <= 16, "leaf_size must be in 1..16");
  return l;
}

int thg_bvh_build_mesh(thg_mesh mesh, const thg_build_opts *opts, void *stream, thg_bvh *out) {
  return guarded([&] {
    require(mesh && out, "mesh/out is NULL");
    Mesh *m = reinterpret_cast<Mesh *>(mesh);
    set_device(m->ctx);
    int leaf = leaf_from(opts);
    std::unique_ptr<Bvh> b(new Bvh());
    b->ctx = m->ctx;
    b->mesh = m;
    build_bvh(*b, m->verts.as<float>(), m->tris.as<int32_t>(), m->nt, false, leaf, as_stream(stream));
    THG_HIP(hipStreamSynchronize(as_stream(stream)));
    *out = reinterpret_cast<thg_bvh>(b.release());
  });
}

int thg_bvh_build_points(thg_ctx ctx, const float *xyz, int64_t n, int32_t mem, const thg_build_opts *opts,
                         void *stream, thg_bvh *out) {
  return guarded([&] {
    require(ctx && out, "ctx/out is NULL");
    require(n >= 0 && (n == 0 || xyz), "bad point array");
    require(n <= (int64_t)0x7ffffff0, "too many points for int32 indices");
    check_mem(mem);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    set_device(c);
    int leaf = leaf_from(opts);
    hipStream_t s = as_stream(stream);
    std::unique_ptr<Bvh> b(new Bvh());
    b->ctx = c;
    DevBuf staged;
    const float *dp = xyz;
    if (mem == THG_MEM_HOST && n > 0) {
      staged.alloc(n * 3 * sizeof(float));
      THG_HIP(hipMemcpyAsync(staged.p, xyz, n * 3 * sizeof(float), hipMemcpyHostToDevice, s));
      dp = staged.as<float>();
    }
    build_bvh(*b, dp, nullptr, n, true, leaf, s);
    THG_HIP(hipStreamSynchronize(s));
    *out = reinterpret_cast<thg_bvh>(b.release());
  });
}

int thg_bvh_build_points_f64(thg_ctx ctx, const double *xyz, int64_t n, int32_t mem, const thg_build_opts *opts,
                             void *stream, thg_bvh *out) {
  return guarded([&] {
    require(ctx && out, "ctx/out is NULL");
    require(n >= 0 && (n == 0 || xyz), "bad point array");
    require(n <= (int64_t)0x7ffffff0, "too many points for int32 indices");
    check_mem(mem);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    set_device(c);
    int leaf = leaf_from(opts);
    hipStream_t s = as_stream(stream);
    std::unique_ptr<Bvh> b(new Bvh());
    b->ctx = c;
    DevBuf staged;
    const double *dp = xyz;
    if (mem == THG_MEM_HOST && n > 0) {
      staged.alloc(n * 3 * sizeof(double));
      THG_HIP(hipMemcpyAsync(staged.p, xyz, n * 3 * sizeof(double), hipMemcpyHostToDevice, s));
      dp = staged.as<double>();
    }
    build_points64(*b, dp, n, leaf, s);
    *out = reinterpret_cast<thg_bvh>(b.release());
  });
}

int thg_bvh_destroy(thg_bvh bvh) {
  return guarded([&] { delete reinterpret_cast<Bvh *>(bvh); });
}

int thg_bvh_info(thg_bvh bvh, int64_t out[6]) {
  return guarded([&] {
    require(bvh && out, "bvh/out is NULL");
    Bvh *b = reinterpret_cast<Bvh *>(bvh);
    out[0] = b->num_nodes;
    out[1] = b->num_leaves;
    out[2] = b->max_depth;
    out[3] = b->n;
    out[4] = (int64_t)b->device_bytes();
    out[5] = b->leaf_size;
  });
}

int thg_bvh_bounds(thg_bvh bvh, float lo[3], float hi[3]) {
  return guarded([&] {
    require(bvh && lo && hi, "bvh/lo/hi is NULL");
    const Bvh *b = reinterpret_cast<const Bvh *>(bvh);
    lo[0] = b->lo.x, lo[1] = b->lo.y, lo[2] = b->lo.z;
    hi[0] = b->hi.x, hi[1] = b->hi.y, hi[2] = b->hi.z;
  });
}

__global__ void k_record_ids(int64_t n, const void *__restrict__ elems, int points, int32_t *__restrict__ ids) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  ids[i] = points ? __float_as_int(((const PointRec *)elems)[i].p.w) : trirec_id(((const TriRec *)elems)[i].pl);
}

int thg_bvh_export(thg_bvh bvh, void *nodes, int32_t *elem_ids) {
  return guarded([&] {
    require(bvh != nullptr, "bvh is NULL");
    Bvh *b = reinterpret_cast<Bvh *>(bvh);
    set_device(b->ctx);
    if (nodes && b->num_nodes > 0)
      THG_HIP(hipMemcpy(nodes, b->nodes.p, (size_t)b->num_nodes * sizeof(BNode), hipMemcpyDeviceToHost));
    if (elem_ids && b->n > 0) {
      DevBuf ids(b->n * 4);
      hipLaunchKernelGGL(k_record_ids, dim3(ceil_div(b->n, 256)), dim3(256), 0, 0, b->n, b->elems.p, b->points ? 1 : 0,
                         ids.as<int32_t>());
      launch_check("k_record_ids");
      THG_HIP(hipMemcpy(elem_ids, ids.p, b->n * 4, hipMemcpyDeviceToHost));
    }
  });
}

int thg_bvh_export_ray_tree(thg_bvh bvh, void *qnodes, int64_t *num_qnodes) {
  return guarded([&] {
    require(bvh != nullptr, "bvh is NULL");
    require(num_qnodes != nullptr, "num_qnodes is NULL");
    Bvh *b = reinterpret_cast<Bvh *>(bvh);
    set_device(b->ctx);
    *num_qnodes = b->qnodes.p ? b->num_qnodes : 0;
    if (qnodes && *num_qnodes > 0)
      THG_HIP(hipMemcpy(qnodes, b->qnodes.p, (size_t)*num_qnodes * sizeof(QNode4), hipMemcpyDeviceToHost));
  });
}

// Fill per-query "not found" outputs for an empty BVH.
__global__ void k_fill_missing(int64_t n, int k, int32_t *idx, float *a, float *b, float *cp3) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n * k) return;
  if (idx) idx[i] = -1;
  if (a) a[i] = -1.0f;
  if (b) b[i] = -1.0f;
  if (cp3) { cp3[3 * i] = 0; cp3[3 * i + 1] = 0; cp3[3 * i + 2] = 0; }
}

static void fill_missing(int64_t n, int k, int32_t *idx, float *a, float *b, float *cp3, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_missing, dim3(ceil_div(n * k, 256)), dim3(256), 0, s, n, k, idx, a, b, cp3);
  launch_check("k_fill_missing");
}

int thg_closest_point(thg_bvh bvh, const float *q, int64_t n, double dist_bound, int32_t *idx, float *cp, float *dist,
                      float *d2, int32_t mem, void *stream) {
  return guarded([&] {
    require(bvh != nullptr, "bvh is NULL");
    require(n >= 0, "negative query count");
    require(n == 0 || (q && idx), "q/idx is NULL");
    require(!std::isnan(dist_bound), "dist_bound is NaN");
    check_mem(mem);
    Bvh *b = reinterpret_cast<Bvh *>(bvh);
    set_device(b->ctx);
    hipStream_t s = as_stream(stream);
    Stage st(s, mem == THG_MEM_HOST);
    const float *dq = st.in(q, n * 3);
    int32_t *didx = st.out(idx, n);
    float *dcp = st.out(cp, n * 3), *ddist = st.out(dist, n), *dd2 = st.out(d2, n);
    if (b->n == 0) {
      fill_missing(n, 1, didx, ddist, dd2, dcp, s);
    } else {
      QueryScratch ws;
      launch_closest(*b, dq, n, dist_bound, 0, didx, dcp, ddist, dd2, ws, s);
    }
    st.finish();
  });
}

int thg_closest_point_f64(thg_bvh bvh, const double *q, int64_t n, double dist_bound, int32_t *idx, double *cp,
                          double *dist, double *d2, int32_t mem, void *stream) {
  return guarded([&] {
    require(bvh != nullptr, "bvh is NULL");
    require(n >= 0, "negative query count");
    require(n == 0 || (q && idx), "q/idx is NULL");
    require(!std::isnan(dist_bound), "dist_bound is NaN");
    check_mem(mem);
    Bvh *b = reinterpret_cast<Bvh *>(bvh);
    require(b->f64, "not a double-precision point BVH (build it with thg_bvh_build_points_f64)");
    set_device(b->ctx);
    hipStream_t s = as_stream(stream);
    Stage st(s, mem == THG_MEM_HOST);
    const double *dq = st.in(q, n * 3);
    int32_t *didx = st.out(idx, n);
    double *dcp = st.out(cp, n * 3), *ddist = st.out(dist, n), *dd2 = st.out(d2, n);
    QueryScratch ws;
    launch_closest64(*b, dq, n, dist_bound, didx, dcp, ddist, dd2, ws, s);
    st.finish();
  });
}

int thg_closest_point_noalloc(thg_bvh bvh, const float *q, int64_t n, double dist_bound, int32_t *idx, float *cp,
                              float *dist, float *d2, void *stream) {
  return guarded([&] {
    require(bvh != nullptr, "bvh is NULL");
    require(n >= 0, "negative query count");
    require(n <= THG_NOALLOC_MAX_QUERIES,
            "n = " + std::to_string(n) + " exceeds THG_NOALLOC_MAX_QUERIES (" +
                std::to_string((int64_t)THG_NOALLOC_MAX_QUERIES) + "); split the batch");
    require(n == 0 || (q && idx), "q/idx is NULL");
    require(!std::isnan(dist_bound), "dist_bound is NaN");
    static_assert(NOALLOC_WORDS >= 4 * (size_t)THG_NOALLOC_MAX_QUERIES + 256, "sort memory too small");
    Bvh *b = reinterpret_cast<Bvh *>(bvh);
    set_device(b->ctx);
    require(sort_scratch_words(THG_NOALLOC_MAX_QUERIES) <= NOALLOC_WORDS, "sort memory smaller than the sort needs");
    hipStream_t s = as_stream(stream);
    if (n == 0) return;
    if (b->n == 0) {
      fill_missing(n, 1, idx, dist, d2, cp, s);
      return;
    }
    QueryScratch ws;
    ws.ext = noalloc_ptr(b->ctx->device);
    ws.ext_cap = THG_NOALLOC_MAX_QUERIES;
    ws.timed = false;  // timing events are host objects created per launch; this path creates none
    launch_closest(*b, q, n, dist_bound, 0, idx, cp, dist, d2, ws, s);
  });
}

int thg_knn_wave_stats(thg_bvh bvh, const float *q, int64_t n, int32_t exclude_self, void *stream,
                       int64_t *per_wave) {
  return guarded([&] {
    require(bvh && per_wave, "bvh/per_wave is NULL");
    require(n >= 0 && (n == 0 || q), "bad query array");
    Bvh *b = reinterpret_cast<Bvh *>(bvh);
    set_device(b->ctx);
    QueryScratch ws;
    knn_wave_stats(*b, q, n, exclude_self, per_wave, ws, as_stream(stream));
  });
}

int thg_closest_stats(thg_bvh bvh, const float *q, int64_t n, void *stream, int64_t out[8]) {
  return guarded([&] {
    require(bvh && out, "bvh/out is NULL");
    require(n >= 0 && (n == 0 || q), "bad query array");
    Bvh *b = reinterpret_cast<Bvh *>(bvh);
    set_device(b->ctx);
    hipStream_t s = as_stream(stream);
    AsyncBuf st(8 * sizeof(unsigned long long), s);
    THG_HIP(hipMemsetAsync(st.p, 0, 8 * sizeof(unsigned long long), s));
    {
      QueryScratch ws;
      closest_wave_stats(*b, q, n, st.as<unsigned long long>(), ws, s);
    }
    unsigned long long h[8];
    THG_HIP(hipMemcpyAsync(h, st.p, sizeof(h), hipMemcpyDeviceToHost, s));
    THG_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < 8; ++i) out[i] = (int64_t)h[i];
  });
}

int thg_ray_stats(thg_bvh bvh, const float *origins, const float *dirs, int64_t n, void *stream, int32_t *idx,
                  int32_t *counts) {
  return guarded([&] {
    require(bvh != nullptr, "bvh is NULL");
    require(n >= 0 && (n == 0 || (origins && dirs && idx && counts)), "bad ray or output array");
    Bvh *b = reinterpret_cast<Bvh *>(bvh);
    require(!b->points, "ray queries need a triangle BVH");
    set_device(b->ctx);
    ray_stats(*b, origins, dirs, n, idx, counts, as_stream(stream));
  });
}

int thg_k_closest(thg_bvh bvh, const float *q, int64_t n, int32_t k, double dist_bound, int32_t exclude_self,
                  int32_t *idx, float *dist, float *d2, int32_t mem, void *stream) {
  return guarded([&] {
    require(bvh != nullptr, "bvh is NULL");
    require(n >= 0, "negative query count");
    require(k >= 1 && k <= 32, "k must be in 1..32");
    require(n == 0 || (q && idx), "q/idx is NULL");
    require(!std::isnan(dist_bound), "dist_bound is NaN");
    check_mem(mem);
    Bvh *b = reinterpret_cast<Bvh *>(bvh);
    require(!exclude_self || n <= b->n, "exclude_self needs the BVH's own points as queries");
    set_device(b->ctx);
    hipStream_t s = as_stream(stream);
    Stage st(s, mem == THG_MEM_HOST);
    const float *dq = st.in(q, n * 3);
    int32_t *didx = st.out(idx, n * k);
    float *ddist = st.out(dist, n * k), *dd2 = st.out(d2, n * k);
    if (b->n == 0) {
      fill_missing(n, k, didx, ddist, dd2, nullptr, s);
    } else {
      QueryScratch ws;
      launch_knn(*b, dq, n, k, dist_bound, exclude_self ? 1 : 0, didx, ddist, dd2, ws, s);
    }
    st.finish();
  });
}

int thg_ray_first_hit(thg_bvh bvh, const float *origins, const float *dirs, int64_t n, float max_time, int32_t *idx,
                      float *t, float *normal, int32_t mem, void *stream) {
  return guarded([&] {
    require(bvh != nullptr, "bvh is NULL");
    require(n >= 0, "negative ray count");
    require(n == 0 || (origins && dirs && idx), "origins/dirs/idx is NULL");
    require(!std::isnan(max_time), "max_time is NaN");
    check_mem(mem);
    Bvh *b = reinterpret_cast<Bvh *>(bvh);
    require(!b->points, "ray queries need a triangle BVH");
    set_device(b->ctx);
    hipStream_t s = as_stream(stream);
    Stage st(s, mem == THG_MEM_HOST);
    const float *dor = st.in(origins, n * 3), *ddir = st.in(dirs, n * 3);
    int32_t *didx = st.out(idx, n);
    float *dt = st.out(t, n), *dn = st.out(normal, n * 3);
    if (b->n == 0) {
      fill_missing(n, 1, didx, dt, nullptr, dn, s);
    } else {
      QueryScratch ws;
      launch_ray(*b, dor, ddir, n, max_time, false, didx, dt, dn, nullptr, ws, s);
    }
    st.finish();
  });
}

__global__ void k_fill_u8(int64_t n, uint8_t *p) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0;
}

int thg_ray_any_hit(thg_bvh bvh, const float *origins, const float *dirs, int64_t n, float max_time, uint8_t *hit,
                    int32_t mem, void *stream) {
  return guarded([&] {
    require(bvh != nullptr, "bvh is NULL");
    require(n >= 0, "negative ray count");
    require(n == 0 || (origins && dirs && hit), "origins/dirs/hit is NULL");
    require(!std::isnan(max_time), "max_time is NaN");
    check_mem(mem);
    Bvh *b = reinterpret_cast<Bvh *>(bvh);
    require(!b->points, "ray queries need a triangle BVH");
    set_device(b->ctx);
    hipStream_t s = as_stream(stream);
    Stage st(s, mem == THG_MEM_HOST);
    const float *dor = st.in(origins, n * 3), *ddir = st.in(dirs, n * 3);
    uint8_t *dh = st.out(hit, n);
    if (b->n == 0) {
      if (n > 0) {
        hipLaunchKernelGGL(k_fill_u8, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, dh);
        launch_check("k_fill_u8");
      }
    } else {
      QueryScratch ws;
      launch_ray(*b, dor, ddir, n, max_time, true, nullptr, nullptr, nullptr, dh, ws, s);
    }
    st.finish();
  });
}

int thg_sample_area(thg_mesh mesh, int64_t n, uint64_t seed, uint64_t stream_id, float *pos, float *normal,
                    int32_t *tri, int64_t *n_out, int32_t mem, void *stream) {
  return guarded([&] {
    require(mesh != nullptr, "mesh is NULL");
    require(n >= 0, "negative sample count");
    require(n == 0 || pos, "pos is NULL");
    check_mem(mem);
    Mesh *m = reinterpret_cast<Mesh *>(mesh);
    set_device(m->ctx);
    hipStream_t s = as_stream(stream);
    mesh_cdf(*m, s);
    int64_t got = (n > 0 && m->nt > 0 && m->total_area > 0) ? n : 0;
    Stage st(s, mem == THG_MEM_HOST);
    float *dp = st.out(pos, got * 3), *dn = st.out(normal, got * 3);
    int32_t *dt = st.out(tri, got);
    launch_sample(*m, got, seed, stream_id, dp, dn, dt, s);
    st.finish();
    if (n_out) *n_out = got;
  });
}

int thg_sample_area_approx(thg_mesh mesh, int64_t n, uint64_t seed, uint64_t stream_id, int64_t capacity, float *pos,
                           float *normal, int32_t *tri, int64_t *n_out, int32_t mem, void *stream) {
  return guarded([&] {
    require(mesh != nullptr && n_out != nullptr, "mesh/n_out is NULL");
    require(n >= 0, "negative sample count");
    require(pos || (!normal && !tri), "normal/tri need pos");
    check_mem(mem);
    Mesh *m = reinterpret_cast<Mesh *>(mesh);
    set_device(m->ctx);
    hipStream_t s = as_stream(stream);
    mesh_cdf(*m, s);
    *n_out = 0;
    const int64_t total = launch_sample_approx(*m, n, seed, stream_id, 0, nullptr, nullptr, nullptr, s);
    if (!pos || total == 0) {
      *n_out = total;
      return;
    }
    require(total <= capacity, "APPROXIMATE sampling produced " + std::to_string(total) +
                                   " samples, more than the output capacity " + std::to_string(capacity));
    Stage st(s, mem == THG_MEM_HOST);
    float *dp = st.out(pos, total * 3), *dn = st.out(normal, total * 3);
    int32_t *dt = st.out(tri, total);
    launch_sample_approx(*m, n, seed, stream_id, capacity, dp, dn, dt, s);
    st.finish();
    *n_out = total;
  });
}

}  // extern "C"

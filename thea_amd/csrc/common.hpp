// Shared host-side infrastructure of libthea_gpu: status handling, device buffers, handle types.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/thea_gpu.h"
#include "device_types.hpp"

namespace thg {

// A failed call throws Status; the C ABI layer catches it, stores the text as the thread's last
// error and returns the code. Thea's equivalent is Error (Error.hpp:27) thrown from BvhN::init.
struct Status {
  int code;
  std::string msg;
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw Status{code, msg}; }

inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    // Clear the sticky non-fatal error so later calls are not blamed for it.
    (void)hipGetLastError();
    fail(e == hipErrorOutOfMemory ? THG_E_OOM : THG_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  }
}

#define THG_HIP(call) ::thg::hip_check((call), #call)

// Integer tuning override from the environment (unset or empty: the default).
inline int env_int(const char *name, int dflt) {
  const char *e = std::getenv(name);
  return e && *e ? std::atoi(e) : dflt;
}

inline void require(bool ok, const std::string &msg) {
  if (!ok) fail(THG_E_INVALID, msg);
}

// Owning device allocation.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { reset(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { reset(); }
  void alloc(size_t n) {
    reset();
    if (n == 0) return;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      fail(THG_E_OOM, "hipMalloc(" + std::to_string(n) + " bytes) failed: " + hipGetErrorString(e));
    }
    bytes = n;
  }
  void ensure(size_t n) { if (n > bytes) alloc(n); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T> T *as() const { return static_cast<T *>(p); }
};

// Stream-ordered device allocation (hipMallocAsync): freed on its stream, so the memory is reused
// only after the work queued before the free has finished. No device-wide synchronisation.
struct AsyncBuf {
  void *p = nullptr;
  size_t bytes = 0;
  hipStream_t s = nullptr;
  AsyncBuf() = default;
  AsyncBuf(size_t n, hipStream_t st) { ensure(n, st); }
  AsyncBuf(const AsyncBuf &) = delete;
  AsyncBuf &operator=(const AsyncBuf &) = delete;
  ~AsyncBuf() { reset(); }
  void ensure(size_t n, hipStream_t st) {
    if (n <= bytes && st == s) return;
    reset();
    s = st;
    if (n == 0) return;
    hipError_t e = hipMallocAsync(&p, n, s);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      fail(THG_E_OOM, "hipMallocAsync(" + std::to_string(n) + " bytes) failed: " + hipGetErrorString(e));
    }
    bytes = n;
  }
  void reset() {
    if (p) (void)hipFreeAsync(p, s);
    p = nullptr;
    bytes = 0;
  }
  template <typename T> T *as() const { return static_cast<T *>(p); }
};

// Host-memory staging for one call: inputs copied to device buffers, outputs written to device
// buffers and copied back in finish().
struct Stage {
  hipStream_t s;
  bool host;
  std::vector<std::unique_ptr<AsyncBuf>> bufs;
  struct Out { void *host; void *dev; size_t bytes; };
  std::vector<Out> outs;
  Stage(hipStream_t st, bool h) : s(st), host(h) {}
  template <typename T> const T *in(const T *p, size_t count) {
    if (!host || !p || count == 0) return p;
    bufs.emplace_back(new AsyncBuf(count * sizeof(T), s));
    THG_HIP(hipMemcpyAsync(bufs.back()->p, p, count * sizeof(T), hipMemcpyHostToDevice, s));
    return bufs.back()->as<T>();
  }
  template <typename T> T *out(T *p, size_t count) {
    if (!host || !p || count == 0) return p;
    bufs.emplace_back(new AsyncBuf(count * sizeof(T), s));
    outs.push_back({p, bufs.back()->p, count * sizeof(T)});
    return bufs.back()->as<T>();
  }
  void finish() {
    if (!host) return;
    for (auto &o : outs) THG_HIP(hipMemcpyAsync(o.host, o.dev, o.bytes, hipMemcpyDeviceToHost, s));
    THG_HIP(hipStreamSynchronize(s));
  }
};

// Grow-only scratch owned by a context, one per (context, purpose). Calls that use the same
// context's scratch are serialised by `lock`.
struct Scratch {
  std::mutex lock;
  DevBuf a, b, c, d, e, f, g, h;
};

// Event pairs recorded around kernels while timing is enabled on a context.
struct TimedLaunch {
  std::string kernel;
  hipEvent_t start, stop;
};

struct Ctx {
  int device = 0;
  int num_cus = 256;
  Scratch scratch;       // host-pointer staging and per-call temporaries
  bool timing = false;
  std::mutex timing_lock;
  std::vector<TimedLaunch> timed;
};

// Records a start event on construction and a stop event in stop() when timing is enabled.
struct KernelTimer {
  Ctx *ctx;
  hipStream_t s;
  const char *name;
  hipEvent_t start = nullptr;
  KernelTimer(Ctx *c, hipStream_t st, const char *nm) : ctx(c), s(st), name(nm) {
    if (ctx && ctx->timing) {
      THG_HIP(hipEventCreate(&start));
      THG_HIP(hipEventRecord(start, s));
    }
  }
  void stop() {
    if (!start) return;
    hipEvent_t e;
    THG_HIP(hipEventCreate(&e));
    THG_HIP(hipEventRecord(e, s));
    std::lock_guard<std::mutex> g(ctx->timing_lock);
    ctx->timed.push_back({name, start, e});
    start = nullptr;
  }
};

struct Mesh {
  Ctx *ctx = nullptr;
  int64_t nv = 0, nt = 0;
  DevBuf verts;          // float[3 nv]
  DevBuf tris;           // int32[3 nt]
  DevBuf area;           // float[nt], Thea's per-triangle area (Triangle3.hpp:154)
  DevBuf cdf;            // double[nt], prefix sums of area in triangle order (lazy)
  bool cdf_ready = false;
  bool cdf_exact_scan = false;  // CDF built by the parallel exact scan (else sequentially)
  double total_area = 0;
  std::mutex cdf_lock;
};

struct Bvh {
  Ctx *ctx = nullptr;
  Mesh *mesh = nullptr;  // null for point BVHs
  bool points = false;
  int64_t n = 0;         // elements
  int64_t num_nodes = 0, num_leaves = 0;
  int32_t max_depth = 0;
  int32_t leaf_size = 4;
  int32_t wave_g = 1;    // lanes per query of the wave-cooperative kernels
  bool wave_pre = true;   // per-element box prefilter in the wave-cooperative kernels
  bool wave_cmp = true;   // compacted (lane, triangle) evaluation queue
  int32_t child_order = 0;  // child visiting order of the compacted kernel (0: majority of lanes)
  float scale = 0;       // max |coordinate| over the elements
  float3 lo{}, hi{};     // scene bounds (element boxes)
  DevBuf nodes;          // BNode[num_nodes]
  DevBuf elems;          // TriRec[n] or PointRec[n], leaf order
  DevBuf rays;           // RayTri[n] (mesh BVHs), leaf order
  DevBuf qnodes;         // QNode4[num_qnodes] (mesh BVHs): the 4-wide quantized tree for rays
  DevBuf lbounds;        // LeafBound[n] (mesh BVHs), at each leaf's first record position
  DevBuf screen;         // ScreenRec[n] (mesh BVHs), leaf order
  int64_t num_qnodes = 0;
  int32_t q_depth = 0;   // QNode4 levels
  bool f64 = false;      // point BVH built from double coordinates
  DevBuf elems64;        // PointRecD[n] (f64 only), leaf order
  size_t device_bytes() const {
    return nodes.bytes + elems.bytes + rays.bytes + qnodes.bytes + lbounds.bytes + screen.bytes + elems64.bytes;
  }
};

inline int float_to_ord_host(float f) {
  int i;
  std::memcpy(&i, &f, 4);
  return i >= 0 ? i : i ^ 0x7fffffff;
}

// Thread-local last error text for thg_last_error().
void set_last_error(const std::string &s);

// Device attribute helpers.
int device_cus(int device);

}  // namespace thg

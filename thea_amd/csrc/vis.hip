// External visibility: Thea's Visibility::compute (Algorithms/SurfaceFeatures/Local/Visibility.hpp:
// 99-125) for a batch of surface points against a mesh BVH.
//
// Per point: num_rays directions uniform on the sphere by Knop's method (Random.cpp:226-250, CACM
// Algorithm 381: x, y = 2 u - 1 until 1e-10 <= x^2 + y^2 <= 1, z = 2 m2 - 1, (x, y) *= 2 sqrt(1 -
// m2)), each ray from position + (0.001 scale) u, and the fraction of rays that hit nothing. Thea
// draws its uniforms from the time-seeded Random::common(); here ray r of point i draws them from
// Philox4x32-10 at counter (i * num_rays + r, VIS_STREAM, attempt) with the caller's seed as key,
// two (x, y) tries per counter, so a batch gives the same values on every run.
#include <cmath>

#include "common.hpp"
#include "device_types.hpp"
#include "kernels_common.hpp"
#include "philox.hpp"
#include "query.hpp"
#include "tri_math.hpp"

namespace thg {

constexpr uint32_t VIS_STREAM = 0x56495342u;  // "VISB"
constexpr int VIS_MAX_ATTEMPTS = 64;           // 128 tries; each is accepted with probability pi/4
constexpr int64_t VIS_CHUNK_RAYS = (int64_t)1 << 26;

// Ray g's direction (Knop's method over Philox uniforms). Exhausting the attempts (probability
// below 1e-85) gives +z.
__host__ __device__ inline float3 vis_direction(uint64_t g, uint64_t seed) {
  for (uint32_t a = 0; a < (uint32_t)VIS_MAX_ATTEMPTS; ++a) {
    uint32_t w[4];
    philox4x32_10((uint32_t)g, (uint32_t)(g >> 32), VIS_STREAM, a, (uint32_t)seed, (uint32_t)(seed >> 32), w);
    for (int h = 0; h < 2; ++h) {
      float x = 2 * u01(w[2 * h]) - 1;
      float y = 2 * u01(w[2 * h + 1]) - 1;
      const float m2 = x * x + y * y;
      if (m2 > 1 || (double)m2 < 1e-10) continue;
      const float z = 2 * m2 - 1;
      const float s = 2 * sqrtf(1 - m2);
      x *= s;
      y *= s;
      return f3(x, y, z);
    }
  }
  return f3(0.0f, 0.0f, 1.0f);
}

// One thread per ray of points [p0, p0 + np): origin position + (0.001 scale) u, direction u.
__global__ void __launch_bounds__(256) k_vis_rays(const float *__restrict__ pos, int64_t p0, int64_t np, int32_t R,
                                                  float scale, uint64_t seed, float *__restrict__ org,
                                                  float *__restrict__ dir) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= np * R) return;
  const int64_t i = p0 + g / R;
  const float3 u = vis_direction((uint64_t)(i * R + g % R), seed);
  const float f = 0.001f * scale;  // Visibility.hpp:115 (0.001f * scale * u)
  org[3 * g] = pos[3 * i] + f * u.x;
  org[3 * g + 1] = pos[3 * i + 1] + f * u.y;
  org[3 * g + 2] = pos[3 * i + 2] + f * u.z;
  dir[3 * g] = u.x;
  dir[3 * g + 1] = u.y;
  dir[3 * g + 2] = u.z;
}

__device__ __forceinline__ uint32_t vis_spread7(uint32_t v) {  // 7 bits -> every third bit
  uint32_t r = 0;
  for (int i = 0; i < 7; ++i) r |= ((v >> i) & 1u) << (3 * i);
  return r;
}

// Tracing order of the rays: key = direction octant (3 bits) above the 21-bit Morton code of the
// origin (7 bits per axis in the scene box), so a wave traces rays of one octant from nearby
// points rather than 64 random directions from one point (74.4 -> 67.6 ms per 100M C3 rays; an
// 8 x 8 octahedral direction grid over a 6-bit-per-axis origin code measured 70.9 ms).
__global__ void __launch_bounds__(256) k_vis_keys(const float *__restrict__ org, const float *__restrict__ dir,
                                                  int64_t m, float3 lo, float3 inv, uint32_t *__restrict__ keys,
                                                  uint32_t *__restrict__ vals) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= m) return;
  const float x = (org[3 * g] - lo.x) * inv.x, y = (org[3 * g + 1] - lo.y) * inv.y, z = (org[3 * g + 2] - lo.z) * inv.z;
  const uint32_t ix = (uint32_t)fminf(fmaxf(x * 128.0f, 0.0f), 127.0f);
  const uint32_t iy = (uint32_t)fminf(fmaxf(y * 128.0f, 0.0f), 127.0f);
  const uint32_t iz = (uint32_t)fminf(fmaxf(z * 128.0f, 0.0f), 127.0f);
  const uint32_t oct = (dir[3 * g] < 0 ? 4u : 0u) | (dir[3 * g + 1] < 0 ? 2u : 0u) | (dir[3 * g + 2] < 0 ? 1u : 0u);
  keys[g] = (oct << 21) | (vis_spread7(ix) << 2) | (vis_spread7(iy) << 1) | vis_spread7(iz);
  vals[g] = (uint32_t)g;
}

// One thread per point: the fraction of its rays that escaped.
__global__ void __launch_bounds__(256) k_vis_count(const uint8_t *__restrict__ hit, int64_t np, int32_t R,
                                                   double *__restrict__ vis) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= np) return;
  int64_t escaped = 0;
  for (int32_t r = 0; r < R; ++r) escaped += hit[i * R + r] ? 0 : 1;
  vis[i] = (double)escaped / R;
}

}  // namespace thg

using namespace thg;

template <typename F> static int vis_guarded(F f) {
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

extern "C" int thg_visibility(thg_bvh mesh_bvh, const float *positions, int64_t n, int32_t num_rays, uint64_t seed,
                              float scale, int32_t mem, void *stream, double *visibility) {
  return vis_guarded([&] {
    require(mesh_bvh != nullptr, "bvh is NULL");
    require(n >= 0 && (n == 0 || (positions && visibility)), "bad point or result array");
    require(mem == THG_MEM_DEVICE || mem == THG_MEM_HOST, "mem must be THG_MEM_DEVICE or THG_MEM_HOST");
    Bvh *b = reinterpret_cast<Bvh *>(mesh_bvh);
    require(!b->points, "visibility needs a mesh BVH");
    THG_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (n == 0) return;
    const int32_t R = num_rays > 0 ? num_rays : 100;  // Visibility.hpp:103-106
    if (!(scale > 0)) {  // Visibility.hpp:85: the diagonal of BvhN::getBounds (element box grown 1.05x)
      float e[3];
      const float lo[3] = {b->lo.x, b->lo.y, b->lo.z}, hi[3] = {b->hi.x, b->hi.y, b->hi.z};
      for (int a = 0; a < 3; ++a) {
        const float c = 0.5f * (hi[a] + lo[a]);
        e[a] = (1.05f * (hi[a] - c) + c) - (1.05f * (lo[a] - c) + c);
      }
      scale = sqrtf(e[0] * e[0] + (e[1] * e[1] + e[2] * e[2]));
    }
    Stage st(s, mem == THG_MEM_HOST);
    const float *d_pos = st.in(positions, (size_t)n * 3);
    double *d_vis = st.out(visibility, (size_t)n);
    const int64_t per = std::max<int64_t>(1, VIS_CHUNK_RAYS / R);
    const int64_t cap = std::min<int64_t>(n, per) * R;
    AsyncBuf org(cap * 12, s), dir(cap * 12, s), hit(cap, s), keys(cap * 4, s), vals(cap * 4, s), tk(cap * 4, s),
        tv(cap * 4, s), hist(radix_hist_words(cap) * 4, s);
    const float3 ext = f3(b->hi.x - b->lo.x, b->hi.y - b->lo.y, b->hi.z - b->lo.z);
    const float3 inv = f3(ext.x > 0 ? 1.0f / ext.x : 0.0f, ext.y > 0 ? 1.0f / ext.y : 0.0f, ext.z > 0 ? 1.0f / ext.z : 0.0f);
    QueryScratch ws;
    for (int64_t p0 = 0; p0 < n; p0 += per) {
      const int64_t np = std::min(per, n - p0), m = np * R;
      hipLaunchKernelGGL(k_vis_rays, dim3(ceil_div(m, 256)), dim3(256), 0, s, d_pos, p0, np, R, scale, seed,
                         org.as<float>(), dir.as<float>());
      launch_check("k_vis_rays");
      hipLaunchKernelGGL(k_vis_keys, dim3(ceil_div(m, 256)), dim3(256), 0, s, org.as<float>(), dir.as<float>(), m, b->lo,
                         inv, keys.as<uint32_t>(), vals.as<uint32_t>());
      launch_check("k_vis_keys");
      radix_sort_pairs(keys.as<uint32_t>(), vals.as<uint32_t>(), tk.as<uint32_t>(), tv.as<uint32_t>(),
                       hist.as<uint32_t>(), m, 0, 24, s);
      launch_ray(*b, org.as<float>(), dir.as<float>(), m, -1.0f, true, nullptr, nullptr, nullptr, hit.as<uint8_t>(), ws,
                 s, /*sort_rays=*/false, vals.as<uint32_t>());
      hipLaunchKernelGGL(k_vis_count, dim3(ceil_div(np, 256)), dim3(256), 0, s, hit.as<uint8_t>(), np, R,
                         d_vis + p0);
      launch_check("k_vis_count");
    }
    st.finish();
  });
}

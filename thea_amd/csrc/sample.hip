// Area-weighted surface sampling (Thea's MeshSampler::sampleEvenlyByArea, EXACT count mode,
// Algorithms/MeshSampler.hpp:100-204, with TriangleN::randomPoint, Triangle3.hpp:179-195).
//
//  k_areas       per-triangle float area with Thea's formula (Triangle3.hpp:154).
//  CDF           Thea sums areas sequentially in double (MeshSampler.hpp:147-149, :166-169). When
//                every partial sum is exactly representable (all areas are multiples of the smallest
//                area ulp 2^e and n * max area < 2^(e + 53)) any association gives the same bits,
//                so a parallel scan is used; otherwise a single lane sums in order.
//  k_sample      per sample: Philox4x32-10 words -> uniform01 (Random.hpp:126-131) -> lower_bound
//                over the CDF, clamped to the last triangle -> reflected barycentric point.
#include "kernels_common.hpp"
#include "philox.hpp"
#include "query.hpp"
#include "tri_math.hpp"

namespace thg {

constexpr int S_BLOCK = 256;

__device__ __forceinline__ float3 vtx(const float *__restrict__ v, int64_t i) {
  return f3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
}

__global__ void __launch_bounds__(S_BLOCK) k_areas(const float *__restrict__ verts, const int32_t *__restrict__ tris,
                                                   int64_t nt, float *__restrict__ area, int *__restrict__ stats) {
  // stats[0]: min ulp exponent over nonzero areas (as int), stats[1]: max area (ordered int bits).
  // Grid-stride loop, one pair of atomics per block.
  int ue = 1 << 20;
  float mx = 0;
  for (int64_t i = (int64_t)blockIdx.x * S_BLOCK + threadIdx.x; i < nt; i += (int64_t)gridDim.x * S_BLOCK) {
    float3 v0 = vtx(verts, tris[3 * i]), v1 = vtx(verts, tris[3 * i + 1]), v2 = vtx(verts, tris[3 * i + 2]);
    const float a = tri_area(v0, v1, v2);
    area[i] = a;
    int e = 1 << 20;
    if (a > 0) {
      if (!isfinite(a)) e = -(1 << 20);  // forces the sequential path
      else if (a < 1.17549435e-38f) e = -149;
      else e = ilogbf(a) - 23;
    } else if (!(a == 0)) {
      e = -(1 << 20);  // NaN
    }
    ue = min(ue, e);
    if (a == a) mx = fmaxf(mx, a);
  }
  __shared__ int s_ue[S_BLOCK / WAVE];
  __shared__ float s_mx[S_BLOCK / WAVE];
  const float wm = wave_max(mx);
  const int wu = (int)wave_min((float)ue);
  if (lane_id() == 0) {
    s_ue[threadIdx.x / WAVE] = wu;
    s_mx[threadIdx.x / WAVE] = wm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int bu = s_ue[0];
    float bm = s_mx[0];
    for (int w = 1; w < S_BLOCK / WAVE; ++w) {
      bu = min(bu, s_ue[w]);
      bm = fmaxf(bm, s_mx[w]);
    }
    atomicMin(&stats[0], bu);
    atomicMax(&stats[1], float_to_ord(bm));
  }
}

// Three-phase exact scan: block sums, scan of block sums, block rescan.
__global__ void __launch_bounds__(S_BLOCK) k_dsum_blocks(const float *__restrict__ a, int64_t n, double *__restrict__ bs) {
  __shared__ double part[S_BLOCK / WAVE];
  int64_t i = (int64_t)blockIdx.x * S_BLOCK + threadIdx.x;
  double v = i < n ? (double)a[i] : 0.0;
  double s = wave_sum(v);
  if (lane_id() == 0) part[threadIdx.x / WAVE] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < S_BLOCK / WAVE; ++w) t += part[w];
    bs[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(1024) k_dscan_excl(double *__restrict__ v, int64_t n) {
  __shared__ double part[1024 / WAVE];
  __shared__ double carry;
  int t = threadIdx.x;
  if (t == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < n; base += 1024) {
    int64_t i = base + t;
    double x = i < n ? v[i] : 0.0;
    double inc = wave_incl_scan(x);
    int w = t / WAVE, l = t % WAVE;
    if (l == WAVE - 1) part[w] = inc;
    __syncthreads();
    if (w == 0) {
      double pv = l < 1024 / WAVE ? part[l] : 0.0;
      double pinc = wave_incl_scan(pv);
      if (l < 1024 / WAVE) part[l] = pinc - pv;
    }
    __syncthreads();
    double excl = carry + part[w] + (inc - x);
    if (i < n) v[i] = excl;
    __syncthreads();
    if (t == 1023) carry = excl + x;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(S_BLOCK) k_dscan_blocks(const float *__restrict__ a, int64_t n,
                                                          const double *__restrict__ boff, double *__restrict__ cdf) {
  __shared__ double part[S_BLOCK / WAVE];
  int64_t i = (int64_t)blockIdx.x * S_BLOCK + threadIdx.x;
  double v = i < n ? (double)a[i] : 0.0;
  double inc = wave_incl_scan(v);
  int w = threadIdx.x / WAVE;
  if (lane_id() == WAVE - 1) part[w] = inc;
  __syncthreads();
  double off = boff[blockIdx.x];
  for (int k = 0; k < w; ++k) off += part[k];
  if (i < n) cdf[i] = off + inc;
}

// Sequential double prefix sum in Thea's order (MeshSampler.hpp:166-169), one lane.
__global__ void k_dscan_seq(const float *__restrict__ a, int64_t n, double *__restrict__ cdf) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double c = 0;
  int64_t i = 0;
  for (; i + 8 <= n; i += 8) {
    float x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = a[i + k];
#pragma unroll
    for (int k = 0; k < 8; ++k) { c += (double)x[k]; cdf[i + k] = c; }
  }
  for (; i < n; ++i) { c += (double)a[i]; cdf[i] = c; }
}

void mesh_cdf(Mesh &m, hipStream_t s) {
  std::lock_guard<std::mutex> g(m.cdf_lock);
  if (m.cdf_ready) return;
  if (m.nt <= 0) {
    m.total_area = 0;
    m.cdf_ready = true;
    return;
  }
  m.area.alloc(m.nt * sizeof(float));
  DevBuf stats(8);
  int init[2] = {1 << 20, float_to_ord_host(-INFINITY)};
  THG_HIP(hipMemcpyAsync(stats.p, init, 8, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_areas, dim3(std::min<uint32_t>(ceil_div(m.nt, S_BLOCK), 1024)), dim3(S_BLOCK), 0, s, m.verts.as<float>(),
                     m.tris.as<int32_t>(), m.nt, m.area.as<float>(), stats.as<int>());
  launch_check("k_areas");
  int st[2];
  THG_HIP(hipMemcpyAsync(st, stats.p, 8, hipMemcpyDeviceToHost, s));
  THG_HIP(hipStreamSynchronize(s));
  float max_area = ord_to_float(st[1]);
  m.cdf.alloc(m.nt * sizeof(double));
  bool exact = st[0] > -(1 << 19) && st[0] < (1 << 19) &&
               (double)m.nt * (double)max_area < std::ldexp(1.0, st[0] + 53);
  if (st[0] == (1 << 20)) exact = true;  // every area is zero
  if (exact) {
    uint32_t nb = ceil_div(m.nt, S_BLOCK);
    DevBuf bs(nb * sizeof(double));
    hipLaunchKernelGGL(k_dsum_blocks, dim3(nb), dim3(S_BLOCK), 0, s, m.area.as<float>(), m.nt, bs.as<double>());
    launch_check("k_dsum_blocks");
    hipLaunchKernelGGL(k_dscan_excl, dim3(1), dim3(1024), 0, s, bs.as<double>(), (int64_t)nb);
    launch_check("k_dscan_excl");
    hipLaunchKernelGGL(k_dscan_blocks, dim3(nb), dim3(S_BLOCK), 0, s, m.area.as<float>(), m.nt, bs.as<double>(),
                       m.cdf.as<double>());
    launch_check("k_dscan_blocks");
    THG_HIP(hipStreamSynchronize(s));
  } else {
    hipLaunchKernelGGL(k_dscan_seq, dim3(1), dim3(64), 0, s, m.area.as<float>(), m.nt, m.cdf.as<double>());
    launch_check("k_dscan_seq");
  }
  double total = 0;
  THG_HIP(hipMemcpyAsync(&total, m.cdf.as<double>() + (m.nt - 1), sizeof(double), hipMemcpyDeviceToHost, s));
  THG_HIP(hipStreamSynchronize(s));
  m.total_area = total;
  m.cdf_exact_scan = exact;
  m.cdf_ready = true;
}

__global__ void __launch_bounds__(S_BLOCK) k_sample(const float *__restrict__ verts, const int32_t *__restrict__ tris,
                                                    const double *__restrict__ cdf, int64_t nt, double total, int64_t n,
                                                    uint32_t k0, uint32_t k1, uint32_t s0, uint32_t s1,
                                                    float *__restrict__ pos, float *__restrict__ nrm,
                                                    int32_t *__restrict__ tri_out) {
  int64_t i = (int64_t)blockIdx.x * S_BLOCK + threadIdx.x;
  if (i >= n) return;
  uint32_t w[4];
  philox4x32_10((uint32_t)i, (uint32_t)((uint64_t)i >> 32), s0, s1, k0, k1, w);
  double t = (double)u01(w[0]) * total;
  // std::lower_bound: first c with !(cdf[c] < t)
  int64_t first = 0, count = nt;
  while (count > 0) {
    int64_t step = count >> 1, mid = first + step;
    if (cdf[mid] < t) { first = mid + 1; count -= step + 1; }
    else count = step;
  }
  int64_t ti = first < nt - 1 ? first : nt - 1;
  float3 v0 = vtx(verts, tris[3 * ti]), v1 = vtx(verts, tris[3 * ti + 1]), v2 = vtx(verts, tris[3 * ti + 2]);
  float3 e01 = sub3(v1, v0), e02 = sub3(v2, v0);
  float a = u01(w[1]), b = u01(w[2]);
  if (a + b > 1) { b = 1 - b; a = 1 - a; }
  float3 p = add3(add3(scl3(a, e01), scl3(b, e02)), v0);
  pos[3 * i] = p.x;
  pos[3 * i + 1] = p.y;
  pos[3 * i + 2] = p.z;
  if (nrm) {
    float3 nn = plane_normal(v0, v1, v2);
    nrm[3 * i] = nn.x;
    nrm[3 * i + 1] = nn.y;
    nrm[3 * i + 2] = nn.z;
  }
  if (tri_out) tri_out[i] = (int32_t)ti;
}

// APPROXIMATE count mode (MeshSampler.hpp:206-240): triangle i gets frac = (n * area) / total
// samples in expectation: while frac > 0, a last sample (frac < 1) is kept only if a uniform draw is
// <= frac, and frac drops by 1 per sample. Subtracting 1 from a double below 2^52 is exact, so that
// is floor(frac) samples plus one more when r = frac - floor(frac) > 0 and the draw is <= r. The
// draw for triangle i is word 0 of Philox4x32-10 at counter (i, stream) with the seed's low key word
// xor-ed with APPROX_KEY_TAG; sample j (in triangle order) takes its barycentric uniforms from words
// 1 and 2 at counter (j, stream), as in EXACT mode.
constexpr uint32_t APPROX_KEY_TAG = 0x41505052u;

__global__ void __launch_bounds__(S_BLOCK) k_approx_count(const float *__restrict__ area, int64_t nt, double desired,
                                                          double total, uint32_t k0, uint32_t k1, uint32_t s0,
                                                          uint32_t s1, uint32_t *__restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * S_BLOCK + threadIdx.x;
  if (i >= nt) return;
  const double frac = (desired * (double)area[i]) / total;
  uint32_t c = 0;
  if (frac > 0) {
    const double whole = floor(frac);
    c = (uint32_t)whole;
    const double r = frac - whole;
    if (r > 0) {
      uint32_t w[4];
      philox4x32_10((uint32_t)i, (uint32_t)((uint64_t)i >> 32), s0, s1, k0 ^ APPROX_KEY_TAG, k1, w);
      if (!((double)u01(w[0]) > r)) ++c;
    }
  }
  count[i] = c;
}

__global__ void __launch_bounds__(S_BLOCK) k_approx_sample(const float *__restrict__ verts,
                                                           const int32_t *__restrict__ tris, int64_t nt,
                                                           const uint32_t *__restrict__ count,
                                                           const uint32_t *__restrict__ offset, uint32_t k0,
                                                           uint32_t k1, uint32_t s0, uint32_t s1,
                                                           float *__restrict__ pos, float *__restrict__ nrm,
                                                           int32_t *__restrict__ tri_out) {
  const int64_t i = (int64_t)blockIdx.x * S_BLOCK + threadIdx.x;
  if (i >= nt) return;
  const uint32_t c = count[i];
  if (c == 0) return;
  const float3 v0 = vtx(verts, tris[3 * i]), v1 = vtx(verts, tris[3 * i + 1]), v2 = vtx(verts, tris[3 * i + 2]);
  const float3 e01 = sub3(v1, v0), e02 = sub3(v2, v0);
  const float3 nn = plane_normal(v0, v1, v2);
  for (uint32_t q = 0; q < c; ++q) {
    const int64_t j = (int64_t)offset[i] + q;
    uint32_t w[4];
    philox4x32_10((uint32_t)j, (uint32_t)((uint64_t)j >> 32), s0, s1, k0, k1, w);
    float a = u01(w[1]), b = u01(w[2]);
    if (a + b > 1) { b = 1 - b; a = 1 - a; }
    const float3 p = add3(add3(scl3(a, e01), scl3(b, e02)), v0);
    pos[3 * j] = p.x;
    pos[3 * j + 1] = p.y;
    pos[3 * j + 2] = p.z;
    if (nrm) {
      nrm[3 * j] = nn.x;
      nrm[3 * j + 1] = nn.y;
      nrm[3 * j + 2] = nn.z;
    }
    if (tri_out) tri_out[j] = (int32_t)i;
  }
}

// Counts (and, with pos set, draws) the APPROXIMATE samples; returns their number.
int64_t launch_sample_approx(Mesh &m, int64_t n, uint64_t seed, uint64_t stream_id, int64_t capacity, float *pos,
                             float *nrm, int32_t *tri, hipStream_t s) {
  if (n <= 0 || m.nt <= 0 || !(m.total_area > 0)) return 0;
  require(n < ((int64_t)1 << 31), "APPROXIMATE sampling supports fewer than 2^31 samples");
  // the scan writes the total after the last offset: off holds nt + 1 words
  AsyncBuf cnt(m.nt * 4, s), off((m.nt + 1) * 4, s), tmp(scan_tmp_words(m.nt) * 4, s);
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32), s0 = (uint32_t)stream_id,
                 s1 = (uint32_t)(stream_id >> 32);
  hipLaunchKernelGGL(k_approx_count, dim3(ceil_div(m.nt, S_BLOCK)), dim3(S_BLOCK), 0, s, m.area.as<float>(), m.nt,
                     (double)n, m.total_area, k0, k1, s0, s1, cnt.as<uint32_t>());
  launch_check("k_approx_count");
  scan_excl_u32(cnt.as<uint32_t>(), off.as<uint32_t>(), m.nt, tmp.as<uint32_t>(), s);
  uint32_t tot = 0;
  THG_HIP(hipMemcpyAsync(&tot, off.as<uint32_t>() + m.nt, 4, hipMemcpyDeviceToHost, s));
  THG_HIP(hipStreamSynchronize(s));
  const int64_t total = (int64_t)tot;
  if (!pos) return total;
  require(total <= capacity, "APPROXIMATE sampling produced " + std::to_string(total) +
                                 " samples, more than the output capacity " + std::to_string(capacity));
  KernelTimer kt(m.ctx, s, "sample");
  hipLaunchKernelGGL(k_approx_sample, dim3(ceil_div(m.nt, S_BLOCK)), dim3(S_BLOCK), 0, s, m.verts.as<float>(),
                     m.tris.as<int32_t>(), m.nt, cnt.as<uint32_t>(), off.as<uint32_t>(), k0, k1, s0, s1, pos, nrm, tri);
  launch_check("k_approx_sample");
  kt.stop();
  return total;
}

void launch_sample(Mesh &m, int64_t n, uint64_t seed, uint64_t stream_id, float *pos, float *nrm, int32_t *tri,
                   hipStream_t s) {
  if (n <= 0 || m.nt <= 0 || !(m.total_area > 0)) return;
  KernelTimer kt(m.ctx, s, "sample");
  hipLaunchKernelGGL(k_sample, dim3(ceil_div(n, S_BLOCK)), dim3(S_BLOCK), 0, s, m.verts.as<float>(), m.tris.as<int32_t>(),
                     m.cdf.as<double>(), m.nt, m.total_area, n, (uint32_t)seed, (uint32_t)(seed >> 32),
                     (uint32_t)stream_id, (uint32_t)(stream_id >> 32), pos, nrm, tri);
  launch_check("k_sample");
  kt.stop();
}

}  // namespace thg

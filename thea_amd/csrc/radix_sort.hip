// Stable LSD radix sort of 32-bit key/value pairs, 8-bit digits, LDS-staged.
//
// Pass structure (per 8-bit digit):
//   1. k_radix_hist: each 4096-key tile counts its digits in LDS and writes a 256-entry histogram
//      in digit-major order (hist[d * tiles + tile]).
//   2. k_radix_rowscan: one workgroup per digit turns its row of tile counts into exclusive
//      offsets and a row total; the scatter kernel scans the 256 totals itself.
//   3. k_radix_scatter: each tile ranks its keys stably (wave ballots give the rank among equal
//      digits inside a wave, an LDS table the offsets across waves and rounds) and writes every key
//      to its final position.
// Used for the 30-bit Morton codes of BVH elements and for spatially ordering query batches.
#include "kernels_common.hpp"

namespace thg {

constexpr int RS_BLOCK = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;
constexpr int RS_WAVES = RS_BLOCK / WAVE;

size_t radix_hist_words(int64_t n) { return (size_t)256 * ceil_div((uint64_t)std::max<int64_t>(n, 1), RS_TILE) + 256; }

__global__ void __launch_bounds__(RS_BLOCK) k_radix_hist(const uint32_t *__restrict__ keys, int64_t n, int shift,
                                                         uint32_t tiles, uint32_t *__restrict__ hist) {
  __shared__ uint32_t h[256];
  int t = threadIdx.x;
  h[t] = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
  for (int r = 0; r < RS_ITEMS; ++r) {
    int64_t i = base + (int64_t)r * RS_BLOCK + t;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFF], 1u);
  }
  __syncthreads();
  hist[(size_t)t * tiles + blockIdx.x] = h[t];
}

// Row scan: block d turns row d of the digit-major histogram (one count per tile) into exclusive
// offsets within the row and writes the row total to totals[d].
__global__ void __launch_bounds__(RS_BLOCK) k_radix_rowscan(uint32_t *__restrict__ hist, uint32_t tiles,
                                                            uint32_t *__restrict__ totals) {
  __shared__ uint32_t part[RS_WAVES];
  __shared__ uint32_t carry;
  int t = threadIdx.x;
  uint32_t *row = hist + (size_t)blockIdx.x * tiles;
  if (t == 0) carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < tiles; base += RS_BLOCK) {
    uint32_t i = base + t;
    uint32_t v = i < tiles ? row[i] : 0;
    uint32_t inc = wave_incl_scan(v);
    int w = t / WAVE, l = t % WAVE;
    if (l == WAVE - 1) part[w] = inc;
    __syncthreads();
    uint32_t off = carry;
    for (int k = 0; k < w; ++k) off += part[k];
    uint32_t excl = off + inc - v;
    if (i < tiles) row[i] = excl;
    __syncthreads();
    if (t == RS_BLOCK - 1) carry = excl + v;
    __syncthreads();
  }
  if (t == 0) totals[blockIdx.x] = carry;
}

__global__ void __launch_bounds__(RS_BLOCK) k_radix_scatter(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                            uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                            int64_t n, int shift, uint32_t tiles,
                                                            const uint32_t *__restrict__ offs,
                                                            const uint32_t *__restrict__ totals) {
  __shared__ uint32_t base[256];
  __shared__ uint32_t cnt[RS_WAVES][256];
  __shared__ uint32_t part[RS_WAVES];
  int t = threadIdx.x;
  int w = t / WAVE;
  {
    // digit base = exclusive scan of the 256 row totals (recomputed per tile; 256 values)
    uint32_t v = totals[t];
    uint32_t inc = wave_incl_scan(v);
    if (lane_id() == WAVE - 1) part[w] = inc;
    __syncthreads();
    uint32_t off = 0;
    for (int k = 0; k < w; ++k) off += part[k];
    base[t] = off + inc - v + offs[(size_t)t * tiles + blockIdx.x];
  }
  int64_t tile0 = (int64_t)blockIdx.x * RS_TILE;
  uint64_t lt = lanemask_lt();
  for (int r = 0; r < RS_ITEMS; ++r) {
#pragma unroll
    for (int k = 0; k < RS_WAVES; ++k) cnt[k][t] = 0;
    __syncthreads();
    int64_t i = tile0 + (int64_t)r * RS_BLOCK + t;
    bool valid = i < n;
    uint32_t key = valid ? kin[i] : 0u;
    uint32_t val = valid ? vin[i] : 0u;
    uint32_t d = (key >> shift) & 0xFF;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      uint64_t bb = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bb : ~bb;
    }
    uint32_t rank = (uint32_t)__popcll(peers & lt);
    if (valid && rank == 0) cnt[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    {
      uint32_t run = base[t];
#pragma unroll
      for (int k = 0; k < RS_WAVES; ++k) {
        uint32_t c = cnt[k][t];
        cnt[k][t] = run;
        run += c;
      }
      base[t] = run;
    }
    __syncthreads();
    if (valid) {
      uint32_t pos = cnt[w][d] + rank;
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
  }
}

void radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *tmp_keys, uint32_t *tmp_vals, uint32_t *hist,
                      int64_t n, int begin_bit, int end_bit, hipStream_t s) {
  if (n <= 1) return;
  uint32_t tiles = ceil_div((uint64_t)n, RS_TILE);
  uint32_t *ka = keys, *va = vals, *kb = tmp_keys, *vb = tmp_vals;
  int passes = 0;
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(RS_BLOCK), 0, s, ka, n, shift, tiles, hist);
    launch_check("k_radix_hist");
    uint32_t *totals = hist + (size_t)256 * tiles;
    hipLaunchKernelGGL(k_radix_rowscan, dim3(256), dim3(RS_BLOCK), 0, s, hist, tiles, totals);
    launch_check("k_radix_rowscan");
    hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(RS_BLOCK), 0, s, ka, va, kb, vb, n, shift, tiles, hist,
                       totals);
    launch_check("k_radix_scatter");
    std::swap(ka, kb);
    std::swap(va, vb);
    ++passes;
  }
  if (passes & 1) {
    THG_HIP(hipMemcpyAsync(keys, ka, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    THG_HIP(hipMemcpyAsync(vals, va, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  }
}

}  // namespace thg

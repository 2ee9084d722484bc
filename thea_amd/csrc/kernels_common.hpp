// Device helpers shared by the kernel files: wave-level primitives, ordered-float atomics,
// XCD-aware block remapping, launch checking.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace thg {

constexpr int WAVE = 64;  // CDNA wavefront width

__device__ __forceinline__ int lane_id() { return __lane_id(); }
// Wave-wide ballot of a lane predicate straight from the compare mask (no 0/1 round trip).
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ __forceinline__ uint64_t lanemask_lt() {
  int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Order-preserving float <-> int maps for integer atomicMin/atomicMax on floats.
__device__ __forceinline__ int float_to_ord(float f) {
  int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7fffffff;
}
__host__ __device__ __forceinline__ float ord_to_float(int i) {
  int j = i >= 0 ? i : i ^ 0x7fffffff;
  union { int i; float f; } u;
  u.i = j;
  return u.f;
}

template <typename T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, WAVE);
  return v;
}
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fminf(v, __shfl_xor(v, off, WAVE));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, WAVE));
  return v;
}

// A unique slot of `counter` for each calling lane (callable in divergent code): one atomic per
// wave, by the lowest active lane, instead of one per lane on the same address.
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter) {
  const uint64_t act = __ballot(1);
  const int leader = __ffsll((unsigned long long)act) - 1;
  uint32_t base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(act));
  base = __shfl(base, leader, WAVE);
  return base + (uint32_t)__popcll(act & lanemask_lt());
}

// Inclusive scan across the wave.
template <typename T> __device__ __forceinline__ T wave_incl_scan(T v) {
  int l = lane_id();
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    T o = __shfl_up(v, off, WAVE);
    if (l >= off) v += o;
  }
  return v;
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD b % 8). Remap so the
// blocks an XCD runs cover one contiguous range of logical blocks: neighbouring (spatially sorted)
// work then shares that XCD's L2. `nb` must be a multiple of 8; logical ids >= the real block count
// are skipped by the caller.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb) {
  const uint32_t X = 8;
  uint32_t per = nb / X;
  return (b % X) * per + b / X;
}

inline uint32_t round_up8(uint64_t nb) { return (uint32_t)((nb + 7) / 8 * 8); }

// A pointer read from a descriptor in device memory reaches the compiler as a generic (flat)
// pointer, and loads through it go to the vector memory path even when every lane reads the same
// address. The trees, records and queries a traversal reads are device memory that no wave writes
// while the kernel runs: reading the descriptor's field as a pointer to the constant address space
// (device memory that is read-only for the kernel's duration) lets the compiler carry that into the
// inlined traversal, so its wave-uniform node and record loads become scalar loads, as they are when
// the same pointers arrive as const __restrict__ kernel arguments. Loads that differ across lanes
// stay vector loads.
template <typename T, typename F> __device__ __forceinline__ const T *readonly_ptr(F *const &field) {
  typedef const __attribute__((address_space(4))) T *cptr;
  return (const T *)*reinterpret_cast<const cptr *>(reinterpret_cast<const void *>(&field));
}

inline void launch_check(const char *name) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) fail(THG_E_DEVICE, std::string("kernel launch ") + name + ": " + hipGetErrorString(e));
}

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Radix sort of (key, value) pairs over key bits [begin_bit, end_bit), stable, LSD with 8-bit
// digits. keys/vals are sorted in place (using tmp_keys/tmp_vals as the ping-pong partner);
// `hist` must hold radix_hist_words(n) uint32.
void radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *tmp_keys, uint32_t *tmp_vals, uint32_t *hist,
                      int64_t n, int begin_bit, int end_bit, hipStream_t s);
size_t radix_hist_words(int64_t n);

}  // namespace thg

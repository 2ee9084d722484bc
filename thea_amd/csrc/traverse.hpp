// Device helpers shared by the traversal kernels: the conservative box lower bound, the canonical
// (value, index) order, and the sorted K-list insertion.
#pragma once

#include "kernels_common.hpp"
#include "tri_math.hpp"

namespace thg {

// Relative half-width of the box a query is widened to (2^-16 of max(|q|, scene scale)), so that box
// lower bounds stay below the float distances of the elements inside (Thea pads nodes by x1.05 for
// the same purpose, BvhN.hpp:2458).
constexpr float SLACK_REL = 1.0f / 65536.0f;

// Squared distance between the query box [qlo, qhi] and a node box: a lower bound of the squared
// distance from the query to anything in the node. Only pruning reads it, so it may fuse its
// multiply-adds: the query box's slack (2^-16 of the scale per axis) is far wider than their rounding.
__device__ __forceinline__ float box_lb(float3 qlo, float3 qhi, float lx, float ly, float lz, float hx, float hy,
                                        float hz) {
  float dx = fmaxf(fmaxf(lx - qhi.x, qlo.x - hx), 0.0f);
  float dy = fmaxf(fmaxf(ly - qhi.y, qlo.y - hy), 0.0f);
  float dz = fmaxf(fmaxf(lz - qhi.z, qlo.z - hz), 0.0f);
  return fmaf(dx, dx, fmaf(dy, dy, dz * dz));
}

// Lower bound on the squared distance from a query p to a triangle, from the triangle's plane and a
// disc holding it: with c the disc's centre (a point of the plane), sd = n.(p - c) (unit normal n)
// is p's distance across the plane, and p's foot on the plane lies sqrt(|p - c|^2 - sd^2) from c,
// so at least that less the radius r from every point of the triangle; the two distances are
// orthogonal and their squares add. Each is lowered by the slack sl, which covers the rounding of
// the stored plane and centre (c is within 2^-23 |c| of the plane, the triangle within 2^-20 of its
// largest coordinate) and of this arithmetic: dc2 = |p - c|^2 may round up by a relative 4 2^-24, which the
// relative 2^-20 added to |sd| (in the term subtracted from it) more than covers across the plane
// and which moves the in-plane distance by far less than sl, as does the hardware square root
// (1 ulp). A negative r marks a triangle whose plane is not tight (degenerate or near-collinear; c
// is then its box centre): the bound is the bounding sphere's, |p - c| - |r|. A NaN sd gives 0,
// which prunes nothing.
__device__ __forceinline__ float tri_lb2(float dc2, float sd, float r, float sl) {
  const bool tight = __float_as_int(r) >= 0;
  const float asd = fabsf(sd);
  const float a = tight ? fmaxf(asd - sl, 0.0f) : 0.0f;
  const float u = tight ? fmaf(asd, 1.0f + 1.0f / 1048576.0f, sl) : 0.0f;
  const float ip2 = fmaf(-u, u, dc2);
  const float b = fmaxf(__builtin_amdgcn_sqrtf(fmaxf(ip2, 0.0f)) - fabsf(r) - sl, 0.0f);
  return fmaf(a, a, b * b);
}

// Lower bound on the squared distance from a query to a leaf's triangles from its LeafBound: with
// d = p - c, dc2 = |d|^2 and sd = n.d, the leaf lies within t of the plane and within r of the
// axis, so the distance across is at least |sd| - t and the distance along the plane at least
// sqrt(dc2 - sd^2) - r; the same slack and margins as tri_lb2 cover the rounding (the stored normal
// is a unit vector to float precision).
__device__ __forceinline__ float leaf_lb2(float dc2, float sd, float t, float r, float sl) {
  const float asd = fabsf(sd);
  const float a = fmaxf(asd - t - sl, 0.0f);
  const float u = fmaf(asd, 1.0f + 1.0f / 1048576.0f, sl);
  const float ip2 = fmaf(-u, u, dc2);
  const float b = fmaxf(__builtin_amdgcn_sqrtf(fmaxf(ip2, 0.0f)) - r - sl, 0.0f);
  return fmaf(a, a, b * b);
}

// Per-lane seed of the wave traversals: descend from `root` to the leaf whose box is nearest the
// query (no backtracking) and return it as [first, first + count). The elements found there give
// each lane a real upper bound before the shared walk starts, so a wave whose queries are far apart
// does not admit every node some other lane has not bounded yet. Depth is below 64 for the wave
// kernels (use_wave_traversal), which bounds the loop.
__device__ __forceinline__ void seed_leaf(const BNode *__restrict__ nodes, int32_t root, float3 qlo, float3 qhi,
                                          int32_t &first, int32_t &count) {
  int32_t nd_i = root;
  first = 0;
  count = 0;
  for (int it = 0; it < 64; ++it) {
    const BNode nd = nodes[nd_i];
    const float l0 = box_lb(qlo, qhi, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y);
    const float l1 = box_lb(qlo, qhi, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w);
    const bool one = l1 < l0;
    const int32_t r = one ? nd.d.y : nd.d.x;
    if (r >= 0) {
      nd_i = r;
      continue;
    }
    first = ~r;
    count = one ? nd.d.w : nd.d.z;
    return;
  }
}

// Lexicographic (value, index) order: exact ties go to the lower index.
__device__ __forceinline__ bool better(float d, int32_t i, float bd, int32_t bi) {
  return d < bd || (d == bd && i < bi);
}

// Insert (d, id) into the ascending list kd/ki of K pairs (the last pair drops out).
template <int K>
__device__ __forceinline__ void knn_insert(float (&kd)[K], int32_t (&ki)[K], float d, int32_t id) {
#pragma unroll
  for (int j = K - 1; j >= 0; --j) {
    bool prev = j > 0 ? better(d, id, kd[j > 0 ? j - 1 : 0], ki[j > 0 ? j - 1 : 0]) : false;
    if (better(d, id, kd[j], ki[j])) {
      kd[j] = prev ? kd[j > 0 ? j - 1 : 0] : d;
      ki[j] = prev ? ki[j > 0 ? j - 1 : 0] : id;
    }
  }
}

}  // namespace thg

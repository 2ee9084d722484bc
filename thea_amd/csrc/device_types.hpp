// Device data layout of libthea_gpu (shared by all kernels).
//
// HBM layout per BVH:
//   BNode[num_nodes]  64 B each: both children's boxes + refs, one 64-B segment per visit.
//   TriRec[n]        128 B each (one L2 line): a triangle with every per-triangle constant the
//                     closest-point and ray tests need, in leaf order, so a leaf is a contiguous run.
//   ScreenRec[n]      32 B each (mesh BVHs): bounding disc and normal, leaf order, for the
//                     closest-point walk's per-triangle reject.
//   LeafBound[n]      32 B per record position (mesh BVHs): the leaf-level reject, at each leaf's
//                     first record.
//   RayTri[n]         48 B each (mesh BVHs): v0 and the edges of each triangle, leaf order, for rays.
//   QNode4[nq]        64 B each (mesh BVHs): the same tree with every other level folded in and the
//                     child boxes quantized to bytes, for rays.
//   PointRec[n]       16 B each: point + original index, in leaf order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace thg {

// Binary BVH node. Child k's box is [lo_k, hi_k]; ref_k >= 0 is an internal node index, ref_k < 0
// is a leaf whose elements are [~ref_k, ~ref_k + cnt_k) in leaf order (cnt_k may be 0).
struct __align__(64) BNode {
  float4 a;  // lo0.x lo0.y lo0.z hi0.x
  float4 b;  // hi0.y hi0.z lo1.x lo1.y
  float4 c;  // lo1.z hi1.x hi1.y hi1.z
  int4 d;    // ref0, ref1, cnt0, cnt1
};
static_assert(sizeof(BNode) == 64, "BNode must be 64 bytes");

// Triangle record. v0/v1/v2 are the input vertices; n, pd are Thea's plane (Hyperplane3.hpp:34-41);
// `inv` is the reciprocal of the in-plane 2-D area of isPointInsideTriangle (Triangle3.hpp:1436-1447);
// dir_k/len_k are the perimeter edge directions and lengths of closestPointOnTrianglePerimeter
// (Triangle3.hpp:1401-1409). `meta` packs the primary axis (bits 0-1) and a flag for a zero 2-D area
// (bit 2); `id` is the caller's triangle index. (cx, cy, cz, r) is a bounding sphere of the
// triangle (radius rounded up), used only to skip elements that cannot win.
struct __align__(128) TriRec {
  float4 p0;   // v0.xyz, n.x
  float4 p1;   // v1.xyz, n.y
  float4 p2;   // v2.xyz, n.z
  float4 pl;   // pd, inv, meta (int bits), id (int bits)
  float4 e0;   // dir0.xyz, len0   (edge v0 -> v1)
  float4 e1;   // dir1.xyz, len1   (edge v1 -> v2)
  float4 e2;   // dir2.xyz, len2   (edge v2 -> v0)
  float4 sb;   // bounding sphere: centre xyz (in the triangle's plane), radius
};
static_assert(sizeof(TriRec) == 128, "TriRec must be 128 bytes");

__host__ __device__ __forceinline__ int trirec_axis(const float4 &pl) {
  union { float f; int i; } m;
  m.f = pl.z;
  return m.i & 3;
}
__host__ __device__ __forceinline__ bool trirec_flat(const float4 &pl) {
  union { float f; int i; } m;
  m.f = pl.z;
  return (m.i & 4) != 0;
}
__host__ __device__ __forceinline__ int trirec_id(const float4 &pl) {
  union { float f; int i; } m;
  m.f = pl.w;
  return m.i;
}

// Screening record of a triangle, in leaf order beside its TriRec: the bounding disc (TriRec.sb) and
// the unit normal, the only fields the closest-point walk's per-triangle reject reads. 32 bytes, so
// one 128-byte cache line holds four consecutive triangles of a leaf instead of one.
struct __align__(32) ScreenRec {
  float4 c;  // disc centre xyz, radius (negated when the plane is not tight)
  float4 n;  // unit normal xyz, 0
};

// Ray record of a triangle, in leaf order beside its TriRec: v0 and the two edges Eberly's test
// starts from (e01 = v1 - v0, e02 = v2 - v0, the same float subtractions as Triangle3.hpp:151-152),
// and the caller's index. 48 bytes, so a ray's triangle tests read 48 of every 128 bytes a TriRec
// would take and neighbouring leaves share cache lines. b.w flags the last record of a leaf, so a
// traversal holding only a leaf's first index knows where it ends.
struct __align__(16) RayTri {
  float4 a;  // v0.xyz, id (int bits)
  float4 b;  // e01.xyz, last-of-leaf flag (int bits: 1 or 0)
  float4 c;  // e02.xyz, 0
};
static_assert(sizeof(RayTri) == 48, "RayTri must be 48 bytes");

// Quantized 4-wide node of the ray BVH (64 bytes, one per visit as for BNode, but four child boxes
// instead of two). A child box is stored as 8-bit offsets on a per-node, per-axis power-of-two grid:
// lo = fmaf(qlo, 2^e, origin), hi = fmaf(qhi, 2^e, origin), rounded outward at build so the decoded
// box (with exactly that fmaf) contains the child's exact box. Children 0..nchild-1 are valid;
// refs follow BNode (>= 0 inner QNode4, < 0 leaf ~first in RayTri order, ended by its last flag).
// The leaf counts sit in the last 8 bytes, which ray traversal does not load (56 of 64 bytes).
struct __align__(64) QNode4 {
  float4 a;  // origin xyz, meta (int bits): biased exponents ex | ey << 8 | ez << 16 | nchild << 24
  uint4 b;   // qlo_x, qlo_y, qlo_z, qhi_x (byte k: child k)
  int4 c;    // qhi_y, qhi_z (uint bits), ref0, ref1
  int4 d;    // ref2, ref3, cnt0 | cnt1 << 16, cnt2 | cnt3 << 16 (leaf element counts, 0 for inner)
};
static_assert(sizeof(QNode4) == 64, "QNode4 must be 64 bytes");

// Leaf bound (mesh BVHs), stored at the position of the leaf's first triangle record: every point
// of the leaf's triangles lies within `t` of the plane through c with normal n (n.(x - c), n as
// stored) and within `r` of the axis through c along n. Only pruning reads it.
struct __align__(16) LeafBound {
  float4 a;    // c.xyz, r
  float4 b;    // n.xyz, t
};

struct __align__(16) PointRec {
  float4 p;    // xyz, id (int bits)
};

}  // namespace thg

// Per-element float arithmetic of Thea's query path, restated for HIP device code.
//
// Every expression keeps Thea's operation order so results are bit-identical to Thea's float
// path compiled with Code/Build/Common/BuildSetup.cmake:96-111 (-O2, no fast-math, SSE2, no FMA):
//  - no contraction of a*b+c into FMA (pragma below, and -ffp-contract=off on the command line);
//  - IEEE-correct division and square root (HIP default, -fhip-fp32-correctly-rounded-divide-sqrt);
//  - 3-vector reductions in Eigen 3.4's fixed-size order x0 + (x1 + x2) (Thea's Vector3 is Eigen);
//  - Eigen normalized(): each component divided by sqrt(squaredNorm()), only if squaredNorm() > 0.
// What is precomputed per triangle (plane, 2-D inside-test area and its reciprocal, perimeter edge
// directions and lengths) is computed with exactly the operations Thea performs per query, so
// caching changes no bits.
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_types.hpp"

namespace thg {

#define THG_HD __host__ __device__ __forceinline__

THG_HD float3 f3(float x, float y, float z) { float3 r; r.x = x; r.y = y; r.z = z; return r; }
THG_HD float3 sub3(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
THG_HD float3 add3(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
THG_HD float3 scl3(float s, float3 a) { return f3(s * a.x, s * a.y, s * a.z); }
THG_HD float dot3(float3 a, float3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
THG_HD float sqn3(float3 a) { return dot3(a, a); }
THG_HD float3 cross3(float3 a, float3 b) {
  return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
THG_HD float comp(float3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
THG_HD float fabs_(float x) { return x < 0 ? -x : (x == 0 ? 0.0f : x); }

// Eigen normalized(): each component divided by sqrt(squaredNorm()), unchanged if that is 0.
THG_HD float3 normalize3(float3 c) {
  float nn = sqn3(c);
  if (nn > 0) {
    float s = sqrtf(nn);
    return f3(c.x / s, c.y / s, c.z / s);
  }
  return c;
}

// Hyperplane3.hpp:34-41 fromThreePoints normal, with Eigen normalized().
THG_HD float3 plane_normal(float3 v0, float3 v1, float3 v2) { return normalize3(cross3(sub3(v1, v0), sub3(v2, v0))); }

// MatrixUtil.hpp:108-115 maxAbsAxis (first maximum wins, as Eigen maxCoeff).
THG_HD int max_abs_axis(float3 n) {
  float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
  int k = 0;
  float m = ax;
  if (ay > m) { m = ay; k = 1; }
  if (az > m) k = 2;
  return k;
}

// Triangle3.hpp:154 area = 0.5 |n . (e01 x e02)|.
THG_HD float tri_area(float3 v0, float3 v1, float3 v2) {
  float3 n = plane_normal(v0, v1, v2);
  float3 e01 = sub3(v1, v0), e02 = sub3(v2, v0);
  return 0.5f * fabsf(dot3(n, cross3(e01, e02)));
}

// In-plane axes of isPointInsideTriangle (Triangle3.hpp:1426-1431).
THG_HD void plane_axes(int axis, int &i, int &j) {
  if (axis == 1) { i = 2; j = 0; }
  else if (axis == 2) { i = 0; j = 1; }
  else { i = 1; j = 2; }
}

THG_HD TriRec make_trirec(float3 v0, float3 v1, float3 v2, int32_t id) {
  TriRec r;
  float3 n = plane_normal(v0, v1, v2);
  float pd = dot3(n, v0);
  int axis = max_abs_axis(n);
  int i, j;
  plane_axes(axis, i, j);
  // THEA_TRI3_AREA2(v0, v1, v2)
  float ar = (comp(v1, i) - comp(v0, i)) * (comp(v2, j) - comp(v0, j)) -
             (comp(v2, i) - comp(v0, i)) * (comp(v1, j) - comp(v0, j));
  float inv = (ar == 0) ? 0.0f : 1.0f / ar;
  float3 d0 = sub3(v1, v0), d1 = sub3(v2, v1), d2 = sub3(v0, v2);
  float l0 = sqrtf(sqn3(d0)), l1 = sqrtf(sqn3(d1)), l2 = sqrtf(sqn3(d2));
  d0 = f3(d0.x / l0, d0.y / l0, d0.z / l0);
  d1 = f3(d1.x / l1, d1.y / l1, d1.z / l1);
  d2 = f3(d2.x / l2, d2.y / l2, d2.z / l2);
  r.p0 = make_float4(v0.x, v0.y, v0.z, n.x);
  r.p1 = make_float4(v1.x, v1.y, v1.z, n.y);
  r.p2 = make_float4(v2.x, v2.y, v2.z, n.z);
  union { int i; float f; } m, idb;
  m.i = axis | (ar == 0 ? 4 : 0);
  idb.i = id;
  r.pl = make_float4(pd, inv, m.f, idb.f);
  r.e0 = make_float4(d0.x, d0.y, d0.z, l0);
  r.e1 = make_float4(d1.x, d1.y, d1.z, l1);
  r.e2 = make_float4(d2.x, d2.y, d2.z, l2);
  // Bounding sphere, centred in the triangle's plane when that plane is tight: the box centre is
  // moved along the normal onto n.x = pd (double) and the radius is the largest vertex distance
  // (double), rounded up, so the triangle lies in the disc of that radius about the centre. The
  // plane is tight when |n|^2 is within 2^-18 of 1 and every vertex is within 2^-20 of the largest
  // vertex coordinate of it; a degenerate or near-collinear triangle's float normal can be far from
  // perpendicular to its edges, and then the centre stays at the box centre and the radius is
  // stored negated, which tells the pruning tests to use the sphere alone.
  double cx = 0.5 * ((double)fminf(v0.x, fminf(v1.x, v2.x)) + (double)fmaxf(v0.x, fmaxf(v1.x, v2.x)));
  double cy = 0.5 * ((double)fminf(v0.y, fminf(v1.y, v2.y)) + (double)fmaxf(v0.y, fmaxf(v1.y, v2.y)));
  double cz = 0.5 * ((double)fminf(v0.z, fminf(v1.z, v2.z)) + (double)fmaxf(v0.z, fmaxf(v1.z, v2.z)));
  const float3 vv[3] = {v0, v1, v2};
  const double nn = (double)n.x * n.x + (double)n.y * n.y + (double)n.z * n.z;
  bool tight = nn > 1.0 - 1.0 / 262144.0 && nn < 1.0 + 1.0 / 262144.0;
  double vmax = 0, off = 0;
  for (int k = 0; k < 3; ++k) {
    vmax = fmax(vmax, fmax(fabs((double)vv[k].x), fmax(fabs((double)vv[k].y), fabs((double)vv[k].z))));
    off = fmax(off, fabs((double)n.x * vv[k].x + (double)n.y * vv[k].y + (double)n.z * vv[k].z - (double)pd));
  }
  tight = tight && off <= vmax * (1.0 / 1048576.0);
  if (tight) {
    const double s = ((double)n.x * cx + (double)n.y * cy + (double)n.z * cz - (double)pd) / nn;
    cx -= s * n.x;
    cy -= s * n.y;
    cz -= s * n.z;
  }
  float fcx = (float)cx, fcy = (float)cy, fcz = (float)cz;
  double rr = 0;
  for (int k = 0; k < 3; ++k) {
    double dx = (double)vv[k].x - fcx, dy = (double)vv[k].y - fcy, dz = (double)vv[k].z - fcz;
    double dd = dx * dx + dy * dy + dz * dz;
    rr = dd > rr ? dd : rr;
  }
  float rad = (float)(sqrt(rr) * (1.0 + 1e-6));
  if (!tight) rad = -rad;
  r.sb = make_float4(fcx, fcy, fcz, rad);
  return r;
}

// Triangle3.hpp:1310-1338 closestPointOnLineSegment.
THG_HD float3 seg_closest(float3 a, float3 b, float3 dir, float len, float3 p) {
  float3 c = sub3(p, a);
  float t = dot3(dir, c);
  if (t <= 0) return a;
  if (t >= len) return b;
  return add3(a, scl3(t, dir));
}

// Thea's TriangleN::closestPoint (Triangle3.hpp:294-304) followed by MetricL2's
// (cp - p).squaredNorm() (MetricL2.hpp:209-214). Returns d^2, writes the closest point.
THG_HD float tri_closest(const float4 p0, const float4 p1, const float4 p2, const float4 pl, const float4 e0,
                         const float4 e1, const float4 e2, float3 p, float3 &cp) {
  float3 v0 = f3(p0.x, p0.y, p0.z), v1 = f3(p1.x, p1.y, p1.z), v2 = f3(p2.x, p2.y, p2.z);
  float3 n = f3(p0.w, p1.w, p2.w);
  // HyperplaneN.hpp:145-160: signedDistance = p.n - d; closestPoint = p - sd n.
  float sd = dot3(p, n) - pl.x;
  float3 proj = sub3(p, scl3(sd, n));
  int i, j;
  plane_axes(trirec_axis(pl), i, j);
  bool inside;
  // isPointInsideTriangle (Triangle3.hpp:1414-1462)
  if (trirec_flat(pl)) {
    inside = (v0.x == proj.x) && (v0.y == proj.y) && (v0.z == proj.z);
  } else {
    float pi = comp(proj, i), pj = comp(proj, j);
    float v0i = comp(v0, i), v0j = comp(v0, j), v1i = comp(v1, i), v1j = comp(v1, j);
    float v2i = comp(v2, i), v2j = comp(v2, j);
    float b0 = ((v1i - pi) * (v2j - pj) - (v2i - pi) * (v1j - pj)) * pl.y;
    inside = !(b0 < 0 || b0 > 1);
    if (inside) {
      float b1 = ((pi - v0i) * (v2j - v0j) - (v2i - v0i) * (pj - v0j)) * pl.y;
      inside = !(b1 < 0 || b1 > 1);
      if (inside) {
        float b2 = 1.0f - b0 - b1;
        inside = (b2 >= 0) && (b2 <= 1);
      }
    }
  }
  if (inside) {
    cp = proj;
  } else {
    // closestPointOnTrianglePerimeter (Triangle3.hpp:1340-1412), cached directions/lengths.
    float3 r0 = seg_closest(v0, v1, f3(e0.x, e0.y, e0.z), e0.w, p);
    float3 r1 = seg_closest(v1, v2, f3(e1.x, e1.y, e1.z), e1.w, p);
    float3 r2 = seg_closest(v2, v0, f3(e2.x, e2.y, e2.z), e2.w, p);
    float d0 = sqn3(sub3(r0, p)), d1 = sqn3(sub3(r1, p)), d2 = sqn3(sub3(r2, p));
    if (d0 < d1) cp = (d0 < d2) ? r0 : r2;
    else cp = (d1 < d2) ? r1 : r2;
  }
  return sqn3(sub3(cp, p));
}

// The same arithmetic as tri_closest with the primary axis known at compile time (no per-lane
// component selects), taking the plane-distance term sd = p.n - pd precomputed by the caller, and
// returning the squared distance the perimeter search already computed for the chosen edge
// (Thea recomputes (cp - p).squaredNorm() with the same operations, so the bits are the same).
template <int AX>
__device__ __forceinline__ float tri_closest_ax(const float4 &p0, const float4 &p1, const float4 &p2, const float4 &pl,
                                                const float4 &e0, const float4 &e1, const float4 &e2, float3 p,
                                                float sd, float3 &cp) {
  constexpr int I = AX == 1 ? 2 : (AX == 2 ? 0 : 1);
  constexpr int J = AX == 1 ? 0 : (AX == 2 ? 1 : 2);
  const float3 v0 = f3(p0.x, p0.y, p0.z), v1 = f3(p1.x, p1.y, p1.z), v2 = f3(p2.x, p2.y, p2.z);
  const float3 proj = f3(p.x - sd * p0.w, p.y - sd * p1.w, p.z - sd * p2.w);
  bool inside;
  if (trirec_flat(pl)) {
    inside = (v0.x == proj.x) && (v0.y == proj.y) && (v0.z == proj.z);
  } else {
    const float pi = comp(proj, I), pj = comp(proj, J);
    const float v0i = comp(v0, I), v0j = comp(v0, J), v1i = comp(v1, I), v1j = comp(v1, J);
    const float v2i = comp(v2, I), v2j = comp(v2, J);
    const float b0 = ((v1i - pi) * (v2j - pj) - (v2i - pi) * (v1j - pj)) * pl.y;
    const float b1 = ((pi - v0i) * (v2j - v0j) - (v2i - v0i) * (pj - v0j)) * pl.y;
    const float b2 = 1.0f - b0 - b1;
    inside = !(b0 < 0 || b0 > 1) && !(b1 < 0 || b1 > 1) && (b2 >= 0) && (b2 <= 1);
  }
  if (inside) {
    cp = proj;
    return sqn3(sub3(proj, p));
  }
  const float3 r0 = seg_closest(v0, v1, f3(e0.x, e0.y, e0.z), e0.w, p);
  const float3 r1 = seg_closest(v1, v2, f3(e1.x, e1.y, e1.z), e1.w, p);
  const float3 r2 = seg_closest(v2, v0, f3(e2.x, e2.y, e2.z), e2.w, p);
  const float d0 = sqn3(sub3(r0, p)), d1 = sqn3(sub3(r1, p)), d2 = sqn3(sub3(r2, p));
  if (d0 < d1) {
    if (d0 < d2) { cp = r0; return d0; }
    cp = r2;
    return d2;
  }
  if (d1 < d2) { cp = r1; return d1; }
  cp = r2;
  return d2;
}

THG_HD float tri_closest(const TriRec &r, float3 p, float3 &cp) {
  return tri_closest(r.p0, r.p1, r.p2, r.pl, r.e0, r.e1, r.e2, p, cp);
}

// Triangle3.hpp:1464-1524 rayTriangleIntersectionTime (Eberly) with edges e01 = v1 - v0 and
// e02 = v2 - v0 (Triangle3.hpp:151-152), then TriangleN::rayIntersection's max_time rule
// (Triangle3.hpp:542-552). Returns t >= 0 on a hit, -1 otherwise.
// Same test from v0 and the edges (RayTri): e01 and e02 computed beforehand with the same
// subtractions give the same bits.
THG_HD float tri_ray_e(float3 v0, float3 e01, float3 e02, float3 o, float3 dir, float max_time) {
  const float EPS = 1.19209290e-07f;  // FLT_EPSILON, Math::eps<float>()
  float3 diff = sub3(o, v0);
  float3 normal = cross3(e01, e02);
  float DdN = dot3(dir, normal);
  float sign;
  if (DdN > EPS) sign = 1.0f;
  else if (DdN < -EPS) { sign = -1.0f; DdN = -DdN; }
  else return -1.0f;
  float DdQxE2 = sign * dot3(dir, cross3(diff, e02));
  if (DdQxE2 >= 0) {
    float DdE1xQ = sign * dot3(dir, cross3(e01, diff));
    if (DdE1xQ >= 0) {
      if (DdQxE2 + DdE1xQ <= DdN) {
        float QdN = (-sign) * dot3(diff, normal);
        if (QdN >= 0) {
          float t = QdN / DdN;
          if (t >= 0 && (max_time < 0 || t <= max_time)) return t;
        }
      }
    }
  }
  return -1.0f;
}

THG_HD float tri_ray(float3 v0, float3 v1, float3 v2, float3 o, float3 dir, float max_time) {
  return tri_ray_e(v0, sub3(v1, v0), sub3(v2, v0), o, dir, max_time);
}

}  // namespace thg

// Double-precision point BVH (BvhN<Vector3d, 3, double> over points): records and host entry points.
#pragma once

#include "common.hpp"
#include "query.hpp"

namespace thg {

// A point with the caller's double coordinates and index, in leaf order beside the float PointRec.
struct __align__(32) PointRecD {
  double x, y, z;
  int32_t id;
  int32_t pad;
};
static_assert(sizeof(PointRecD) == 32, "PointRecD must be 32 bytes");

// Builds the float LBVH of the points rounded to float and the leaf-order double records.
void build_points64(Bvh &b, const double *xyz, int64_t n, int leaf_size, hipStream_t s);
// Largest double below dist_bound^2 (or +inf when unbounded): admits exactly d^2 < dist_bound^2.
double bound_to_best64(double dist_bound);
// Nearest point of each double query (lowest index on exact ties); outputs nullable except idx.
void launch_closest64(const Bvh &b, const double *q, int64_t n, double dist_bound, int32_t *idx, double *cp,
                      double *dist, double *d2, QueryScratch &ws, hipStream_t s);
void order_queries64(const Bvh &b, const double *q, int64_t n, QueryScratch &ws, hipStream_t s);

}  // namespace thg

// thea_gpu_bvh.hpp -- header-only C++ host API over libthea_gpu's C ABI (thea_gpu.h), keeping the
// names and argument conventions of Thea's BVH query and sampling path so templated callers
// (Algorithms/Icp3.hpp:282-296, ShapeDiameter.hpp:214-217, Tools/MeshSample, SampleGraph, ShapeAlign)
// compile against it with a typedef change:
//
//   Thea                                                  here
//   MeshBvh<Mesh>::closestElement<MetricL2>(q, bound,     MeshBvh::closestElement<MetricL2>(q, bound,
//       compat, &dist, &closest_point)                        compat, &dist, &closest_point)
//     (Algorithms/BvhN.hpp:1050-1063)
//   BvhN::closestPair<MetricL2>(q, bound, compat, true)   MeshBvh::closestPair<MetricL2>(...)
//   BvhN::kClosestPairs<MetricL2>(q, set, bound, compat)  MeshBvh::kClosestPairs<MetricL2>(q, set, bound)
//   BvhN::rayStructureIntersection<Tester>(ray, max_t)    MeshBvh::rayStructureIntersection<Tester>(ray, max_t)
//   BvhN::rayIntersectionTime / rayIntersects             same names
//   MeshSampler::sampleEvenlyByArea(n, pos, nrm, tris)    MeshSampler::sampleEvenlyByArea(n, pos, nrm, tris)
//
// Single-query calls run as a batch of one (correct, latency bound); the batch overloads
// (closestElements, kClosestElements, rayIntersections, ...) are the fast path. Vector types only
// need operator[] (Thea's Eigen Vector3 works). Failures throw thea_gpu::Error (Thea throws
// Thea::Error, Error.hpp:27); "no result" is index -1 / time -1 as in Thea.
#ifndef THEA_GPU_BVH_HPP
#define THEA_GPU_BVH_HPP

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <type_traits>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "thea_gpu.h"

namespace thea_gpu {

typedef std::int64_t intx;

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int status) {
  if (status != THG_OK) throw Error(status, thg_last_error());
}

// Tag types mirroring Thea's template arguments (Algorithms/MetricL2.hpp, Filter.hpp:68-73,
// RayIntersectionTester.hpp). Only the L2 metric and universal compatibility exist on this path.
struct MetricL2 {};
struct UniversalCompatibility {};
struct RayIntersectionTester {};

// Result of a proximity query (Thea's NeighborPair, ProximityQueryStructureN.hpp:48-144). Pairs
// order by (squared distance, query index, target index), the lowest-index rule of this library.
struct NeighborPair {
  intx query_index = -1, target_index = -1;
  double mon_approx_dist = -1;  // squared distance (MetricL2's monotone approximation)
  float query_point[3] = {0, 0, 0}, target_point[3] = {0, 0, 0};  // filled when closest points are asked for
  bool isValid() const { return query_index >= 0 && target_index >= 0; }
  intx getTargetIndex() const { return target_index; }
  intx getQueryIndex() const { return query_index; }
  double getMonotoneApproxDistance() const { return mon_approx_dist; }
  template <typename MetricT = MetricL2> double getDistance() const { return mon_approx_dist < 0 ? -1 : std::sqrt(mon_approx_dist); }
  const float *getQueryPoint() const { return query_point; }
  const float *getTargetPoint() const { return target_point; }
  bool operator<(NeighborPair const &o) const {
    if (mon_approx_dist != o.mon_approx_dist) return mon_approx_dist < o.mon_approx_dist;
    if (query_index != o.query_index) return query_index < o.query_index;
    return target_index < o.target_index;
  }
  bool operator==(NeighborPair const &o) const { return query_index == o.query_index && target_index == o.target_index; }
};

// Largest k the GPU k-nearest kernels keep per query (thg_k_closest).
constexpr int MAX_K = 32;

// Fixed-capacity sorted array (Thea's BoundedSortedArrayN, BoundedSortedArrayN.hpp): keeps the N
// smallest values inserted, ascending. Thea's own BoundedSortedArrayN / BoundedSortedArray work with
// kClosestPairs too; any type with getCapacity(), clear() and insert() does.
template <std::size_t N, typename T> class BoundedSortedArrayN {
 public:
  static constexpr std::size_t getCapacity() { return N; }
  std::size_t size() const { return n_; }
  bool empty() const { return n_ == 0; }
  T const &operator[](std::size_t i) const { return v_[i]; }
  T const &front() const { return v_[0]; }
  T const &back() const { return v_[n_ - 1]; }
  const T *begin() const { return v_; }
  const T *end() const { return v_ + n_; }
  void clear() { n_ = 0; }
  bool isInsertable(T const &t) const { return n_ < N || t < back(); }
  std::size_t insert(T const &t) {
    if (!isInsertable(t)) return N;
    std::size_t pos = (std::size_t)(std::upper_bound(v_, v_ + n_, t) - v_);
    std::size_t last = n_ < N ? n_ : N - 1;
    for (std::size_t i = last; i > pos; --i) v_[i] = v_[i - 1];
    v_[pos] = t;
    if (n_ < N) ++n_;
    return pos;
  }

 private:
  T v_[N];
  std::size_t n_ = 0;
};

// Result of a ray query (Thea's RayStructureIntersectionN, RayQueryStructureN.hpp:29-53).
struct RayStructureIntersection3 {
  float time = -1;
  float normal[3] = {0, 0, 0};
  intx element_index = -1;
  bool isValid() const { return time >= 0; }
  float getTime() const { return time; }
  intx getElementIndex() const { return element_index; }
};

// Thea's Ray3 stand-in: origin and direction with operator[].
struct Ray3 {
  float o[3], d[3];
  Ray3(float ox, float oy, float oz, float dx, float dy, float dz) : o{ox, oy, oz}, d{dx, dy, dz} {}
  template <typename V> Ray3(V const &origin, V const &dir)
      : o{(float)origin[0], (float)origin[1], (float)origin[2]}, d{(float)dir[0], (float)dir[1], (float)dir[2]} {}
};

class Context {
 public:
  explicit Context(int device = 0) { check(thg_ctx_create(device, &h_)); }
  ~Context() { if (h_) thg_ctx_destroy(h_); }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  thg_ctx handle() const { return h_; }

 private:
  thg_ctx h_ = nullptr;
};

// Triangle mesh (vertices float3, triangles int32 triplets), resident on the GPU.
class Mesh {
 public:
  Mesh(Context &ctx, const float *verts, intx nv, const std::int32_t *tris, intx nt) {
    check(thg_mesh_create(ctx.handle(), verts, nv, tris, nt, THG_MEM_HOST, &h_));
  }
  ~Mesh() { if (h_) thg_mesh_destroy(h_); }
  Mesh(const Mesh &) = delete;
  Mesh &operator=(const Mesh &) = delete;
  thg_mesh handle() const { return h_; }
  intx numTriangles() const { return thg_mesh_num_triangles(h_); }

 private:
  thg_mesh h_ = nullptr;
};

class BvhBase;
class PointBvh;

namespace detail {
// bounded neighbour sets (Thea's BoundedSortedArrayN / BoundedSortedArray) expose getCapacity()
template <typename T, typename = void> struct has_capacity : std::false_type {};
template <typename T>
struct has_capacity<T, decltype((void)std::declval<T const &>().getCapacity(), void())> : std::true_type {};
template <typename T> using if_bounded_set = typename std::enable_if<has_capacity<T>::value, int>::type;
template <typename T> using if_plain_list = typename std::enable_if<!has_capacity<T>::value, int>::type;
template <typename T> using if_not_bvh =
    typename std::enable_if<!std::is_base_of<BvhBase, typename std::decay<T>::type>::value, int>::type;
}  // namespace detail

// Shared query methods of mesh and point BVHs (BvhN.hpp:1050-1146).
class BvhBase {
 public:
  typedef thea_gpu::NeighborPair NeighborPair;  // BvhN::NeighborPair
  BvhBase(const BvhBase &) = delete;
  BvhBase &operator=(const BvhBase &) = delete;
  virtual ~BvhBase() { if (h_) thg_bvh_destroy(h_); }
  thg_bvh handle() const { return h_; }

  // BvhN::getBounds (BvhN.hpp:991): the union of the element boxes scaled by 1.05 about its centre
  // (BvhN.hpp:1437, AxisAlignedBoxN.hpp:274,316-324), in float arithmetic.
  void getBounds(float lo[3], float hi[3]) const {
    check(thg_bvh_bounds(h_, lo, hi));
    std::int64_t info[6];
    check(thg_bvh_info(h_, info));
    if (info[3] == 0) return;
    for (int a = 0; a < 3; ++a) {
      const float c = 0.5f * (hi[a] + lo[a]);
      lo[a] = 1.05f * (lo[a] - c) + c;
      hi[a] = 1.05f * (hi[a] - c) + c;
    }
  }

  // BvhN::enableNearestNeighborAcceleration / disableNearestNeighborAcceleration (BvhN.hpp:853-888).
  // Thea's proxy structure only seeds the search bound with an upper bound on the nearest distance,
  // so it never changes a result; the GPU kernels seed each query from its nearest leaf instead.
  // Accepted and ignored.
  void enableNearestNeighborAcceleration(intx num_acceleration_samples = -1) { (void)num_acceleration_samples; }
  void disableNearestNeighborAcceleration(bool deallocate_memory = true) { (void)deallocate_memory; }

  // Batch: idx[n], optional dist[n], closest points cp[3n].
  void closestElements(intx n, const float *q, std::int32_t *idx, float *dist = nullptr, float *cp = nullptr,
                       double dist_bound = -1) const {
    check(thg_closest_point(h_, q, n, dist_bound, idx, cp, dist, nullptr, THG_MEM_HOST, nullptr));
  }

  // BvhN::closestElement (BvhN.hpp:1050-1063): index of the nearest element or -1.
  template <typename MetricT = MetricL2, typename QueryT, typename CompatT = UniversalCompatibility,
            typename VectorT = QueryT>
  intx closestElement(QueryT const &query, double dist_bound = -1, CompatT = CompatT(), double *dist = nullptr,
                      VectorT *closest_point = nullptr) const {
    float q[3] = {(float)query[0], (float)query[1], (float)query[2]}, cp[3], d;
    std::int32_t idx;
    closestElements(1, q, &idx, &d, cp, dist_bound);
    if (idx >= 0) {
      if (dist) *dist = (double)d;
      if (closest_point) { (*closest_point)[0] = cp[0]; (*closest_point)[1] = cp[1]; (*closest_point)[2] = cp[2]; }
    }
    return idx;
  }

  // BvhN::closestElementDistance-style helper: distance or -1.
  template <typename MetricT = MetricL2, typename QueryT>
  double closestElementDistance(QueryT const &query, double dist_bound = -1) const {
    double d = -1;
    return closestElement<MetricT>(query, dist_bound, UniversalCompatibility(), &d) >= 0 ? d : -1;
  }

  // BvhN::closestPair (BvhN.hpp:1082-1114) with a point query.
  template <typename MetricT = MetricL2, typename QueryT, typename CompatT = UniversalCompatibility,
            detail::if_not_bvh<QueryT> = 0>
  NeighborPair closestPair(QueryT const &query, double dist_bound = -1, CompatT = CompatT(),
                           bool get_closest_points = false) const {
    (void)get_closest_points;
    float q[3] = {(float)query[0], (float)query[1], (float)query[2]}, d2;
    std::int32_t idx;
    check(thg_closest_point(h_, q, 1, dist_bound, &idx, nullptr, nullptr, &d2, THG_MEM_HOST, nullptr));
    NeighborPair p;
    if (idx >= 0) { p.query_index = 0; p.target_index = idx; p.mon_approx_dist = d2; }
    return p;
  }

  // Batch k-NN: idx[n*k], dist[n*k] (rows sorted by (distance, index); -1 padded).
  void kClosestElements(intx n, const float *q, int k, std::int32_t *idx, float *dist = nullptr, double dist_bound = -1,
                        bool exclude_self = false) const {
    check(thg_k_closest(h_, q, n, k, dist_bound, exclude_self ? 1 : 0, idx, dist, nullptr, THG_MEM_HOST, nullptr));
  }

  // BvhN::kClosestPairs (BvhN.hpp:1139-1146) with Thea's signature: the capacity of the bounded set
  // (Thea's BoundedSortedArrayN<k, NeighborPair> or BoundedSortedArray, or the one above) is k, at
  // most MAX_K; larger capacities fail with Error(THG_E_UNSUPPORTED). Pairs go in through
  // insert(), so pairs already in the set compete with the new ones as in Thea. Returns the set's size.
  template <typename MetricT = MetricL2, typename QueryT, typename BoundedSetT, typename CompatT = UniversalCompatibility,
            detail::if_bounded_set<BoundedSetT> = 0, detail::if_not_bvh<QueryT> = 0>
  intx kClosestPairs(QueryT const &query, BoundedSetT &k_closest_pairs, double dist_bound = -1, CompatT = CompatT(),
                     bool get_closest_points = false) const {
    (void)get_closest_points;
    const int k = checked_k((std::size_t)k_closest_pairs.getCapacity());
    float q[3] = {(float)query[0], (float)query[1], (float)query[2]};
    std::vector<std::int32_t> idx(k);
    std::vector<float> d2(k);
    check(thg_k_closest(h_, q, 1, k, dist_bound, 0, idx.data(), nullptr, d2.data(), THG_MEM_HOST, nullptr));
    for (int j = 0; j < k && idx[j] >= 0; ++j) {
      NeighborPair p;
      p.query_index = 0;
      p.target_index = idx[j];
      p.mon_approx_dist = d2[j];
      k_closest_pairs.insert(p);
    }
    return (intx)k_closest_pairs.size();
  }

  // Explicit-k form into a plain container with push_back() (e.g. std::vector<NeighborPair>).
  template <typename MetricT = MetricL2, typename QueryT, typename PairListT, typename CompatT = UniversalCompatibility,
            detail::if_plain_list<PairListT> = 0>
  intx kClosestPairs(QueryT const &query, PairListT &k_closest_pairs, int k, double dist_bound = -1,
                     CompatT = CompatT()) const {
    k = checked_k((std::size_t)(k > 0 ? k : 0));
    float q[3] = {(float)query[0], (float)query[1], (float)query[2]};
    std::vector<std::int32_t> idx(k);
    std::vector<float> d2(k);
    check(thg_k_closest(h_, q, 1, k, dist_bound, 0, idx.data(), nullptr, d2.data(), THG_MEM_HOST, nullptr));
    intx found = 0;
    for (int j = 0; j < k; ++j) {
      if (idx[j] < 0) break;
      NeighborPair p;
      p.query_index = 0;
      p.target_index = idx[j];
      p.mon_approx_dist = d2[j];
      k_closest_pairs.push_back(p);
      ++found;
    }
    return found;
  }

  // BvhN::closestPair and kClosestPairs with another point BVH as the query (BvhN.hpp:1082-1146,
  // TestBvh.cpp:256-297): the nearest pair, or the k nearest pairs, between the query BVH's points
  // and this BVH's elements. query_index is the query point's index, target_index this BVH's
  // element; ties go to the lower (query, target) index pair. Defined after PointBvh.
  template <typename MetricT = MetricL2, typename CompatT = UniversalCompatibility>
  NeighborPair closestPair(PointBvh const &query, double dist_bound = -1, CompatT = CompatT(),
                           bool get_closest_points = false) const;
  template <typename MetricT = MetricL2, typename BoundedSetT, typename CompatT = UniversalCompatibility,
            detail::if_bounded_set<BoundedSetT> = 0>
  intx kClosestPairs(PointBvh const &query, BoundedSetT &k_closest_pairs, double dist_bound = -1, CompatT = CompatT(),
                     bool get_closest_points = false) const;

 protected:
  BvhBase() = default;
  virtual const float *hostPoints() const { return nullptr; }  // host copy of point elements, if any
  static int checked_k(std::size_t k) {
    if (k == 0) throw Error(THG_E_INVALID, "kClosestPairs: the neighbour set has capacity 0");
    if (k > (std::size_t)MAX_K)
      throw Error(THG_E_UNSUPPORTED, "kClosestPairs: capacity " + std::to_string(k) +
                                         " exceeds the GPU k-nearest limit of " + std::to_string(MAX_K));
    return (int)k;
  }
  thg_bvh h_ = nullptr;
};

// BVH over a triangle mesh (Thea's MeshBvh<Mesh>, Algorithms/MeshBvh.hpp:31-165).
class MeshBvh : public BvhBase {
 public:
  explicit MeshBvh(Mesh &mesh, int max_elems_in_leaf = 4) : mesh_(&mesh) {
    thg_build_opts o = {};
    o.leaf_size = max_elems_in_leaf;
    check(thg_bvh_build_mesh(mesh.handle(), &o, nullptr, &h_));
  }

  // Batch first hits: idx[n], optional t[n], normal[3n].
  void rayIntersections(intx n, const float *origins, const float *dirs, std::int32_t *idx, float *t = nullptr,
                        float *normal = nullptr, float max_time = -1) const {
    check(thg_ray_first_hit(h_, origins, dirs, n, max_time, idx, t, normal, THG_MEM_HOST, nullptr));
  }

  // BvhN::rayStructureIntersection (BvhN.hpp:1291-1316).
  template <typename TesterT = RayIntersectionTester>
  RayStructureIntersection3 rayStructureIntersection(Ray3 const &ray, float max_time = -1) const {
    RayStructureIntersection3 r;
    std::int32_t idx;
    float t, nrm[3];
    rayIntersections(1, ray.o, ray.d, &idx, &t, nrm, max_time);
    if (idx >= 0) { r.time = t; r.element_index = idx; r.normal[0] = nrm[0]; r.normal[1] = nrm[1]; r.normal[2] = nrm[2]; }
    return r;
  }

  // BvhN::rayIntersectionTime (BvhN.hpp:1271-1289): hit time or -1.
  template <typename TesterT = RayIntersectionTester>
  float rayIntersectionTime(Ray3 const &ray, float max_time = -1) const {
    return rayStructureIntersection<TesterT>(ray, max_time).time;
  }

  // BvhN::rayIntersects (BvhN.hpp:1266-1269).
  template <typename TesterT = RayIntersectionTester> bool rayIntersects(Ray3 const &ray, float max_time = -1) const {
    std::uint8_t hit = 0;
    check(thg_ray_any_hit(h_, ray.o, ray.d, 1, max_time, &hit, THG_MEM_HOST, nullptr));
    return hit != 0;
  }

  Mesh &mesh() const { return *mesh_; }

 private:
  Mesh *mesh_;
};

// BVH over points (Thea's BvhN<Vector3, 3>). Keeps a host copy of the points, which serve as the
// queries when this BVH is the query side of closestPair / kClosestPairs.
class PointBvh : public BvhBase {
 public:
  PointBvh(Context &ctx, const float *xyz, intx n, int max_elems_in_leaf = 4) : xyz_(xyz, xyz + 3 * (n > 0 ? n : 0)) {
    thg_build_opts o = {};
    o.leaf_size = max_elems_in_leaf;
    check(thg_bvh_build_points(ctx.handle(), xyz, n, THG_MEM_HOST, &o, nullptr, &h_));
  }
  intx numElements() const { return (intx)(xyz_.size() / 3); }
  const float *points() const { return xyz_.data(); }

 protected:
  const float *hostPoints() const override { return xyz_.data(); }

 private:
  std::vector<float> xyz_;
};

// Thea's BvhN<Vector3d, 3, double> over a point set (the target of ShapeAlign's ICP,
// Tools/ShapeAlign/ShapeAlign.cpp:339): closestElement / closestElements search in double. The
// float methods of BvhBase also work on it and search the points rounded to float.
class PointBvhD : public BvhBase {
 public:
  PointBvhD(Context &ctx, const double *xyz, intx n, int max_elems_in_leaf = 4) : n_(n > 0 ? n : 0) {
    thg_build_opts o = {};
    o.leaf_size = max_elems_in_leaf;
    check(thg_bvh_build_points_f64(ctx.handle(), xyz, n, THG_MEM_HOST, &o, nullptr, &h_));
  }
  intx numElements() const { return n_; }

  // Batch: idx[n], optional dist[n] and closest points cp[3n], all double.
  void closestElements(intx n, const double *q, std::int32_t *idx, double *dist = nullptr, double *cp = nullptr,
                       double dist_bound = -1) const {
    check(thg_closest_point_f64(h_, q, n, dist_bound, idx, cp, dist, nullptr, THG_MEM_HOST, nullptr));
  }

  // BvhN<Vector3d>::closestElement (BvhN.hpp:1050-1063): index of the nearest point or -1.
  template <typename MetricT = MetricL2, typename QueryT, typename CompatT = UniversalCompatibility,
            typename VectorT = QueryT>
  intx closestElement(QueryT const &query, double dist_bound = -1, CompatT = CompatT(), double *dist = nullptr,
                      VectorT *closest_point = nullptr) const {
    const double q[3] = {(double)query[0], (double)query[1], (double)query[2]};
    double cp[3], d;
    std::int32_t idx;
    closestElements(1, q, &idx, &d, cp, dist_bound);
    if (idx >= 0) {
      if (dist) *dist = d;
      if (closest_point) { (*closest_point)[0] = cp[0]; (*closest_point)[1] = cp[1]; (*closest_point)[2] = cp[2]; }
    }
    return idx;
  }

 private:
  intx n_;
};

template <typename MetricT, typename CompatT>
NeighborPair BvhBase::closestPair(PointBvh const &query, double dist_bound, CompatT, bool get_closest_points) const {
  const intx n = query.numElements();
  std::vector<std::int32_t> idx((std::size_t)n);
  std::vector<float> d2((std::size_t)n), cp(3 * (std::size_t)n);
  check(thg_closest_point(h_, query.points(), n, dist_bound, idx.data(), cp.data(), nullptr, d2.data(), THG_MEM_HOST,
                          nullptr));
  NeighborPair best;
  for (intx i = 0; i < n; ++i) {
    if (idx[i] < 0) continue;
    NeighborPair p;
    p.query_index = i;
    p.target_index = idx[i];
    p.mon_approx_dist = d2[i];
    if (!best.isValid() || p < best) {
      if (get_closest_points)
        for (int c = 0; c < 3; ++c) { p.query_point[c] = query.points()[3 * i + c]; p.target_point[c] = cp[3 * i + c]; }
      best = p;
    }
  }
  return best;
}

template <typename MetricT, typename BoundedSetT, typename CompatT, detail::if_bounded_set<BoundedSetT>>
intx BvhBase::kClosestPairs(PointBvh const &query, BoundedSetT &k_closest_pairs, double dist_bound, CompatT,
                            bool get_closest_points) const {
  // the k nearest pairs overall are among the k nearest targets of each query point
  const int k = checked_k((std::size_t)k_closest_pairs.getCapacity());
  const intx n = query.numElements();
  std::vector<std::int32_t> idx((std::size_t)n * k);
  std::vector<float> d2((std::size_t)n * k);
  check(thg_k_closest(h_, query.points(), n, k, dist_bound, 0, idx.data(), nullptr, d2.data(), THG_MEM_HOST, nullptr));
  std::vector<NeighborPair> cand;
  for (intx i = 0; i < n; ++i)
    for (int j = 0; j < k && idx[(std::size_t)i * k + j] >= 0; ++j) {
      NeighborPair p;
      p.query_index = i;
      p.target_index = idx[(std::size_t)i * k + j];
      p.mon_approx_dist = d2[(std::size_t)i * k + j];
      cand.push_back(p);
    }
  const std::size_t keep = std::min<std::size_t>(cand.size(), (std::size_t)k);
  std::partial_sort(cand.begin(), cand.begin() + keep, cand.end());
  const float *targets = hostPoints();
  for (std::size_t j = 0; j < keep; ++j) {
    NeighborPair p = cand[j];
    if (get_closest_points) {
      // point elements: the closest points are the points themselves (target points are known
      // on the host for a PointBvh target; a MeshBvh target leaves them zero)
      for (int c = 0; c < 3; ++c) {
        p.query_point[c] = query.points()[3 * p.query_index + c];
        if (targets) p.target_point[c] = targets[3 * p.target_index + c];
      }
    }
    k_closest_pairs.insert(p);
  }
  return (intx)k_closest_pairs.size();
}

// Thea's MeshSampler (Algorithms/MeshSampler.hpp:33-336), area-weighted sampling in both count modes.
class MeshSampler {
 public:
  // MeshSampler::CountMode (MeshSampler.hpp:48-54): EXACT returns exactly n samples; APPROXIMATE
  // gives each triangle floor(n * area / total) samples plus one more with probability equal to the
  // fractional part (MeshSampler.hpp:205-241), so the count varies around n.
  enum class CountMode { EXACT, APPROXIMATE };
  explicit MeshSampler(Mesh &mesh, std::uint64_t seed = 0, std::uint64_t stream = 0)
      : mesh_(&mesh), seed_(seed), stream_(stream) {}

  // sampleEvenlyByArea (MeshSampler.hpp:100-105): positions (and optional face normals and
  // triangle indices) of the samples; returns their number (0 for an empty or zero-area mesh). Each
  // call advances the RNG stream so successive calls draw fresh samples.
  template <typename VectorT>
  intx sampleEvenlyByArea(intx n, std::vector<VectorT> &positions,
                          typename std::common_type<std::vector<VectorT>>::type *face_normals = nullptr,
                          std::vector<intx> *triangles = nullptr, CountMode count_mode = CountMode::EXACT,
                          bool verbose = false) {
    (void)verbose;
    const std::uint64_t stream = stream_++;
    std::int64_t cap = n > 0 ? n : 0;
    if (count_mode == CountMode::APPROXIMATE) {
      // the count is a function of (seed, stream): ask for it, then fill exactly that many rows
      check(thg_sample_area_approx(mesh_->handle(), n, seed_, stream, 0, nullptr, nullptr, nullptr, &cap, THG_MEM_HOST,
                                   nullptr));
    }
    std::vector<float> pos(3 * (size_t)cap), nrm(face_normals ? pos.size() : 0);
    std::vector<std::int32_t> tri(triangles ? (size_t)cap : 0);
    std::int64_t got = 0;
    if (count_mode == CountMode::APPROXIMATE) {
      if (cap > 0)
        check(thg_sample_area_approx(mesh_->handle(), n, seed_, stream, cap, pos.data(),
                                     face_normals ? nrm.data() : nullptr, triangles ? tri.data() : nullptr, &got,
                                     THG_MEM_HOST, nullptr));
    } else {
      check(thg_sample_area(mesh_->handle(), n, seed_, stream, pos.data(), face_normals ? nrm.data() : nullptr,
                            triangles ? tri.data() : nullptr, &got, THG_MEM_HOST, nullptr));
    }
    positions.resize((size_t)got);
    if (face_normals) face_normals->resize((size_t)got);
    if (triangles) triangles->resize((size_t)got);
    for (std::int64_t i = 0; i < got; ++i) {
      for (int c = 0; c < 3; ++c) {
        positions[i][c] = pos[3 * i + c];
        if (face_normals) (*face_normals)[i][c] = nrm[3 * i + c];
      }
      if (triangles) (*triangles)[i] = tri[i];
    }
    return got;
  }

  template <typename VectorT>
  intx sampleUniformRandom(intx n, std::vector<VectorT> &positions, std::vector<VectorT> *face_normals = nullptr,
                           std::vector<intx> *triangles = nullptr) {
    return sampleEvenlyByArea(n, positions, face_normals, triangles);
  }

 private:
  Mesh *mesh_;
  std::uint64_t seed_, stream_;
};

// Thea's Icp3<double> (Algorithms/Icp3.hpp:34-476) over a BVH as the proximity structure. align()
// returns the row-major 3x4 transform [L | t]; alignBatch() runs one alignment per initial transform
// in one call (the search over starting rotations of Tools/ShapeAlign/ShapeAlign.cpp:343-414).
class Icp3 {
 public:
  using Transform = std::array<double, 12>;
  explicit Icp3(double fractional_error_threshold = -1, intx min_iterations = -1, intx max_iterations = -1,
                bool verbose = false) {
    (void)verbose;
    opts_ = thg_icp_opts{};
    opts_.fractional_error_threshold = fractional_error_threshold;
    opts_.min_iterations = (std::int32_t)min_iterations;
    opts_.max_iterations = (std::int32_t)max_iterations;
  }
  void setUpVector(const double up[3]) {
    opts_.has_up = 1;
    for (int a = 0; a < 3; ++a) opts_.up[a] = up[a];
  }
  void clearUpVector() { opts_.has_up = 0; }
  bool hasUpVector() const { return opts_.has_up != 0; }
  // Least-squares rotation instead of Thea's U V^T (its transpose).
  void setLeastSquaresFit(bool on) { opts_.kabsch = on ? 1 : 0; }

  template <typename VectorT>
  Transform align(intx n, const VectorT *from, const BvhBase &to, double *error = nullptr) const {
    std::vector<Transform> tr;
    std::vector<double> err;
    alignBatch(n, from, to, std::vector<Transform>(), tr, &err);
    if (error) *error = err[0];
    return tr[0];
  }

  template <typename VectorT>
  void alignBatch(intx n, const VectorT *from, const BvhBase &to, const std::vector<Transform> &initial,
                  std::vector<Transform> &transforms, std::vector<double> *errors = nullptr,
                  std::vector<std::int32_t> *iterations = nullptr) const {
    std::vector<double> xyz(3 * (size_t)n);
    for (intx i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) xyz[3 * i + c] = (double)from[i][c];
    const int k = initial.empty() ? 1 : (int)initial.size();
    transforms.resize(k);
    std::vector<double> err(k);
    std::vector<std::int32_t> its(k);
    check(thg_icp_align(to.handle(), xyz.data(), n, k, initial.empty() ? nullptr : initial[0].data(), &opts_,
                        THG_MEM_HOST, nullptr, transforms[0].data(), err.data(), its.data()));
    if (errors) *errors = err;
    if (iterations) *iterations = its;
  }

 private:
  thg_icp_opts opts_;
};

// Thea's ShapeDiameter (Algorithms/SurfaceFeatures/Local/ShapeDiameter.hpp:40-265) over a
// caller-built MeshBvh, as its precomputed-BVH constructor (:110-120): a normalisation scale <= 0
// defaults to the diagonal of the BVH's bounds. compute() evaluates one point (a batch of one);
// computeBatch() is the fast path.
class ShapeDiameter {
 public:
  explicit ShapeDiameter(MeshBvh const *bvh, float normalization_scale = -1) : bvh_(bvh), scale_(normalization_scale) {
    if (!bvh_) throw Error(THG_E_INVALID, "ShapeDiameter: precomputed BVH cannot be null");
    if (scale_ <= 0) {
      float lo[3], hi[3];
      bvh_->getBounds(lo, hi);
      const float e[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
      scale_ = std::sqrt(e[0] * e[0] + (e[1] * e[1] + e[2] * e[2]));
    }
  }
  float getScale() const { return scale_; }

  // compute(position): the normal of the nearest face (:162-176).
  template <typename VectorT> double compute(VectorT const &position, bool only_hit_interior_surfaces = true) const {
    const float p[3] = {(float)position[0], (float)position[1], (float)position[2]};
    double sdf = -1;
    computeBatch(1, p, nullptr, &sdf, only_hit_interior_surfaces);
    return sdf;
  }

  // compute(position, normal) with an outward normal (:185-256).
  template <typename VectorT>
  double compute(VectorT const &position, VectorT const &normal, bool only_hit_interior_surfaces = true) const {
    const float p[3] = {(float)position[0], (float)position[1], (float)position[2]};
    const float nm[3] = {(float)normal[0], (float)normal[1], (float)normal[2]};
    double sdf = -1;
    computeBatch(1, p, nm, &sdf, only_hit_interior_surfaces);
    return sdf;
  }

  // positions[3n], normals[3n] or nullptr (nearest face normals), sdf[n].
  void computeBatch(intx n, const float *positions, const float *normals, double *sdf,
                    bool only_hit_interior_surfaces = true) const {
    check(thg_shape_diameter(bvh_->handle(), positions, normals, n, scale_, only_hit_interior_surfaces ? 1 : 0,
                             THG_MEM_HOST, nullptr, sdf, nullptr));
  }

 private:
  MeshBvh const *bvh_;
  float scale_;
};

}  // namespace thea_gpu

#endif  // THEA_GPU_BVH_HPP

// Exercises include/thea_gpu_bvh.hpp the way a Thea caller would (single-query template calls and
// batch calls) on the unit cube of Data/Models/cube.obj; prints one line per result for
// tests/test_gpu_shim.py to compare with the Python binding. With a file argument (TestBvh's inputs:
// 10,000 points, 10,000 more points, 1,000 triangles and the ray direction, raw float32), it also
// runs the call sites of Thea's Test/TestBvh.cpp (:204-297, :338-360, :405-439) unchanged but for
// the BVH typedef.
#include <array>
#include <cstdio>
#include <vector>

#include "../../include/thea_gpu_bvh.hpp"

struct V3 {  // minimal vector with operator[] (Thea's Eigen Vector3 works the same way)
  float v[3];
  float &operator[](int i) { return v[i]; }
  float operator[](int i) const { return v[i]; }
};

static int test_bvh_calls(thea_gpu::Context &ctx, const char *path) {
  using namespace thea_gpu;
  const int NUM_POINTS = 10000, NUM_TRIANGLES = 1000;
  std::vector<float> buf(2 * 3 * NUM_POINTS + 9 * NUM_TRIANGLES + 3);
  FILE *f = std::fopen(path, "rb");
  if (!f || std::fread(buf.data(), sizeof(float), buf.size(), f) != buf.size()) {
    std::printf("error reading %s\n", path);
    if (f) std::fclose(f);
    return 2;
  }
  std::fclose(f);
  const float *points = buf.data(), *new_points = points + 3 * NUM_POINTS, *tri_xyz = new_points + 3 * NUM_POINTS;
  const float *ray_dir = tri_xyz + 9 * NUM_TRIANGLES;

  // points (TestBvh.cpp:159-236, 243-297) with `typedef BvhN<MyCustomPoint, 3> Bvh` -> PointBvh
  typedef PointBvh Bvh;
  Bvh bvh(ctx, points, NUM_POINTS);
  V3 query{{0.5f, 0.5f, 0.5f}};
  double dist_bound = 0.25;
  double dist = 0;
  intx nn_index = bvh.closestElement<MetricL2>(query, dist_bound, UniversalCompatibility(), &dist);
  std::printf("tb_nn %lld %.9g\n", (long long)nn_index, dist);
  BoundedSortedArrayN<3, Bvh::NeighborPair> nbrs;
  intx num_nbrs = bvh.kClosestPairs<MetricL2>(query, nbrs, dist_bound, UniversalCompatibility());
  for (size_t i = 0; i < nbrs.size(); ++i)
    std::printf("tb_k3 %lld %lld %.9g\n", (long long)num_nbrs, (long long)nbrs[i].getTargetIndex(),
                nbrs[i].getMonotoneApproxDistance());
  Bvh new_bvh(ctx, new_points, NUM_POINTS);
  typedef Bvh::NeighborPair NeighborPair;
  NeighborPair nn_pair = bvh.closestPair<MetricL2>(new_bvh, -1, UniversalCompatibility(), true);
  std::printf("tb_pair %lld %lld %.9g %.9g %.9g %.9g\n", (long long)nn_pair.getQueryIndex(),
              (long long)nn_pair.getTargetIndex(), nn_pair.getMonotoneApproxDistance(), nn_pair.getTargetPoint()[0],
              nn_pair.getTargetPoint()[1], nn_pair.getTargetPoint()[2]);
  nbrs.clear();
  num_nbrs = bvh.kClosestPairs<MetricL2>(new_bvh, nbrs, -1, UniversalCompatibility(), true);
  for (size_t i = 0; i < nbrs.size(); ++i)
    std::printf("tb_kpairs %lld %lld %lld %.9g\n", (long long)num_nbrs, (long long)nbrs[i].getQueryIndex(),
                (long long)nbrs[i].getTargetIndex(), nbrs[i].getMonotoneApproxDistance());
  // a neighbour set larger than the GPU k-nearest limit is refused with a clear error
  BoundedSortedArrayN<MAX_K + 8, NeighborPair> big;
  try {
    bvh.kClosestPairs<MetricL2>(query, big, -1, UniversalCompatibility());
    std::printf("tb_bigk accepted\n");
  } catch (const Error &e) {
    std::printf("tb_bigk %d %s\n", e.code(), e.what());
  }

  // triangles (TestBvh.cpp:311-439): a triangle soup, NN acceleration switched on as TestBvh does
  std::vector<std::int32_t> tri_idx(3 * NUM_TRIANGLES);
  for (int i = 0; i < 3 * NUM_TRIANGLES; ++i) tri_idx[i] = i;
  Mesh soup(ctx, tri_xyz, 3 * NUM_TRIANGLES, tri_idx.data(), NUM_TRIANGLES);
  MeshBvh tbvh(soup);
  tbvh.enableNearestNeighborAcceleration();
  V3 closest_point{};
  nn_index = tbvh.closestElement<MetricL2>(query, dist_bound, UniversalCompatibility(), &dist, &closest_point);
  std::printf("tb_tnn %lld %.9g %.9g %.9g %.9g\n", (long long)nn_index, dist, closest_point[0], closest_point[1],
              closest_point[2]);
  V3 zero{{0, 0, 0}}, dir{{ray_dir[0], ray_dir[1], ray_dir[2]}};
  Ray3 ray(zero, dir);
  float max_time = -1;
  bool hits = tbvh.rayIntersects<RayIntersectionTester>(ray, max_time);
  float hit_time = tbvh.rayIntersectionTime<RayIntersectionTester>(ray, max_time);
  RayStructureIntersection3 isec = tbvh.rayStructureIntersection<RayIntersectionTester>(ray, max_time);
  std::printf("tb_ray %d %.9g %lld %.9g\n", (int)hits, hit_time, (long long)isec.getElementIndex(), isec.getTime());
  tbvh.disableNearestNeighborAcceleration();
  return 0;
}

int main(int argc, char **argv) {
  using namespace thea_gpu;
  const float verts[] = {0, 0, 0, 1, 1, 0, 1, 0, 0, 0, 1, 0, 0, 1, 1, 1, 1, 1, 0, 0, 1, 1, 0, 1};
  const std::int32_t tris[] = {0, 1, 2, 0, 3, 1, 0, 4, 3, 0, 6, 4, 3, 5, 1, 3, 4, 5,
                               2, 1, 5, 2, 5, 7, 0, 2, 7, 0, 7, 6, 6, 7, 5, 6, 5, 4};
  try {
    Context ctx(0);
    Mesh mesh(ctx, verts, 8, tris, 12);
    MeshBvh bvh(mesh);
    for (int i = 0; i < 5; ++i) {
      V3 q{{-0.5f + 0.5f * i, 0.3f, 1.7f - 0.4f * i}}, cp{};
      double dist = -1;
      intx idx = bvh.closestElement<MetricL2>(q, -1, UniversalCompatibility(), &dist, &cp);
      std::printf("closest %d %lld %.9g %.9g %.9g %.9g\n", i, (long long)idx, dist, cp[0], cp[1], cp[2]);
    }
    std::vector<NeighborPair> nbrs;
    V3 q{{0.5f, 0.5f, 2.0f}};
    bvh.kClosestPairs<MetricL2>(q, nbrs, 3);
    for (auto &p : nbrs) std::printf("knn %lld %.9g\n", (long long)p.getTargetIndex(), p.getDistance());
    RayStructureIntersection3 hit = bvh.rayStructureIntersection(Ray3(0.25f, 0.5f, 3.0f, 0.0f, 0.0f, -1.0f));
    std::printf("ray %lld %.9g %d\n", (long long)hit.getElementIndex(), hit.getTime(),
                (int)bvh.rayIntersects(Ray3(0.25f, 0.5f, 3.0f, 0.0f, 0.0f, 1.0f)));
    MeshSampler sampler(mesh, 7, 0);
    std::vector<V3> pos, nrm;
    std::vector<intx> tri;
    intx got = sampler.sampleEvenlyByArea(4, pos, &nrm, &tri);
    for (intx i = 0; i < got; ++i) std::printf("sample %lld %.9g %.9g %.9g\n", (long long)tri[i], pos[i][0], pos[i][1], pos[i][2]);
    // APPROXIMATE count mode (MeshSampler.hpp:205-241), 40 samples requested, stream 5 of seed 7
    MeshSampler approx(mesh, 7, 5);
    std::vector<V3> apos;
    std::vector<intx> atri;
    intx agot = approx.sampleEvenlyByArea(40, apos, nullptr, &atri, MeshSampler::CountMode::APPROXIMATE);
    for (intx i = 0; i < agot; ++i)
      std::printf("approx %lld %.9g %.9g %.9g\n", (long long)atri[i], apos[i][0], apos[i][1], apos[i][2]);
    // ICP of the cube's surface samples, shifted, back onto the cube (ShapeAlign's Icp3 call)
    std::vector<V3> surf;
    sampler.sampleEvenlyByArea(2000, surf);
    std::vector<std::array<double, 3>> moved(surf.size());
    for (size_t i = 0; i < surf.size(); ++i)
      for (int c = 0; c < 3; ++c) moved[i][c] = surf[i][c] + (c == 0 ? 0.05 : 0.0);
    Icp3 icp;
    icp.setLeastSquaresFit(true);
    double err = -1;
    Icp3::Transform tr = icp.align((intx)moved.size(), moved.data(), bvh, &err);
    std::printf("icp %.6f %.6f %.6f %.3g\n", tr[3], tr[7], tr[11], err);
    // ShapeAlign's form (ShapeAlign.cpp:339,448): a double point BVH of the target samples and
    // Icp3<double> with Thea's defaults
    std::vector<double> target(3 * surf.size());
    for (size_t i = 0; i < surf.size(); ++i)
      for (int c = 0; c < 3; ++c) target[3 * i + c] = surf[i][c];
    PointBvhD to_bvh(ctx, target.data(), (intx)surf.size());
    std::array<double, 3> q0{{0.3, 0.2, 1.5}}, c0{};
    double d0 = -1;
    intx i0 = to_bvh.closestElement<MetricL2>(q0, -1, UniversalCompatibility(), &d0, &c0);
    std::printf("closest64 %lld %.17g %.17g %.17g %.17g\n", (long long)i0, d0, c0[0], c0[1], c0[2]);
    Icp3 icp_d;
    double err_d = -1;
    Icp3::Transform tr_d = icp_d.align((intx)moved.size(), moved.data(), to_bvh, &err_d);
    std::printf("icp64 %.17g %.17g %.17g %.17g\n", tr_d[3], tr_d[7], tr_d[11], err_d);
    // Shape diameter at the first samples (ShapeDiameter over a precomputed BVH, default scale)
    ShapeDiameter sdf(&bvh);
    std::printf("scale %.9g\n", sdf.getScale());
    for (intx i = 0; i < got; ++i)
      std::printf("sdf %lld %.17g %.17g\n", (long long)i, sdf.compute(pos[i], nrm[i]), sdf.compute(pos[i]));
    if (argc > 1) return test_bvh_calls(ctx, argv[1]);
  } catch (const Error &e) {
    std::printf("error %d %s\n", e.code(), e.what());
    return 2;
  }
  return 0;
}

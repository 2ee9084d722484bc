// Exercises include/thea_gpu_bvh.hpp the way a Thea caller would (single-query template calls and
// batch calls) on the unit cube of Data/Models/cube.obj; prints one line per result for
// tests/test_gpu_shim.py to compare with the Python binding.
#include <array>
#include <cstdio>
#include <vector>

#include "../../include/thea_gpu_bvh.hpp"

struct V3 {  // minimal vector with operator[] (Thea's Eigen Vector3 works the same way)
  float v[3];
  float &operator[](int i) { return v[i]; }
  float operator[](int i) const { return v[i]; }
};

int main() {
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
    // Shape diameter at the first samples (ShapeDiameter over a precomputed BVH, default scale)
    ShapeDiameter sdf(&bvh);
    std::printf("scale %.9g\n", sdf.getScale());
    for (intx i = 0; i < got; ++i)
      std::printf("sdf %lld %.17g %.17g\n", (long long)i, sdf.compute(pos[i], nrm[i]), sdf.compute(pos[i]));
  } catch (const Error &e) {
    std::printf("error %d %s\n", e.code(), e.what());
    return 2;
  }
  return 0;
}

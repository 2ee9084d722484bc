// Shape diameter function: Thea's ShapeDiameter::compute (Algorithms/SurfaceFeatures/Local/
// ShapeDiameter.hpp:185-256; Gal, Shamir and Cohen-Or, "Pose-Oblivious Shape Signature", TVCG 2007)
// for a batch of surface points against a mesh BVH, entirely on the device.
//
// Per point: in = -normalize(normal); rot = orthonormalBasis(in) (MatrixUtil.hpp:218-246); 30 rays
// from position + 0.001 scale in along rot * CONE_DIRS[i]; a hit counts if its face normal does not
// oppose the ray (only_interior). The reduction follows Thea exactly: the median by std::nth_element
// on the valid hit times (the permutation libstdc++ leaves, restated in select.hpp), the variance of
// all of them, and the weighted mean of the times within one standard deviation of the median, where
// weight i is the cone direction's z of the i-th valid ray while the times have been permuted by
// nth_element (ShapeDiameter.hpp:226-252). Sums run in double in Thea's order.
#include <cmath>

#include "common.hpp"
#include "device_types.hpp"
#include "kernels_common.hpp"
#include "query.hpp"
#include "select.hpp"
#include "tri_math.hpp"

namespace thg {

constexpr int SDF_RAYS = 30;

// ShapeDiameter.hpp:197-228: unit directions in a cone about +z.
__constant__ float c_cone[SDF_RAYS][3] = {
    {0.407497f, 0.134086f, 0.903309f},    {0.128722f, -0.822715f, 0.553688f},   {0.359298f, 0.686129f, 0.632560f},
    {0.813318f, -0.153956f, 0.561081f},   {-0.707169f, 0.296987f, 0.641647f},   {-0.478783f, -0.543441f, 0.689520f},
    {0.561083f, 0.523634f, 0.641087f},    {-0.380898f, -0.655789f, 0.651811f},  {0.722588f, 0.472155f, 0.504912f},
    {-0.535205f, 0.421235f, 0.732200f},   {0.577593f, -0.477896f, 0.661817f},   {0.393652f, -0.144484f, 0.907834f},
    {0.141508f, -0.660777f, 0.737122f},   {-0.274967f, -0.141703f, 0.950954f},  {-0.304935f, 0.474085f, 0.825989f},
    {0.711150f, 0.301383f, 0.635164f},    {-0.696119f, -0.112341f, 0.709082f},  {-0.248327f, 0.381035f, 0.890588f},
    {-0.0772394f, 0.0462501f, 0.995939f}, {0.438024f, 0.656463f, 0.614159f},    {-0.483384f, -0.542477f, 0.687065f},
    {0.498569f, -0.697466f, 0.514752f},   {-0.128902f, -0.0173429f, 0.991506f}, {0.338759f, -0.318772f, 0.885227f},
    {-0.574364f, -0.463887f, 0.674474f},  {-0.459207f, 0.260125f, 0.849390f},   {-0.628057f, -0.528928f, 0.570771f},
    {0.232916f, -0.328273f, 0.915416f},   {0.231667f, -0.796882f, 0.557951f},   {-0.270612f, -0.809654f, 0.520797f},
};
__device__ __forceinline__ float3 normalized3(float3 v) {  // v / |v| (Eigen normalized: divide by norm)
  const float n = sqrtf(sqn3(v));
  return f3(v.x / n, v.y / n, v.z / n);
}

// Face normal of triangle tri[i] (TriangleN::getNormal: the unit plane normal, Triangle3.hpp:143-155).
__global__ void __launch_bounds__(256) k_face_normals(const int32_t *__restrict__ tri, int64_t n,
                                                      const float *__restrict__ verts, const int32_t *__restrict__ tris,
                                                      float *__restrict__ nrm) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int64_t t = tri[i];
  float3 v[3];
  for (int k = 0; k < 3; ++k) {
    const int64_t vi = tris[3 * t + k];
    v[k] = f3(verts[3 * vi], verts[3 * vi + 1], verts[3 * vi + 2]);
  }
  const float3 nn = plane_normal(v[0], v[1], v[2]);
  nrm[3 * i] = nn.x;
  nrm[3 * i + 1] = nn.y;
  nrm[3 * i + 2] = nn.z;
}

// Rays are laid out cone-major over the points in Morton order: ray g = r * n + k is cone
// direction r of point order[k] (k itself without an order). A wave then holds one cone direction
// for 64 neighbouring points, whose rays are nearly parallel, instead of 30 directions of 2 points.
__device__ __forceinline__ int64_t sdf_point(const uint32_t *__restrict__ order, int64_t k) {
  return order ? (int64_t)order[k] : k;
}

// Positions and normals in the ray layout's point order (k-th row = point order[k]).
__global__ void __launch_bounds__(256) k_sdf_gather(int64_t n, const uint32_t *__restrict__ order,
                                                    const float *__restrict__ pos, const float *__restrict__ nrm,
                                                    float *__restrict__ spos, float *__restrict__ snrm) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const int64_t i = sdf_point(order, k);
  for (int c = 0; c < 3; ++c) {
    spos[3 * k + c] = pos[3 * i + c];
    snrm[3 * k + c] = nrm[3 * i + c];
  }
}

// One thread per ray: origin and direction of cone ray r of the k-th point (pos/nrm in that order).
__global__ void __launch_bounds__(256) k_sdf_rays(const float *__restrict__ pos, const float *__restrict__ nrm,
                                                  int64_t n, float scale, float *__restrict__ org,
                                                  float *__restrict__ dir) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= n * SDF_RAYS) return;
  const int r = (int)(g / n);
  const int64_t i = g - (int64_t)r * n;
  const float3 normal = f3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
  const float3 nn = normalized3(normal);
  const float3 in = f3(-nn.x, -nn.y, -nn.z);
  // orthonormalBasis(in): columns u, v, w
  const float3 w = normalized3(in);
  float3 u;
  if (max_abs_axis(in) == 0) u = normalized3(f3(in.y, -in.x, 0.0f));
  else u = normalized3(f3(0.0f, in.z, -in.y));
  const float3 v = cross3(w, u);
  const float c0 = c_cone[r][0], c1 = c_cone[r][1], c2 = c_cone[r][2];
  // rot * c, each row as m0 c0 + (m1 c1 + m2 c2)
  const float3 d = f3(u.x * c0 + (v.x * c1 + w.x * c2), u.y * c0 + (v.y * c1 + w.y * c2), u.z * c0 + (v.z * c1 + w.z * c2));
  const float s = 0.001f * scale;
  const float3 o = f3(pos[3 * i] + s * in.x, pos[3 * i + 1] + s * in.y, pos[3 * i + 2] + s * in.z);
  org[3 * g] = o.x;
  org[3 * g + 1] = o.y;
  org[3 * g + 2] = o.z;
  dir[3 * g] = d.x;
  dir[3 * g + 1] = d.y;
  dir[3 * g + 2] = d.z;
}

// value of ray g: its hit time if it hit and (when only_interior) the face normal does not oppose
// the ray, else -1.
__global__ void __launch_bounds__(256) k_sdf_valid(int64_t m, const int32_t *__restrict__ idx, const float *__restrict__ t,
                                                   const float *__restrict__ hn, const float *__restrict__ dir,
                                                   int only_interior, float *__restrict__ val) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= m) return;
  bool ok = idx[g] >= 0;
  if (ok && only_interior)
    ok = dot3(f3(hn[3 * g], hn[3 * g + 1], hn[3 * g + 2]), f3(dir[3 * g], dir[3 * g + 1], dir[3 * g + 2])) >= 0;
  val[g] = ok ? t[g] : -1.0f;
}

// Per-ray values from the ray layout to the caller's point-major layout (point i, ray r at 30 i + r).
__global__ void __launch_bounds__(256) k_sdf_times(int64_t n, const uint32_t *__restrict__ order,
                                                   const float *__restrict__ val, float *__restrict__ times) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= n * SDF_RAYS) return;
  const int r = (int)(g / n);
  times[sdf_point(order, g - (int64_t)r * n) * SDF_RAYS + r] = val[g];
}

constexpr int SDF_RED_THREADS = 256;

// One thread per point (in ray-layout order k): the point's 30 ray values are read with coalesced
// loads (ray r of the points k.. k+63 are consecutive), compacted in LDS to the valid ones with their
// ray numbers, and reduced as ShapeDiameter.hpp:230-256.
__global__ void __launch_bounds__(SDF_RED_THREADS) k_sdf_reduce(int64_t n, const float *__restrict__ val,
                                                                const uint32_t *__restrict__ order, double scale,
                                                                double *__restrict__ sdf) {
  __shared__ float sv[SDF_RAYS * SDF_RED_THREADS];
  __shared__ uint8_t sr[SDF_RAYS * SDF_RED_THREADS];
  const int t = threadIdx.x;
  const int64_t k = (int64_t)blockIdx.x * SDF_RED_THREADS + t;
  if (k >= n) return;
  float *mine = sv + t * SDF_RAYS;
  for (int r = 0; r < SDF_RAYS; ++r) mine[r] = val[(int64_t)r * n + k];
  const int64_t out = sdf_point(order, k);
  uint8_t *rays = sr + t * SDF_RAYS;
  int nv = 0;
  for (int r = 0; r < SDF_RAYS; ++r) {
    const float x = mine[r];
    if (x >= 0) {
      mine[nv] = x;
      rays[nv] = (uint8_t)r;
      ++nv;
    }
  }
  if (nv <= 0) {
    sdf[out] = -1.0;
    return;
  }
  const int mid = nv / 2;
  nth_element_gcc(StridedF{mine, 1}, mid, nv);
  double sum_values = 0, sum_squares = 0;
  for (int i = 0; i < nv; ++i) {
    const double x = mine[i];
    sum_values += x;
    sum_squares += x * x;
  }
  const double avg = sum_values / nv;
  const double var = sum_squares / nv - avg * avg;
  const double vmid = mine[mid];
  sum_values = 0;
  double sum_weights = 0;
  for (int i = 0; i < nv; ++i) {
    const double x = mine[i];
    const double dv = x - vmid;
    if (dv * dv <= var) {
      const double w = c_cone[rays[i]][2];
      sum_values += w * x;
      sum_weights += w;
    }
  }
  const double y = sum_weights > 0 ? sum_values / (sum_weights * scale) : vmid / scale;
  sdf[out] = fmin(1.0, fmax(0.0, y));
}

}  // namespace thg

using namespace thg;

template <typename F> static int sdf_guarded(F f) {
  try {
    f();
    return THG_OK;
  } catch (const Status &e) {
    set_last_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return THG_E_OOM;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return THG_E_DEVICE;
  }
}

extern "C" int thg_shape_diameter(thg_bvh mesh_bvh, const float *positions, const float *normals, int64_t n,
                                  float scale, int32_t only_interior, int32_t mem, void *stream, double *sdf,
                                  float *ray_times) {
  return sdf_guarded([&] {
    require(mesh_bvh != nullptr, "bvh is NULL");
    require(n >= 0 && (n == 0 || (positions && sdf)), "bad point or result array");
    require(scale > 0, "scale must be positive");
    require(mem == THG_MEM_DEVICE || mem == THG_MEM_HOST, "mem must be THG_MEM_DEVICE or THG_MEM_HOST");
    Bvh *b = reinterpret_cast<Bvh *>(mesh_bvh);
    require(!b->points, "the shape diameter needs a mesh BVH");
    THG_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (n == 0) return;
    const int64_t m = n * SDF_RAYS;
    AsyncBuf pos_d, nrm_d, org(m * 12, s), dir(m * 12, s), ridx(m * 4, s), rt(m * 4, s), rn(m * 12, s), val(m * 4, s);
    const float *d_pos = positions, *d_nrm = normals;
    if (mem == THG_MEM_HOST) {
      pos_d.ensure(n * 12, s);
      THG_HIP(hipMemcpyAsync(pos_d.p, positions, n * 12, hipMemcpyHostToDevice, s));
      d_pos = pos_d.as<float>();
      if (normals) {
        nrm_d.ensure(n * 12, s);
        THG_HIP(hipMemcpyAsync(nrm_d.p, normals, n * 12, hipMemcpyHostToDevice, s));
        d_nrm = nrm_d.as<float>();
      }
    }
    require(b->n > 0, "the mesh has no triangles");
    if (!d_nrm) {  // compute(position): the normal of the nearest face (ShapeDiameter.hpp:162-176)
      require(b->mesh != nullptr, "the BVH has no mesh for face normals");
      QueryScratch ws;
      AsyncBuf tri(n * 4, s);
      launch_closest(*b, d_pos, n, -1.0, 0, tri.as<int32_t>(), nullptr, nullptr, nullptr, ws, s);
      nrm_d.ensure(n * 12, s);
      hipLaunchKernelGGL(k_face_normals, dim3(ceil_div(n, 256)), dim3(256), 0, s, tri.as<int32_t>(), n,
                         b->mesh->verts.as<float>(), b->mesh->tris.as<int32_t>(), nrm_d.as<float>());
      launch_check("k_face_normals");
      d_nrm = nrm_d.as<float>();
    }
    // Morton order of the points (none for small batches), then the cone-major ray layout.
    QueryScratch pws;
    order_queries(*b, d_pos, n, pws, s);
    const uint32_t *pord = pws.order;
    AsyncBuf spos(n * 12, s), snrm(n * 12, s);
    hipLaunchKernelGGL(k_sdf_gather, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, pord, d_pos, d_nrm, spos.as<float>(),
                       snrm.as<float>());
    launch_check("k_sdf_gather");
    hipLaunchKernelGGL(k_sdf_rays, dim3(ceil_div(m, 256)), dim3(256), 0, s, (const float *)spos.as<float>(),
                       (const float *)snrm.as<float>(), n, scale, org.as<float>(), dir.as<float>());
    launch_check("k_sdf_rays");
    QueryScratch ws;
    launch_ray(*b, org.as<float>(), dir.as<float>(), m, -1.0f, false, ridx.as<int32_t>(), rt.as<float>(), rn.as<float>(),
               nullptr, ws, s, /*sort_rays=*/false);
    hipLaunchKernelGGL(k_sdf_valid, dim3(ceil_div(m, 256)), dim3(256), 0, s, m, ridx.as<int32_t>(), rt.as<float>(),
                       rn.as<float>(), dir.as<float>(), only_interior ? 1 : 0, val.as<float>());
    launch_check("k_sdf_valid");
    Stage st(s, mem == THG_MEM_HOST);
    double *d_sdf = st.out(sdf, (size_t)n);
    float *d_times = st.out(ray_times, (size_t)m);
    hipLaunchKernelGGL(k_sdf_reduce, dim3(ceil_div(n, SDF_RED_THREADS)), dim3(SDF_RED_THREADS), 0, s, n,
                       (const float *)val.as<float>(), pord, (double)scale, d_sdf);
    launch_check("k_sdf_reduce");
    if (d_times) {
      hipLaunchKernelGGL(k_sdf_times, dim3(ceil_div(m, 256)), dim3(256), 0, s, n, pord, (const float *)val.as<float>(),
                         d_times);
      launch_check("k_sdf_times");
    }
    st.finish();
  });
}

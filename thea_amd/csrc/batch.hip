// Many meshes in one call (thg_batch_closest): builds every mesh's BVH and answers each mesh's
// closest-point queries. Thea has no analogue: each BvhN is built and queried by its own caller.
#include <chrono>

#include "common.hpp"
#include "kernels_common.hpp"
#include "query.hpp"

using namespace thg;

extern "C" int thg_batch_closest(thg_ctx ctx, int32_t nm, const int64_t *nv, const int64_t *nt, const int64_t *nq,
                                 const float *verts, const int32_t *tris, const float *queries, int32_t leaf_size,
                                 int32_t *idx, float *dist, void *stream, double timings_ms[2]) {
  try {
    require(ctx != nullptr, "ctx is NULL");
    require(nm >= 0 && (nm == 0 || (nv && nt && nq)), "bad mesh counts");
    require(leaf_size >= 0 && leaf_size <= 16, "leaf_size must be in 0..16");
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    THG_HIP(hipSetDevice(c->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::vector<std::unique_ptr<Bvh>> bvhs(nm);
    int64_t vo = 0, to = 0;
    auto t0 = std::chrono::steady_clock::now();
    for (int32_t m = 0; m < nm; ++m) {
      require(nv[m] >= 0 && nt[m] >= 0 && nq[m] >= 0, "negative count");
      bvhs[m].reset(new Bvh());
      bvhs[m]->ctx = c;
      int leaf = leaf_size > 0 ? leaf_size : 4;
      build_bvh(*bvhs[m], verts + 3 * vo, tris + 3 * to, nt[m], false, leaf, s);
      vo += nv[m];
      to += nt[m];
    }
    THG_HIP(hipStreamSynchronize(s));
    auto t1 = std::chrono::steady_clock::now();
    int64_t qo = 0;
    for (int32_t m = 0; m < nm; ++m) {
      QueryScratch ws;
      if (bvhs[m]->n > 0) launch_closest(*bvhs[m], queries + 3 * qo, nq[m], -1.0, 0, idx + qo, nullptr, dist ? dist + qo : nullptr,
                                         nullptr, ws, s);
      qo += nq[m];
    }
    THG_HIP(hipStreamSynchronize(s));
    auto t2 = std::chrono::steady_clock::now();
    if (timings_ms) {
      timings_ms[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
      timings_ms[1] = std::chrono::duration<double, std::milli>(t2 - t1).count();
    }
    return THG_OK;
  } catch (const Status &st) {
    set_last_error(st.msg);
    return st.code;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return THG_E_DEVICE;
  }
}

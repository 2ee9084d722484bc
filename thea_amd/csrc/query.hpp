// Host-side interface of the query kernels (query.hip) and sampler (sample.hip).
#pragma once

#include "common.hpp"

namespace thg {

// Batches smaller than this are processed in input order (sorting would cost more than it saves).
constexpr int64_t ORDER_MIN_QUERIES = 4096;

// Child visiting orders of the compacted closest-point kernel (Bvh::child_order, SegDesc::child_order):
// results never depend on the order, only the work does.
constexpr int CHILD_MAJORITY = 0;  // the child most of the wave's lanes find nearer
constexpr int CHILD_LEADER = 1;    // the child the wave's first active query finds nearer
constexpr int CHILD_BUILD = 2;     // child 0 first

// One batch of a multi-batch launch (k_closest_multi / k_knn_multi / k_ray_multi): the BVH it
// queries, its inputs and outputs, and the first wave of the launch that works on it.
struct SegDesc {
  const void *nodes;
  const void *elems;
  const void *rays;      // RayTri records (ray segments)
  const void *qnodes;    // QNode4 tree (ray segments)
  const float *q;        // queries (closest, k-NN) or ray origins
  const float *q2;       // ray directions
  const uint32_t *order; // optional spatial order of the batch
  const void *screen;    // ScreenRec records (mesh closest-point segments)
  int32_t *idx;
  float *cp;             // closest points (closest) / normals (rays)
  float *dist;           // distances (closest, k-NN) / hit times (rays)
  float *d2;
  int64_t n;
  uint64_t wave_begin;
  float best_init;       // squared distance bound (closest, k-NN) / bound on t (rays)
  float scale;
  float max_time;
  int32_t exclude_self;
  int32_t k;
  int32_t points;
  int32_t root;          // node the traversal starts from
  int32_t nelem;         // elements of the segment's tree (k-NN seed window)
  int32_t lanes_log2;    // lanes per query (log2) of the grouped sparse kernels
  int32_t child_order;   // visiting order of the compacted closest-point kernel (0: majority)
};

// Per-call temporaries of a query batch (spatial order of the queries). By default the sort memory
// is allocated stream-ordered per call; with `ext` set it comes from caller-provided device memory
// (ext_cap queries' worth: 4 * ext_cap words of keys, values and their ping-pong copies, then
// radix_hist_words(ext_cap) histogram words) and nothing is allocated.
struct QueryScratch {
  AsyncBuf keys, vals, tk, tv, hist;
  const uint32_t *order = nullptr;
  uint32_t *ext = nullptr;
  int64_t ext_cap = 0;
  bool timed = true;  // record kernel timing events when the context has timing enabled
  struct Sort {
    uint32_t *keys, *vals, *tk, *tv, *hist;
  };
  Sort sort_memory(int64_t n, hipStream_t s);
  Ctx *timer_ctx(Ctx *c) const { return timed ? c : nullptr; }
};
// Device words a QueryScratch::ext region needs for batches of up to `cap` queries.
size_t sort_scratch_words(int64_t cap);

// Fails with THG_E_INVALID unless every vertex index of the nt triangles (device memory) is in
// [0, nv); synchronises `s`.
void check_tri_indices(const int32_t *tris, int64_t nt, int64_t nv, hipStream_t s);
void build_bvh(Bvh &b, const float *verts, const int32_t *tris, int64_t n, bool points, int leaf_size, hipStream_t s);
// The 4-wide quantized ray tree (Bvh::qnodes) of a built mesh BVH; synchronises `s`.
void build_qbvh4(Bvh &b, hipStream_t s);
void order_queries(const Bvh &b, const float *q, int64_t n, QueryScratch &ws, hipStream_t s);
void order_rays(const Bvh &b, const float *o, const float *d, int64_t n, QueryScratch &ws, hipStream_t s);
float bound_to_best(double dist_bound);
void launch_closest(const Bvh &b, const float *q, int64_t n, double dist_bound, int exclude_self, int32_t *idx,
                    float *cp, float *dist, float *d2, QueryScratch &ws, hipStream_t s);
// Build-pipeline pieces reused by the batched many-mesh plan (batch.hip).
// Leaf bounds of a mesh BVH's leaves into out[first record of each leaf]: nodes [0, nsmall + *kept)
// (kept null: none beyond nsmall), at most max_nodes.
void launch_screen_records(const TriRec *tr, int64_t n, ScreenRec *out, hipStream_t s);
void launch_leaf_bounds(const BNode *nodes, int64_t nsmall, const uint32_t *kept, int64_t max_nodes, const TriRec *tr,
                        LeafBound *out, hipStream_t s);
// Words of the refit's per-span entry counts (and level-list lengths) for n elements, and whether
// the refit of n elements needs its n-word level list.
size_t refit_count_words(int64_t n);
bool refit_needs_list(int64_t n);
void launch_refit(int64_t n, const uint32_t *sorted, const float4 *blo, const float4 *bhi, const int32_t *left,
                  const int32_t *right, const int32_t *rfirst, const int32_t *rlast, const int32_t *nparent,
                  const int32_t *lparent, uint32_t *flags, float *nbox, int32_t *entries, uint32_t *nentries,
                  int32_t *list, hipStream_t s);
void launch_emit(int64_t ni, int32_t base, const uint32_t *keep, const uint32_t *fidx, const int32_t *left,
                 const int32_t *right, const int32_t *rfirst, const int32_t *rlast, const uint32_t *sorted,
                 const float4 *blo, const float4 *bhi, const float *nbox, BNode *out, uint32_t *leaf_count,
                 hipStream_t s);
void launch_depth(int64_t n, const int32_t *lparent, const int32_t *nparent, const uint32_t *keep, const int32_t *rfirst,
                  int32_t *max_depth, hipStream_t s);
// Exclusive scan of n words: out must hold n + 1 words (the total lands in out[n]) and tmp
// scan_tmp_words(n) words.
size_t scan_tmp_words(int64_t n);
void scan_excl_u32(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *tmp, hipStream_t s);

// Launch many closest-point batches (segments) as one kernel; `seg_dev` is a device copy of the
// descriptor table with wave_begin filled in. Requires every segment's BVH to fit the wave stack.
// concurrent: the launch shares the device with other streams' launches (request server), which
// selects the variant that needs no scratch memory.
// block_waves: waves per workgroup (1, 2, 4, 8); spill_build: the 8-resident-wave build.
void launch_closest_multi(const SegDesc *seg_dev, int nseg, uint64_t total_waves, bool points, hipStream_t s,
                          bool concurrent = false, int block_waves = 4, bool spill_build = true);
void launch_knn_multi(const SegDesc *seg_dev, int nseg, uint64_t total_waves, int k, bool points, hipStream_t s);
// One query per lane over a segment table (stack >= the deepest tree of the segments, at most 64).
void launch_closest_lane_multi(const SegDesc *seg_dev, int nseg, uint64_t total_waves, bool points, int stack,
                               hipStream_t s);
void launch_knn_lane_multi(const SegDesc *seg_dev, int nseg, uint64_t total_waves, int k, bool points, int stack,
                           hipStream_t s);
// 2^lanes_log2 lanes per query over a segment table (SegDesc::lanes_log2 per segment, each
// segment spanning ceil(n 2^lanes_log2 / 64) waves); stack >= the deepest tree plus 1, at most 64.
void launch_closest_group_multi(const SegDesc *seg_dev, int nseg, uint64_t total_waves, bool points, int stack,
                                hipStream_t s);
void launch_knn_group_multi(const SegDesc *seg_dev, int nseg, uint64_t total_waves, int k, bool points, int stack,
                            hipStream_t s);
// log2 of the lanes per query the sparse kernels give a batch of n queries on a tree of this depth.
int group_lanes_log2(int64_t n, int depth);
bool use_wave_traversal(const Bvh &b, int64_t n);
// stack: the deepest tree among the segments plus one (32 or 64 entries). q_stack > 0 traverses
// the segments' QNode4 trees instead; it must be at least q4_stack() of every segment's BVH.
void launch_ray_multi(const SegDesc *seg_dev, int nseg, uint64_t total_waves, int stack, hipStream_t s,
                      int q_stack = 0);
// Stack entries a QNode4 traversal of b may need (three per level) and the most the kernels hold.
inline int q4_stack(const Bvh &b) { return 3 * b.q_depth + 1; }
constexpr int Q4_MAX_STACK = 16 + 96;
constexpr int SEG_WAVE_QUERIES = 64;  // queries per wave in the multi-batch kernels
void ray_stats(const Bvh &b, const float *o, const float *d, int64_t n, int32_t *idx, int32_t *counts, hipStream_t s);
void closest_wave_stats(const Bvh &b, const float *q, int64_t n, unsigned long long *stats, QueryScratch &ws,
                        hipStream_t s);
void launch_knn(const Bvh &b, const float *q, int64_t n, int k, double dist_bound, int exclude_self, int32_t *idx,
                float *dist, float *d2, QueryScratch &ws, hipStream_t s);
// Per-wave statistics of the 16-nearest wave kernel on a point BVH (4 int64 per 64-query wave).
void knn_wave_stats(const Bvh &b, const float *q, int64_t n, int exclude_self, int64_t *wstats, QueryScratch &ws,
                    hipStream_t s);
// sort_rays = false traces the rays in the order given by given_order (ray given_order[j] in slot j;
// nullptr: input order), for callers that make the rays coherent themselves. Results are always
// written at the rays' input indices.
void launch_ray(const Bvh &b, const float *o, const float *d, int64_t n, float max_time, bool any, int32_t *idx,
                float *t, float *nrm, uint8_t *hit, QueryScratch &ws, hipStream_t s, bool sort_rays = true,
                const uint32_t *given_order = nullptr);

// Sampler
void mesh_cdf(Mesh &m, hipStream_t s);
int64_t launch_sample_approx(Mesh &m, int64_t n, uint64_t seed, uint64_t stream_id, int64_t capacity, float *pos,
                             float *nrm, int32_t *tri, hipStream_t s);
float graph_build(Bvh *b, const float *d_nrm, int md, int64_t trials, uint64_t seed, hipStream_t s, int32_t *d_idx,
                  float *d_sep, int32_t *d_deg, const Bvh *surface = nullptr);
void launch_sample(Mesh &m, int64_t n, uint64_t seed, uint64_t stream_id, float *pos, float *nrm, int32_t *tri,
                   hipStream_t s);

}  // namespace thg

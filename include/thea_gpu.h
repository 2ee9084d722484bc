/*
 * thea_gpu.h -- C ABI of the MI355X BVH-query and mesh-sampling library (libthea_gpu.so).
 *
 * This is the drop-in boundary for Thea's BvhN / MeshBvh / MeshSampler query path. Thea exposes
 * that path as C++ templates (Algorithms/BvhN.hpp, Algorithms/MeshBvh.hpp,
 * Algorithms/MeshSampler.hpp); each entry point below names the Thea interface it replaces.
 * include/thea_gpu_bvh.hpp wraps these calls in classes that keep Thea's method names.
 *
 * Conventions
 *  - Every function returns THG_OK (0) or a negative THG_E* status; thg_last_error() returns a
 *    thread-local message for the last failure on the calling thread. Nothing calls exit().
 *  - "Not found" is encoded per query as index -1 and time/distance -1, as in Thea
 *    (BvhN.hpp:1062 closestElement returns -1; BvhN.hpp:1315 ray miss is time -1, index -1).
 *  - `mem` says where the caller's query/result pointers live: THG_MEM_DEVICE (device memory,
 *    the call is asynchronous on `stream`) or THG_MEM_HOST (host memory, the call stages
 *    through library-owned device buffers and returns when results are in host memory).
 *  - `stream` is a hipStream_t passed as void*; NULL is the null stream.
 *  - Handles are immutable after creation; queries on one handle may run concurrently on
 *    different streams. Destroying a handle while work on it is in flight is undefined.
 *  - Indices returned are the caller's triangle (or point) indices, in input order, as Thea's
 *    MeshBvh returns mesh triangle order (Algorithms/MeshTriangles.hpp:328-339, BvhN.hpp:720-725).
 *  - Exact ties (bit-identical float distance or hit time) resolve to the LOWEST index.
 */
#ifndef THEA_GPU_H
#define THEA_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define THG_OK 0
#define THG_E_INVALID (-1)     /* bad argument */
#define THG_E_OOM (-2)         /* device allocation failed */
#define THG_E_DEVICE (-3)      /* HIP runtime error */
#define THG_E_BUDGET (-4)      /* request exceeds a memory budget */
#define THG_E_UNSUPPORTED (-5) /* valid request this build does not support */
#define THG_E_CANCELLED (-6)   /* request cancelled or timed out (serving loop) */

#define THG_MEM_DEVICE 0
#define THG_MEM_HOST 1

typedef struct thg_ctx_s *thg_ctx;
typedef struct thg_mesh_s *thg_mesh;
typedef struct thg_bvh_s *thg_bvh;

/* BVH build options. Replaces the `init(method, max_depth, max_elems_in_leaf, save_memory, ...)`
 * arguments of BvhN::init (Algorithms/BvhN.hpp:705-706) and MeshBvh::init
 * (Algorithms/MeshBvh.hpp:104-111). The tree is a GPU LBVH, not Thea's kd-tree; results are
 * defined by the element arithmetic and the lowest-index tie-break, not by tree shape. */
typedef struct thg_build_opts {
  int32_t leaf_size; /* max elements per leaf, 1..16; 0 = default (4) */
  int32_t reserved[7];
} thg_build_opts;

/* Library status text for the calling thread's last failed call (never NULL). */
const char *thg_last_error(void);

/* Version string of the library. */
const char *thg_version(void);

/* Context on one HIP device (the device index as HIP numbers it). One per process per device is
 * typical. Thea has no analogue (its queries run on the calling CPU thread). */
int thg_ctx_create(int32_t device, thg_ctx *out);
int thg_ctx_destroy(thg_ctx ctx);

/* Upload a triangle mesh: `verts` = nv float3 (x,y,z interleaved), `tris` = nt int32 triplets of
 * vertex indices. Replaces MeshBvh::add / addTriangles and MeshTriangles::add
 * (Algorithms/MeshBvh.hpp:55-67, Algorithms/MeshTriangles.hpp:348-422). Vertices and indices are
 * copied; the caller's arrays may be freed afterwards. */
int thg_mesh_create(thg_ctx ctx, const float *verts, int64_t nv, const int32_t *tris, int64_t nt, int32_t mem,
                    thg_mesh *out);
int thg_mesh_destroy(thg_mesh mesh);
int64_t thg_mesh_num_triangles(thg_mesh mesh);
/* Total surface area (double sum of per-triangle float areas, MeshSampler.hpp:147-149). */
int thg_mesh_total_area(thg_mesh mesh, double *area);

/* Build a BVH over the triangles of `mesh` on `stream` (synchronous). Replaces BvhN::init /
 * initTree / createKdTree (Algorithms/BvhN.hpp:705-794, :1366-1525) and MeshBvh::init
 * (Algorithms/MeshBvh.hpp:104-111). The BVH keeps a reference to the mesh; destroy the BVH first. */
int thg_bvh_build_mesh(thg_mesh mesh, const thg_build_opts *opts, void *stream, thg_bvh *out);

/* Build a BVH over n points (float3). Replaces BvhN<Vector3, 3>::init over points (the point BVH
 * of TestBvh.cpp:159-163, PointSet3.hpp:71 SampleBvh). Points are copied. */
int thg_bvh_build_points(thg_ctx ctx, const float *xyz, int64_t n, int32_t mem, const thg_build_opts *opts,
                         void *stream, thg_bvh *out);
int thg_bvh_destroy(thg_bvh bvh);
/* Statistics of a built BVH: out[0] nodes, [1] leaves, [2] max depth, [3] elements,
 * [4] device bytes held by the BVH, [5] leaf size. */
int thg_bvh_info(thg_bvh bvh, int64_t out[6]);
/* The union of the element boxes: exact min/max of the coordinates, without padding (an empty BVH
 * gives zeros). Thea's BvhN::getBounds (BvhN.hpp:991) returns this box scaled by 1.05 about its
 * centre (BOUNDS_EXPANSION_FACTOR, BvhN.hpp:1437,2458). */
int thg_bvh_bounds(thg_bvh bvh, float lo[3], float hi[3]);
/* Copy a built BVH to host memory (diagnostics and tests; synchronous). nodes (nullable):
 * info[0] 64-byte nodes, each {child 0 box lo xyz, hi xyz, child 1 box lo xyz, hi xyz (floats),
 * ref0, ref1, count0, count1 (int32)}; ref >= 0 is a node index, ref < 0 a leaf holding elements
 * [~ref, ~ref + count) of the leaf order. elem_ids (nullable): info[3] caller indices in leaf order.
 * Node 0 is the root. */
int thg_bvh_export(thg_bvh bvh, void *nodes, int32_t *elem_ids);
/* Copy a mesh BVH's ray tree to host memory (diagnostics and tests; synchronous; no counterpart in
 * Thea). *num_qnodes receives the node count; qnodes (nullable) receives that many 64-byte nodes,
 * each {origin xyz (float), meta (uint32: biased exponents ex | ey << 8 | ez << 16 | children << 24),
 * qlo_x, qlo_y, qlo_z, qhi_x, qhi_y, qhi_z (uint32, byte k for child k), ref0..ref3 (int32, >= 0 a
 * node of this tree, < 0 a leaf ~first in leaf order), counts (uint16 x 4)}. Child k's box decodes
 * as fmaf(q, 2^e, origin) per axis and contains the binary tree's box for it. Node 0 is the root;
 * point BVHs have none (*num_qnodes = 0). */
int thg_bvh_export_ray_tree(thg_bvh bvh, void *qnodes, int64_t *num_qnodes);

/* Closest element and closest point for n query points `q` (float3).
 * Replaces BvhN::closestElement<MetricL2>(query, dist_bound, compat, &dist, &closest_point)
 * (Algorithms/BvhN.hpp:1050-1063) and BvhN::closestPair (:1082-1114).
 *  dist_bound: only elements at distance < dist_bound are considered; negative = unbounded.
 *  idx[i]: element index or -1. cp (nullable): closest point on that element (float3).
 *  dist (nullable): distance sqrt(d^2) as float (Thea returns sqrt((double)d^2), MetricL2.hpp:97;
 *  the two agree after rounding to float). d2 (nullable): the exact float squared distance. */
int thg_closest_point(thg_bvh bvh, const float *q, int64_t n, double dist_bound, int32_t *idx, float *cp, float *dist,
                      float *d2, int32_t mem, void *stream);

/* Point BVH over double coordinates, for callers of BvhN<Vector3d, 3, double> over points (the
 * target structure of ShapeAlign's ICP, Tools/ShapeAlign/ShapeAlign.cpp:339; built by BvhN::init,
 * Algorithms/BvhN.hpp:705-794). `xyz` = n double3. Queries on it go through thg_closest_point_f64
 * (thg_closest_point and thg_k_closest also accept it and then search the points rounded to float). */
int thg_bvh_build_points_f64(thg_ctx ctx, const double *xyz, int64_t n, int32_t mem, const thg_build_opts *opts,
                             void *stream, thg_bvh *out);

/* Nearest point in double precision on a thg_bvh_build_points_f64 BVH. Replaces
 * BvhN<Vector3d, 3, double>::closestElement<MetricL2>(query, dist_bound, compat, &dist, &closest_point)
 * (Algorithms/BvhN.hpp:1050-1063; point metric MetricL2.hpp:176-193) as Icp3<double> calls it
 * (Algorithms/Icp3.hpp:282-296,388-404). q: n double3. The answer is the lowest-index point of least
 * double squared distance ((dx^2 + dy^2) + dz^2); only points with distance < dist_bound count
 * (negative = unbounded). idx[i]: point index or -1. cp (nullable): that point's double
 * coordinates. dist (nullable): sqrt(d^2) in double. d2 (nullable): the double squared distance. */
int thg_closest_point_f64(thg_bvh bvh, const double *q, int64_t n, double dist_bound, int32_t *idx, double *cp,
                          double *dist, double *d2, int32_t mem, void *stream);

/* Largest batch thg_closest_point_noalloc accepts. */
#define THG_NOALLOC_MAX_QUERIES 4194304

/* thg_closest_point that allocates nothing: no device or host allocation, no HIP events, no
 * memset. No counterpart in Thea; it is for callers that cannot allocate: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc. Same arguments and results as
 * thg_closest_point with mem = THG_MEM_DEVICE (all pointers device memory, asynchronous on
 * `stream`). The spatial sort of the queries uses a fixed per-device array of the library
 * (68.2 MB, held until the process ends), so n is limited to THG_NOALLOC_MAX_QUERIES: larger n
 * fails with THG_E_INVALID (split the batch). Every call on one device shares that array: calls
 * that overlap on one device (other streams, other threads) race and must be serialised by the
 * caller; calls in order on one stream are safe. A context must have been created on the device. */
int thg_closest_point_noalloc(thg_bvh bvh, const float *q, int64_t n, double dist_bound, int32_t *idx, float *cp,
                              float *dist, float *d2, void *stream);

/* k nearest elements per query, sorted by (distance, index) ascending; rows shorter than k are
 * padded with index -1 and distance -1. Replaces BvhN::kClosestPairs<MetricL2>
 * (Algorithms/BvhN.hpp:1139-1146) with a BoundedSortedArray of capacity k.
 *  k: 1..32. exclude_self: skip element i for query i (queries are the BVH's own points; the
 *  self-exclusion PointSet3 applies with a filter, Algorithms/PointSet3.hpp:576-583).
 *  idx: n*k int32. dist (nullable): n*k float distances. d2 (nullable): n*k float squared. */
int thg_k_closest(thg_bvh bvh, const float *q, int64_t n, int32_t k, double dist_bound, int32_t exclude_self,
                  int32_t *idx, float *dist, float *d2, int32_t mem, void *stream);

/* First hit of n rays (origins, directions: float3; directions need not be unit length).
 * Replaces BvhN::rayStructureIntersection<RayIntersectionTester> (Algorithms/BvhN.hpp:1291-1316),
 * rayIntersectionTime (:1274-1289) and rayIntersects (:1266-1269).
 *  max_time: only hits with t < max_time count (Thea's structure traversal accepts a hit only if it
 *  is strictly earlier than the running bound, which starts at max_time, BvhN.hpp:2247-2258);
 *  negative = unbounded.
 *  idx[i]: triangle index or -1. t[i] (nullable): hit time or -1. normal (nullable): face normal
 *  of the hit triangle (float3, Triangle3.hpp:542-552) or 0 on a miss. */
int thg_ray_first_hit(thg_bvh bvh, const float *origins, const float *dirs, int64_t n, float max_time, int32_t *idx,
                      float *t, float *normal, int32_t mem, void *stream);

/* Any hit: hit[i] = 1 if some triangle is hit with 0 <= t (< max_time when max_time >= 0),
 * else 0. Replaces BvhN::rayIntersects (Algorithms/BvhN.hpp:1266-1269). */
int thg_ray_any_hit(thg_bvh bvh, const float *origins, const float *dirs, int64_t n, float max_time, uint8_t *hit,
                    int32_t mem, void *stream);

/* Area-weighted random samples on the mesh surface, exact count. Replaces
 * MeshSampler::sampleEvenlyByArea(n, positions, face_normals, triangles, CountMode::EXACT)
 * (Algorithms/MeshSampler.hpp:100-204) and TriangleN::randomPoint (Triangle3.hpp:179-195).
 * Randomness: sample i uses Philox4x32-10 with counter (i, stream_id) and key `seed`; its first
 * three 32-bit words drive the triangle pick and the two barycentric draws, each mapped to [0,1]
 * as Thea's Random::uniform01 does (Random.hpp:126-131). The same stream is reproducible on the
 * host. `n_out` receives the number of samples written (0 for an empty or zero-area mesh).
 *  pos: n float3. normal (nullable): n float3 face normals. tri (nullable): n int32. */
int thg_sample_area(thg_mesh mesh, int64_t n, uint64_t seed, uint64_t stream_id, float *pos, float *normal,
                    int32_t *tri, int64_t *n_out, int32_t mem, void *stream);
/* sampleEvenlyByArea in CountMode::APPROXIMATE (Algorithms/MeshSampler.hpp:206-240): triangle i
 * gets floor(frac) samples, frac = (n * area_i) / total_area, plus one more when a uniform draw is
 * at most frac's fractional part; samples come in triangle order. Thea draws from its time-seeded
 * generator; here triangle i's draw is word 0 of Philox4x32-10 at counter (i, stream_id) with key
 * (seed lo ^ 0x41505052, seed hi), and sample j's barycentric uniforms are words 1 and 2 at counter
 * (j, stream_id) with key seed. The count is random: with pos NULL only *n_out is set; otherwise
 * the call fails with THG_E_INVALID if the count exceeds `capacity` (rows of pos/normal/tri). */
int thg_sample_area_approx(thg_mesh mesh, int64_t n, uint64_t seed, uint64_t stream_id, int64_t capacity, float *pos,
                           float *normal, int32_t *tri, int64_t *n_out, int32_t mem, void *stream);

/* Kernel timing (diagnostics; Thea's analogue is Stopwatch, Stopwatch.hpp:41-76). While enabled on
 * a context, each query or sampling call records HIP events around its main kernel on the call's
 * stream. thg_ctx_kernel_time synchronises those events and returns the summed kernel time (ms) and
 * launch count for `kernel` ("closest", "knn", "ray", "any_hit", "sample", "order") since the last
 * reset; reset != 0 clears the record. */
int thg_ctx_set_timing(thg_ctx ctx, int32_t enable);
int thg_ctx_kernel_time(thg_ctx ctx, const char *kernel, double *total_ms, int64_t *launches, int32_t reset);

/* Traversal statistics of the closest-point kernel (diagnostics): runs the wave-cooperative
 * traversal over n queries (device pointer) and returns sums over waves of out[0] waves,
 * [1] queries, [2] nodes loaded, [3] child-box tests, [4] elements screened (those of leaves that
 * some lane keeps after the leaf-level reject), [5] lane slots spent on exact element evaluations
 * (64 per element the whole wave evaluates, or 64 per flush of the compacted queue), [6] (lane,
 * element) pairs that passed the per-lane element rejects,
 * [7] whole-wave exact evaluations whose result came within 2^-8 of some lane's best (0 for the
 * compacted kernel). */
int thg_closest_stats(thg_bvh bvh, const float *q, int64_t n, void *stream, int64_t out[8]);

/* Per-wave statistics of the 16-nearest-neighbour wave kernel on a point BVH (diagnostics; no
 * counterpart in Thea): n queries (device pointer), 64 per wave in spatial order; per_wave (device,
 * 4 int64 per wave) = elapsed wall clock in 10 ns ticks, nodes entered, points tested, insertion
 * rounds. Asynchronous on `stream`. */
int thg_knn_wave_stats(thg_bvh bvh, const float *q, int64_t n, int32_t exclude_self, void *stream,
                       int64_t *per_wave);

/* Per-ray traversal counts of the first-hit kernel (diagnostics), rays in input order (device
 * pointers, asynchronous on `stream`): idx[i] as thg_ray_first_hit, counts[4i..4i+3] = nodes
 * visited, triangles tested, deepest traversal stack, stack pops. */
int thg_ray_stats(thg_bvh bvh, const float *origins, const float *dirs, int64_t n, void *stream, int32_t *idx,
                  int32_t *counts);

/* Many meshes, one call: build a BVH for each of `nm` meshes and answer `nq[m]` closest-point
 * queries for mesh m (Thea has no analogue: each BvhN is built and queried by its own caller).
 * Inputs are concatenated: verts (sum nv float3), tris (sum nt int32 triplets, vertex indices local
 * to their mesh), queries (sum nq float3, grouped by mesh); nv/nt/nq are per-mesh counts (host
 * arrays). Results use local triangle indices. All array pointers are device pointers.
 *
 * A plan sizes every buffer once; its whole pipeline (all meshes' bounds, one Morton sort over all
 * meshes, per-mesh radix trees, one refit, one compaction, records, query sort, and one
 * multi-segment closest-point launch) has no host synchronisation, so the first run captures it into
 * a HIP graph and later runs replay it. leaf_sizes: per mesh, 1..16 (NULL = 4).
 * timings_ms (nullable): when given, the run is eager (not graph) and reports [build, query] ms. */
typedef struct thg_batch_plan_s *thg_batch_plan;
int thg_batch_plan_create(thg_ctx ctx, int32_t nm, const int64_t *nv, const int64_t *nt, const int64_t *nq,
                          const int32_t *leaf_sizes, thg_batch_plan *out);
int thg_batch_plan_run(thg_batch_plan plan, const float *verts, const int32_t *tris, const float *queries,
                       int32_t *idx, float *dist, void *stream, double timings_ms[2]);
int thg_batch_plan_destroy(thg_batch_plan plan);
/* Leaf size per mesh chosen by measurement: for each size class (floor(log2(triangles))) the
 * plan's closest-point kernel is timed on up to two of its meshes for leaf sizes 2..16 and the
 * fastest is kept (cached for the process). Same inputs as thg_batch_closest; writes leaf_sizes[nm]. */
int thg_autotune_leaf(thg_ctx ctx, int32_t nm, const int64_t *nv, const int64_t *nt, const int64_t *nq,
                      const float *verts, const int32_t *tris, const float *queries, int32_t *leaf_sizes);
/* Leaf size and child visiting order per mesh, and the launch shape of the plan, chosen by
 * measurement. Per size class, as thg_autotune_leaf, then at that leaf size the child order
 * (child_orders[m]: 0 = the child most of a wave's queries find nearer, 1 = the child with the least
 * lower bound over the wave's queries, 2 = the build's order; cached per class). Then launch[0] =
 * waves per workgroup of the query launch (1, 2, 4, 8) and launch[1] = 1 for the build compiled for 8
 * resident waves per SIMD (0: the spill-free one), the fastest summed over the sampled meshes.
 * Results never depend on these choices, only the time does. No counterpart in Thea. */
int thg_autotune_plan(thg_ctx ctx, int32_t nm, const int64_t *nv, const int64_t *nt, const int64_t *nq,
                      const float *verts, const int32_t *tris, const float *queries, int32_t *leaf_sizes,
                      int32_t *child_orders, int32_t launch[2]);
/* Child visiting orders (nullable: 0 for every mesh) and launch shape of a plan's query launch, as
 * thg_autotune_plan chooses them; the next run captures its graph again. */
int thg_batch_plan_set_launch(thg_batch_plan plan, const int32_t *child_orders, int32_t block_waves,
                              int32_t spill_build);
/* One-shot form: autotunes leaf sizes, child orders and launch shape when leaf_size == 0, creates a
 * plan, runs it, frees it. */
int thg_batch_closest(thg_ctx ctx, int32_t nm, const int64_t *nv, const int64_t *nt, const int64_t *nq,
                      const float *verts, const int32_t *tris, const float *queries, int32_t leaf_size,
                      int32_t *idx, float *dist, void *stream, double timings_ms[2]);

/* ---------------------------------------------------------------------------------------------
 * Iterative closest point. Replaces Icp3<double>::align(n, from, to_structure, &error)
 * (Algorithms/Icp3.hpp:112-117, :193-263) with the BVH as the proximity structure, and runs
 * num_problems alignments of the same source points from different initial transforms at once
 * (ShapeAlign's search over starting rotations, Tools/ShapeAlign/ShapeAlign.cpp:343-414).
 * --------------------------------------------------------------------------------------------- */
typedef struct thg_icp_opts {
  double fractional_error_threshold; /* convergence threshold on the fractional error change; negative: 0.0001 */
  int32_t min_iterations;            /* negative: 3 */
  int32_t max_iterations;            /* negative: 10 */
  int32_t has_up;                    /* nonzero: rotations about `up` only (Icp3::setUpVector) */
  int32_t kabsch;                    /* 0: Thea's fit, rot = U V^T of cov = sum dp dq^T (Icp3.hpp:349-350),
                                        which is the transpose of the least-squares rotation;
                                        nonzero: the least-squares rotation V U^T */
  double up[3];
  const double *translation_weights; /* nullable: n per-point weights of the means (getTranslationWeight) */
  const double *rotation_weights;    /* nullable: n per-point weights of the fit (getRotationWeight) */
  int64_t reserved[4];
} thg_icp_opts;

/* from: n double3 source points; initial (nullable): num_problems row-major 3x4 transforms applied to
 * `from` first (NULL: identity, one problem per call makes sense then); from, initial and the weights
 * live in `mem` memory. Writes, in host memory, transforms[12 k ...] = the row-major 3x4 transform
 * Icp3 returns for problem k (to be composed after initial[k], as ShapeAlign does), errors[k]
 * (nullable) = its alignment error, iterations[k] (nullable) = the ICP steps taken. The nearest-point
 * searches use the BVH's float arithmetic on the positions rounded to float; fits and errors are
 * double. Fails with THG_E_INVALID if the BVH is empty (Thea throws, Icp3.hpp:249-250). */
int thg_icp_align(thg_bvh to, const double *from, int64_t n, int32_t num_problems, const double *initial,
                  const thg_icp_opts *opts, int32_t mem, void *stream, double *transforms, double *errors,
                  int32_t *iterations);

/* ---------------------------------------------------------------------------------------------
 * Sample proximity graph. Replaces PointSet3::updateGraph() without a surface, with or without
 * oversampling (Algorithms/PointSet3.hpp:324-386, :586-697; PointSet3.cpp:158-178), the graph
 * Tools/SampleGraph builds (Tools/SampleGraph/SampleGraph.cpp:296-334).
 * --------------------------------------------------------------------------------------------- */
typedef struct thg_graph_opts {
  int32_t max_degree;        /* PointSet3::Options::max_degree, 1..32; <= 0: 8 */
  int32_t separation_trials; /* samples of the initial separation estimate; <= 0: 100 */
  uint64_t seed;             /* picks of those samples: pick t = (w * n) >> 32 for the first word w of
                                Philox4x32-10 at counter (t, 0, 0x47524150, 0), key = seed (Thea picks
                                with its time-seeded Random::common()) */
  int64_t num_original;      /* thg_proximity_graph: the first num_original points are the samples and
                                the rest their oversampling (PointSet3::addOversampling); <= 0: all
                                points are samples. Ignored elsewhere. */
  thg_bvh surface;           /* thg_proximity_graph (nullable): a mesh BVH for the reachability test of
                                PointSet3::updateGraph(surface) (PointSet3.hpp:633-641; SampleGraph -r);
                                needs normals. Ignored elsewhere. */
  int64_t reserved[2];
} thg_graph_opts;

/* points: a point BVH over the samples (thg_bvh_build_points); normals (nullable): n float3 in
 * `mem`. With S = n samples, writes, in `mem` memory, nbr_idx[S * max_degree] (neighbours sorted by
 * separation, then index; -1 padded), nbr_sep (nullable, same layout; separations are Thea's
 * fastSqrt of the float squared distance, -1 padded), degree (nullable, S). With oversampling
 * (opts->num_original = S < n, the SampleGraph tool's case below 50,000 samples) the graph is built
 * over all n points and each sample's neighbours are the max_degree nearest other samples by
 * shortest-path distance through it (PointSet3::extractOriginalAdjacencies, PointSet3.cpp:158-178):
 * S rows, separations = float(path length); exact ties in path length are "parity unpinned" (Thea's
 * Fibonacci heap orders them). separation (nullable, host, 2 floats): the initial RMS separation
 * estimate (over all n points) and the final average edge length of the S samples
 * (getAverageSeparation). */
int thg_proximity_graph(thg_bvh points, const float *normals, const thg_graph_opts *opts, int32_t mem, void *stream,
                        int32_t *nbr_idx, float *nbr_sep, int32_t *degree, float separation[2]);

/* ---- Mesh and point-sample files (host only; no device involved) --------------------------------
 * A mesh file read into one flat triangle list: Thea's OFF (ASCII or "OFF BINARY") and OBJ readers
 * (Graphics/MeshCodecOff.hpp:146-430, Graphics/MeshCodecObj.hpp:210-530, with MeshSample's options:
 * no normals or texture coordinates, non-strict), then Thea's triangulation of each face
 * (Algorithms/MeshTriangles.hpp:173-228: triangles as they are, quads split by
 * Polygon3::triangulateQuad, larger polygons by Polygon3::triangulate's ear clipping). Triangles
 * come in face order; triangle_faces holds the index Thea stores for each triangle's face (its
 * position in the file). Arrays are allocated by the library: release them with
 * thg_mesh_file_free. */
typedef struct thg_mesh_file {
  int64_t num_vertices;
  int64_t num_faces;       /* faces kept (>= 3 vertices) */
  int64_t num_triangles;
  float *vertices;         /* 3 * num_vertices, file order */
  int32_t *triangles;      /* 3 * num_triangles, indices into vertices */
  int64_t *triangle_faces; /* num_triangles */
} thg_mesh_file;
/* Reads `path` by its extension (.off or .obj, any case). Replaces MeshGroup::load
 * (Graphics/MeshGroup.hpp:378-415) followed by MeshTriangles::add. */
int thg_mesh_file_read(const char *path, thg_mesh_file *out);
void thg_mesh_file_free(thg_mesh_file *mesh);
/* MeshSample's point file (Tools/MeshSample/MeshSample.cpp:560-601): one line per sample,
 * "x y z nx ny nz [face_id]" in the C++ stream's default float format, values below 1e-35 in
 * magnitude written as 0. nrm and ids nullable. */
int thg_pts_write(const char *path, int64_t n, const float *pos, const float *nrm, const int64_t *ids);
/* The first three numbers of every line that is not empty or a '#' comment (MeshSample.cpp:
 * 314-348); *pos is allocated by the library (3 * *n floats): release it with thg_free. */
int thg_pts_read(const char *path, int64_t *n, float **pos);
void thg_free(void *p);

/* External visibility at n surface points: Visibility::compute(position, num_rays)
 * (Algorithms/SurfaceFeatures/Local/Visibility.hpp:99-125): the fraction of num_rays (<= 0: 100)
 * rays, uniform on the sphere by Knop's method and offset by 0.001 scale along themselves, that hit
 * nothing. Thea draws the directions from its time-seeded Random::common(); here ray r of point i
 * uses the uniforms of Philox4x32-10 at counter (i * num_rays + r, 0x56495342, attempt) with key
 * `seed`. scale <= 0 takes the diagonal of the BVH's bounds grown 1.05x, as Thea's constructors.
 * positions and visibility (n doubles) in `mem`. */
int thg_visibility(thg_bvh mesh_bvh, const float *positions, int64_t n, int32_t num_rays, uint64_t seed, float scale,
                   int32_t mem, void *stream, double *visibility);

/* FurthestPointSampling::subsample with geodesic distances (Algorithms/FurthestPointSampling.cpp:
 * 47-125): the proximity graph of the n points (as thg_proximity_graph without normals; graph_opts
 * nullable), then k picks starting with point 0, each the point furthest along the graph from the
 * picks so far (unreachable points first, the highest index among them; at equal maximum distance
 * the highest index). points in `mem`, selected (k int64) in `mem`; *count (host) = the number of
 * picks made, fewer than k when every remaining point is at distance 0 (Thea reports an error and
 * returns that count). Requires k <= n < 2^30. */
int thg_furthest_point_sampling(thg_ctx ctx, const float *points, int64_t n, int64_t k,
                                const thg_graph_opts *graph_opts, int32_t mem, void *stream, int64_t *selected,
                                int64_t *count);

/* MeshSampler::sampleEvenlyBySeparation (Algorithms/MeshSampler.hpp:270-329): ceil(factor * n)
 * area samples (thg_sample_area with seed/stream_id; factor < 0 selects Thea's 3), then n of them by
 * thg_furthest_point_sampling. pos (3n), normal (nullable, 3n), tri (nullable, n) in `mem`;
 * *n_out = n, or 0 when the mesh has no area or fewer than n picks were possible (Thea returns 0). */
int thg_sample_separation(thg_mesh mesh, int64_t n, float oversampling_factor, uint64_t seed, uint64_t stream_id,
                          const thg_graph_opts *graph_opts, float *pos, float *normal, int32_t *tri, int64_t *n_out,
                          int32_t mem, void *stream);

/* Shape diameter function at n surface points: ShapeDiameter::compute(position, normal,
 * only_hit_interior_surfaces) (Algorithms/SurfaceFeatures/Local/ShapeDiameter.hpp:185-256), or
 * compute(position) with the normal of the nearest face when `normals` is NULL (:162-176). scale:
 * the normalisation length (> 0; Thea defaults it to the bounding-sphere diameter of the mesh).
 * positions, normals, sdf and ray_times all in `mem`. sdf (n doubles): the values in [0, 1] or -1
 * where no ray counts; ray_times (nullable, 30 n floats): each cone ray's hit time, -1 if it does
 * not count. With THG_MEM_DEVICE the call is asynchronous on `stream`. */
int thg_shape_diameter(thg_bvh mesh_bvh, const float *positions, const float *normals, int64_t n, float scale,
                       int32_t only_interior, int32_t mem, void *stream, double *sdf, float *ray_times);

/* ---------------------------------------------------------------------------------------------
 * Request server (Thea has no analogue: its queries run synchronously on the caller's thread).
 * Keeps meshes and their BVHs resident on one GPU under a memory budget and serves closest-point,
 * k-NN and ray requests with continuous batching: whenever one of `num_streams` stream workers is
 * free it takes the queued requests that the memory budget admits (FIFO, up to the batch limits and
 * to its share of the queue: the queued points over `num_streams`, at least 256k points, so a deep
 * queue is spread over the streams) and runs them as one multi-batch launch per query type (each request a segment against its own
 * mesh), so requests that arrive while others run are admitted into the next launch and overlap
 * on the device. Results are bit-identical to the single-call entry points. All request arrays
 * are device pointers owned by the caller and must stay valid until the request completes.
 * ------------------------------------------------------------------------------------------- */
typedef struct thg_server_s *thg_server;

#define THG_REQ_CLOSEST 0
#define THG_REQ_KNN 1
#define THG_REQ_RAY 2

#define THG_PENDING 1 /* thg_server_wait: the request has not completed within the timeout */

typedef struct thg_server_opts {
  int64_t memory_budget_bytes; /* device bytes for resident meshes/BVHs + request working memory;
                                  0 = 90% of the device memory free at creation */
  int32_t num_streams;         /* concurrent batches (stream workers); 0 = 4 */
  int32_t max_batch_requests;  /* requests per launch; 0 = 256 */
  int64_t max_batch_queries;   /* queries per launch; 0 = 16M */
  int32_t coalesce;            /* 1 (default) merge admitted requests into one launch per type;
                                  -1 = one launch per request (still on concurrent streams) */
  int32_t reserved[7];
} thg_server_opts;

typedef struct thg_request {
  int32_t type;         /* THG_REQ_CLOSEST | THG_REQ_KNN | THG_REQ_RAY */
  int32_t mesh;         /* id from thg_server_add_mesh */
  int64_t n;            /* number of queries / rays */
  int32_t k;            /* k-NN: 1..32 */
  int32_t exclude_self; /* k-NN: as thg_k_closest */
  double bound;         /* dist_bound (closest, k-NN) or max_time (rays); negative = none */
  const float *q;       /* n float3: query points or ray origins */
  const float *dirs;    /* rays: n float3 directions */
  int32_t *idx;         /* n (closest, rays) or n*k (k-NN) element indices */
  float *dist;          /* nullable: distances (closest, k-NN) or hit times (rays) */
  float *cp;            /* nullable: closest points (closest) or hit normals (rays), n float3 */
} thg_request;

int thg_server_create(thg_ctx ctx, const thg_server_opts *opts, thg_server *out);
/* Waits for running batches, fails queued requests with THG_E_CANCELLED, frees everything. */
int thg_server_destroy(thg_server srv);
/* Upload a mesh and build its BVH; fails with THG_E_BUDGET if it does not fit the budget. */
int thg_server_add_mesh(thg_server srv, const float *verts, int64_t nv, const int32_t *tris, int64_t nt, int32_t mem,
                        int32_t leaf_size, int32_t *mesh_id);
/* Queue a request (returns at once). THG_E_BUDGET if it can never fit the budget. */
int thg_server_submit(thg_server srv, const thg_request *req, int64_t *req_id);
/* Wait up to timeout_us (negative = forever). Returns the request's status (THG_OK, an error
 * code, THG_E_CANCELLED) once it has finished, or THG_PENDING. latency_us (nullable) receives the
 * time from submit to completion. A finished request is forgotten after a successful wait. */
int thg_server_wait(thg_server srv, int64_t req_id, int64_t timeout_us, double *latency_us);
/* Cancel a request that has not started; THG_E_INVALID if it is running or unknown. */
int thg_server_cancel(thg_server srv, int64_t req_id);
/* out[0] requests completed, [1] launches, [2] mean requests per launch, [3] p50 latency (us),
 * [4] p99 latency (us), [5] resident bytes, [6] peak in-flight working bytes, [7] rejected requests,
 * [8] queries completed, [9] mean latency (us). reset != 0 clears the counters. */
int thg_server_stats(thg_server srv, double out[10], int32_t reset);

#ifdef __cplusplus
}
#endif

#endif /* THEA_GPU_H */

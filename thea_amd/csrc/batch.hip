// Many meshes of uneven size, built and queried together (thg_batch_plan_*, thg_batch_closest,
// thg_autotune_leaf in include/thea_gpu.h). Thea has no analogue: each BvhN is built and queried by
// its own caller.
//
// A plan owns device copies of the concatenated inputs and every intermediate array, sized once.
// Its pipeline runs on one stream with no host synchronisation, so it is captured into a HIP graph
// on the first run and replayed afterwards (inputs are copied into the plan's buffers first):
//   per-mesh bounds (one workgroup per chunk of one mesh, block reduction, ordered-int atomics)
//   -> keys (mesh id << b) | Morton code (32 - b bits) and one stable radix sort over all meshes,
//      which groups the elements by mesh in input order
//   -> Karras radix trees, each over its mesh's key range (global positions and node ids)
//   -> one bottom-up refit, per-mesh leaf sizes, one compaction, one emit, records with local ids
//   -> queries sorted by (mesh, Morton) the same way
//   -> one multi-segment closest-point launch: segment m walks mesh m's tree from its root.
// Small meshes (no more elements than their leaf size) get a single-leaf root node each.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <mutex>

#include "common.hpp"
#include "kernels_common.hpp"
#include "query.hpp"
#include "tri_math.hpp"

namespace thg {

constexpr int BB = 256;
constexpr int CHUNK = 4096;

struct MeshInfo {  // per mesh, device side
  int64_t voff, toff, qoff, ioff;
  int64_t nt, nq, nv;
  int32_t leaf, small_slot;  // small_slot >= 0: index of its single-leaf root node
  uint64_t wave_begin;       // first 64-query wave of the mesh within its traversal group
  int32_t seg_slot;          // row of the segment table (wave-cooperative meshes first)
  int32_t lane;              // > 0: log2 lanes per query (sparse queries), 0: wave-cooperative
  int32_t child_order;       // visiting order of the closest-point kernel on this mesh (CHILD_*)
};

struct BatchPlan {
  Ctx *ctx = nullptr;
  int32_t nm = 0;
  int mesh_bits = 0;
  std::vector<int64_t> nv, nt, nq;
  std::vector<MeshInfo> info;
  int64_t NV = 0, NT = 0, NQ = 0, NI = 0, KMAX = 0;
  int32_t nsmall = 0;
  uint64_t total_waves = 0, lane_waves = 0;
  int32_t nwave_meshes = 0;
  std::vector<int4> tchunks, qchunks;
  DevBuf verts, tris, queries;
  DevBuf d_info, d_tchunks, d_qchunks, mbounds;
  DevBuf center, blo, bhi, keys, vals, tk, tv, hist;
  DevBuf left, right, rfirst, rlast, nparent, lparent, flags, nbox, keep, fidx, stmp;
  DevBuf nodes, recs, screen, seg, out_idx, out_dist, counters, rent, rcount, rlist, bad;
  int block_waves = 4;     // workgroup shape of the query launch (waves per workgroup)
  bool spill_build = true;  // the 8-resident-wave build of the query kernel
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  hipStream_t stream = nullptr;  // plan-owned: graphs cannot be captured on the null stream
  ~BatchPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

__device__ __forceinline__ int mesh_of(const MeshInfo *__restrict__ mi, int nm, int64_t pos, bool by_query,
                                       bool by_node) {
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    int64_t start = by_query ? mi[mid].qoff : (by_node ? mi[mid].ioff : mi[mid].toff);
    if (start <= pos) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Every triangle's vertex indices against its own mesh's vertex count; any index outside
// [0, nv) sets *bad (checked on the host before the pipeline reads vertices through them).
__global__ void __launch_bounds__(BB) k_bcheck(const int32_t *__restrict__ tris, const MeshInfo *__restrict__ mi, int nm,
                                               int64_t nt, int *__restrict__ bad) {
  const int64_t e = (int64_t)blockIdx.x * BB + threadIdx.x;
  if (e >= nt) return;
  const int64_t nv = mi[mesh_of(mi, nm, e, false, false)].nv;
  bool ok = true;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int32_t v = tris[3 * e + c];
    ok = ok && v >= 0 && v < nv;
  }
  if (!ok) atomicOr(bad, 1);
}

// Fill kernels (graph capture records no memsets).
__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * BB + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_init_mbounds(int *mb, int nm) {
  int64_t i = (int64_t)blockIdx.x * BB + threadIdx.x;
  if (i >= (int64_t)nm * 12) return;
  int slot = (int)(i % 12);
  bool is_lo = slot < 3 || (slot >= 6 && slot < 9);
  mb[i] = is_lo ? float_to_ord(INFINITY) : float_to_ord(-INFINITY);
}

// Element boxes/centres and per-mesh bounds; chunk = (mesh, first element, count), one mesh each.
__global__ void __launch_bounds__(BB) k_bprep(const float *__restrict__ verts, const int32_t *__restrict__ tris,
                                              const MeshInfo *__restrict__ mi, const int4 *__restrict__ chunks,
                                              float4 *__restrict__ center, float4 *__restrict__ blo,
                                              float4 *__restrict__ bhi, int *__restrict__ mb) {
  const int4 ch = chunks[blockIdx.x];
  const int m = ch.x;
  const int64_t voff = mi[m].voff, toff = mi[m].toff;
  float mn[6] = {INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY};
  float mx[6] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int j = threadIdx.x; j < ch.z; j += BB) {
    const int64_t e = toff + ch.y + j;
    const int32_t *t3 = tris + 3 * e;
    float3 v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t vi = voff + t3[c];
      v[c] = f3(verts[3 * vi], verts[3 * vi + 1], verts[3 * vi + 2]);
    }
    float3 lo = f3(fminf(v[0].x, fminf(v[1].x, v[2].x)), fminf(v[0].y, fminf(v[1].y, v[2].y)),
                   fminf(v[0].z, fminf(v[1].z, v[2].z)));
    float3 hi = f3(fmaxf(v[0].x, fmaxf(v[1].x, v[2].x)), fmaxf(v[0].y, fmaxf(v[1].y, v[2].y)),
                   fmaxf(v[0].z, fmaxf(v[1].z, v[2].z)));
    float3 s = add3(add3(v[0], v[1]), v[2]);
    float3 c = f3(s.x / 3.0f, s.y / 3.0f, s.z / 3.0f);
    center[e] = make_float4(c.x, c.y, c.z, 0.0f);
    blo[e] = make_float4(lo.x, lo.y, lo.z, 0.0f);
    bhi[e] = make_float4(hi.x, hi.y, hi.z, 0.0f);
    mn[0] = fminf(mn[0], lo.x); mn[1] = fminf(mn[1], lo.y); mn[2] = fminf(mn[2], lo.z);
    mx[0] = fmaxf(mx[0], hi.x); mx[1] = fmaxf(mx[1], hi.y); mx[2] = fmaxf(mx[2], hi.z);
    mn[3] = fminf(mn[3], c.x); mn[4] = fminf(mn[4], c.y); mn[5] = fminf(mn[5], c.z);
    mx[3] = fmaxf(mx[3], c.x); mx[4] = fmaxf(mx[4], c.y); mx[5] = fmaxf(mx[5], c.z);
  }
  __shared__ float red[BB / WAVE][12];
  const int w = threadIdx.x / WAVE;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    float a = wave_min(mn[k]), b = wave_max(mx[k]);
    if (lane_id() == 0) { red[w][k] = a; red[w][6 + k] = b; }
  }
  __syncthreads();
  if (threadIdx.x < 12) {
    const int k = threadIdx.x;
    float v = red[0][k];
    for (int ww = 1; ww < BB / WAVE; ++ww) v = k < 6 ? fminf(v, red[ww][k]) : fmaxf(v, red[ww][k]);
    // slots: elem lo xyz (0-2), elem hi (3-5), centre lo (6-8), centre hi (9-11)
    const int slot = k < 3 ? k : (k < 6 ? 6 + (k - 3) : (k < 9 ? 3 + (k - 6) : 9 + (k - 9)));
    if (k < 6) atomicMin(&mb[m * 12 + slot], float_to_ord(v));
    else atomicMax(&mb[m * 12 + slot], float_to_ord(v));
  }
}

__device__ __forceinline__ uint32_t spread_bits(uint32_t v, int bits) {
  uint32_t r = 0;
  for (int i = 0; i < bits; ++i) r |= ((v >> i) & 1u) << (3 * i);
  return r;
}

// key = (mesh << (32 - mesh_bits)) | Morton code of the point in the mesh's box (32 - mesh_bits bits)
__device__ __forceinline__ uint32_t batch_key(float3 p, const int *__restrict__ mb, int m, int mesh_bits, bool centre) {
  const int o = centre ? 6 : 0;
  float3 lo = f3(ord_to_float(mb[m * 12 + o]), ord_to_float(mb[m * 12 + o + 1]), ord_to_float(mb[m * 12 + o + 2]));
  float3 hi = f3(ord_to_float(mb[m * 12 + o + 3]), ord_to_float(mb[m * 12 + o + 4]), ord_to_float(mb[m * 12 + o + 5]));
  const int mbits = 32 - mesh_bits;
  float c[3] = {p.x, p.y, p.z}, l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z};
  uint32_t code = 0;
  for (int a = 0; a < 3; ++a) {
    // axis a owns code bit positions 3i + (2 - a) below mbits
    int bits = (mbits - (2 - a) + 2) / 3;
    float ext = h[a] - l[a];
    float u = ext > 0 ? (c[a] - l[a]) / ext : 0.0f;
    float cells = (float)(1u << bits);
    uint32_t q = (uint32_t)fminf(fmaxf(u * cells, 0.0f), cells - 1.0f);
    code |= spread_bits(q, bits) << (2 - a);
  }
  return ((uint32_t)m << mbits) | (code & ((mbits >= 32) ? 0xffffffffu : ((1u << mbits) - 1u)));
}

__global__ void __launch_bounds__(BB) k_bkeys(const float4 *__restrict__ center, const MeshInfo *__restrict__ mi, int nm,
                                              const int *__restrict__ mb, int mesh_bits, int64_t n,
                                              uint32_t *__restrict__ keys, uint32_t *__restrict__ vals) {
  int64_t e = (int64_t)blockIdx.x * BB + threadIdx.x;
  if (e >= n) return;
  int m = mesh_of(mi, nm, e, false, false);
  float4 c = center[e];
  keys[e] = batch_key(f3(c.x, c.y, c.z), mb, m, mesh_bits, true);
  vals[e] = (uint32_t)e;
}

__global__ void __launch_bounds__(BB) k_bqkeys(const float *__restrict__ q, const MeshInfo *__restrict__ mi, int nm,
                                               const int *__restrict__ mb, int mesh_bits, int64_t n,
                                               uint32_t *__restrict__ keys, uint32_t *__restrict__ vals) {
  int64_t j = (int64_t)blockIdx.x * BB + threadIdx.x;
  if (j >= n) return;
  int m = mesh_of(mi, nm, j, true, false);
  keys[j] = batch_key(f3(q[3 * j], q[3 * j + 1], q[3 * j + 2]), mb, m, mesh_bits, false);
  vals[j] = (uint32_t)j;
}

// Common prefix of sorted keys at global positions i, j of one mesh [a, b); -1 outside.
__device__ __forceinline__ int bdelta(const uint32_t *__restrict__ k, int64_t a, int64_t b, int64_t i, int64_t j) {
  if (j < a || j >= b) return -1;
  uint32_t x = k[i], y = k[j];
  if (x == y) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
  return __clz(x ^ y);
}

__global__ void __launch_bounds__(BB) k_bkarras(const uint32_t *__restrict__ k, const MeshInfo *__restrict__ mi, int nm,
                                                int64_t NT, int32_t *__restrict__ left, int32_t *__restrict__ right,
                                                int32_t *__restrict__ rfirst, int32_t *__restrict__ rlast,
                                                int32_t *__restrict__ nparent, int32_t *__restrict__ lparent) {
  int64_t pos = (int64_t)blockIdx.x * BB + threadIdx.x;
  if (pos >= NT) return;
  int m = mesh_of(mi, nm, pos, false, false);
  const int64_t a = mi[m].toff, nmesh = mi[m].nt, b = a + nmesh;
  const int64_t i = pos;  // global position; local index = pos - a
  if (nmesh < 2 || pos - a >= nmesh - 1) return;
  const int64_t ibase = mi[m].ioff - a;  // global node id = ibase + global position
  int d = (bdelta(k, a, b, i, i + 1) - bdelta(k, a, b, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = bdelta(k, a, b, i, i - d);
  int64_t lmax = 2;
  while (bdelta(k, a, b, i, i + lmax * d) > dmin) lmax *= 2;
  int64_t l = 0;
  for (int64_t t = lmax / 2; t >= 1; t /= 2)
    if (bdelta(k, a, b, i, i + (l + t) * d) > dmin) l += t;
  int64_t j = i + l * d;
  int dnode = bdelta(k, a, b, i, j);
  int64_t s = 0, div = 2, t;
  do {
    t = (l + div - 1) / div;
    if (t >= 1 && bdelta(k, a, b, i, i + (s + t) * d) > dnode) s += t;
    div *= 2;
  } while (t > 1);
  int64_t g = i + s * d + (d < 0 ? -1 : 0);
  int64_t first = i < j ? i : j, last = i < j ? j : i;
  const int64_t node = ibase + i;
  int32_t lc, rc;
  if (first == g) { lc = (int32_t)(g | 0x80000000u); lparent[g] = (int32_t)node; }
  else { lc = (int32_t)(ibase + g); nparent[ibase + g] = (int32_t)node; }
  if (last == g + 1) { rc = (int32_t)((g + 1) | 0x80000000u); lparent[g + 1] = (int32_t)node; }
  else { rc = (int32_t)(ibase + g + 1); nparent[ibase + g + 1] = (int32_t)node; }
  left[node] = lc;
  right[node] = rc;
  rfirst[node] = (int32_t)first;
  rlast[node] = (int32_t)last;
  if (i == a) nparent[node] = -1;
}

__global__ void __launch_bounds__(BB) k_bkeep(const int32_t *__restrict__ rfirst, const int32_t *__restrict__ rlast,
                                              const MeshInfo *__restrict__ mi, int nm, int64_t NI,
                                              uint32_t *__restrict__ keep) {
  int64_t g = (int64_t)blockIdx.x * BB + threadIdx.x;
  if (g >= NI) return;
  int m = mesh_of(mi, nm, rfirst[g], false, false);
  keep[g] = (rlast[g] - rfirst[g] + 1) > mi[m].leaf ? 1u : 0u;
}

__global__ void __launch_bounds__(BB) k_brecords(const float *__restrict__ verts, const int32_t *__restrict__ tris,
                                                 const uint32_t *__restrict__ sorted, const MeshInfo *__restrict__ mi,
                                                 int nm, int64_t NT, TriRec *__restrict__ out) {
  int64_t pos = (int64_t)blockIdx.x * BB + threadIdx.x;
  if (pos >= NT) return;
  const int64_t e = sorted[pos];
  const int m = mesh_of(mi, nm, e, false, false);
  const int64_t voff = mi[m].voff;
  float3 v[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int64_t vi = voff + tris[3 * e + c];
    v[c] = f3(verts[3 * vi], verts[3 * vi + 1], verts[3 * vi + 2]);
  }
  out[pos] = make_trirec(v[0], v[1], v[2], (int32_t)(e - mi[m].toff));
}

// Single-leaf roots of small meshes and the segment table of the query launch.
__global__ void k_bsegs(const MeshInfo *__restrict__ mi, int nm, const int *__restrict__ mb,
                        const uint32_t *__restrict__ fidx, int32_t nsmall, BNode *__restrict__ nodes,
                        const TriRec *__restrict__ recs, const ScreenRec *__restrict__ scr,
                        const float *__restrict__ q, const uint32_t *__restrict__ qorder,
                        int32_t *__restrict__ out_idx, float *__restrict__ out_dist, SegDesc *__restrict__ seg) {
  int m = blockIdx.x * BB + threadIdx.x;
  if (m >= nm) return;
  const MeshInfo &x = mi[m];
  float lo[3], hi[3], sc = 0;
  for (int a = 0; a < 3; ++a) {
    lo[a] = ord_to_float(mb[m * 12 + a]);
    hi[a] = ord_to_float(mb[m * 12 + 3 + a]);
    sc = fmaxf(sc, fmaxf(fabsf(lo[a]), fabsf(hi[a])));
  }
  int32_t root = 0;
  if (x.small_slot >= 0) {
    BNode nd;
    nd.a = make_float4(lo[0], lo[1], lo[2], hi[0]);
    nd.b = make_float4(hi[1], hi[2], INFINITY, INFINITY);
    nd.c = make_float4(INFINITY, -INFINITY, -INFINITY, -INFINITY);
    nd.d = make_int4(~(int32_t)x.toff, ~(int32_t)x.toff, (int32_t)x.nt, 0);
    nodes[x.small_slot] = nd;
    root = x.small_slot;
  } else if (x.nt > 0) {
    root = nsmall + (int32_t)fidx[x.ioff];
  }
  SegDesc d{};
  d.nodes = nodes;
  d.elems = recs;
  d.screen = scr;
  d.q = q;
  d.order = qorder + x.qoff;
  d.idx = out_idx;
  d.dist = out_dist;
  d.n = x.nt > 0 ? x.nq : 0;
  d.wave_begin = x.wave_begin;
  d.best_init = INFINITY;
  d.scale = sc;
  d.root = root;
  d.nelem = (int32_t)x.nt;
  d.lanes_log2 = x.lane;
  d.child_order = x.child_order;
  seg[x.seg_slot] = d;
}

// Queries of empty meshes get "not found".
__global__ void k_bempty(const MeshInfo *__restrict__ mi, int nm, int64_t NQ, int32_t *__restrict__ out_idx,
                         float *__restrict__ out_dist) {
  int64_t j = (int64_t)blockIdx.x * BB + threadIdx.x;
  if (j >= NQ) return;
  int m = mesh_of(mi, nm, j, true, false);
  if (mi[m].nt == 0) { out_idx[j] = -1; out_dist[j] = -1.0f; }
}

static void make_chunks(const std::vector<int64_t> &counts, std::vector<int4> &out) {
  out.clear();
  for (size_t m = 0; m < counts.size(); ++m)
    for (int64_t s = 0; s < counts[m]; s += CHUNK) out.push_back(make_int4((int)m, (int)s, (int)std::min<int64_t>(CHUNK, counts[m] - s), 0));
}

static void plan_pipeline(BatchPlan &P, hipStream_t s, hipEvent_t mid = nullptr) {
  const int nm = P.nm;
  const MeshInfo *mi = P.d_info.as<MeshInfo>();
  int *mb = P.mbounds.as<int>();
  hipLaunchKernelGGL(k_init_mbounds, dim3(ceil_div((uint64_t)nm * 12, BB)), dim3(BB), 0, s, mb, nm);
  hipLaunchKernelGGL(k_fill_i32, dim3(ceil_div(std::max<int64_t>(P.NT, 1), BB)), dim3(BB), 0, s, P.lparent.as<int32_t>(),
                     P.NT, -1);
  if (P.NI > 0)
    hipLaunchKernelGGL(k_fill_i32, dim3(ceil_div(P.NI, BB)), dim3(BB), 0, s, (int32_t *)P.flags.as<uint32_t>(), P.NI, 0);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(BB), 0, s, P.counters.as<int32_t>(), (int64_t)4, 0);
  launch_check("batch fill");
  if (P.NT > 0) {
    hipLaunchKernelGGL(k_bprep, dim3((uint32_t)P.tchunks.size()), dim3(BB), 0, s, P.verts.as<float>(), P.tris.as<int32_t>(),
                       mi, P.d_tchunks.as<int4>(), P.center.as<float4>(), P.blo.as<float4>(), P.bhi.as<float4>(), mb);
    launch_check("k_bprep");
    hipLaunchKernelGGL(k_bkeys, dim3(ceil_div(P.NT, BB)), dim3(BB), 0, s, P.center.as<float4>(), mi, nm, mb, P.mesh_bits,
                       P.NT, P.keys.as<uint32_t>(), P.vals.as<uint32_t>());
    launch_check("k_bkeys");
    radix_sort_pairs(P.keys.as<uint32_t>(), P.vals.as<uint32_t>(), P.tk.as<uint32_t>(), P.tv.as<uint32_t>(),
                     P.hist.as<uint32_t>(), P.NT, 0, 32, s);
    hipLaunchKernelGGL(k_bkarras, dim3(ceil_div(P.NT, BB)), dim3(BB), 0, s, P.keys.as<uint32_t>(), mi, nm, P.NT,
                       P.left.as<int32_t>(), P.right.as<int32_t>(), P.rfirst.as<int32_t>(), P.rlast.as<int32_t>(),
                       P.nparent.as<int32_t>(), P.lparent.as<int32_t>());
    launch_check("k_bkarras");
    launch_refit(P.NT, P.vals.as<uint32_t>(), P.blo.as<float4>(), P.bhi.as<float4>(), P.left.as<int32_t>(),
                 P.right.as<int32_t>(), P.rfirst.as<int32_t>(), P.rlast.as<int32_t>(), P.nparent.as<int32_t>(),
                 P.lparent.as<int32_t>(), P.flags.as<uint32_t>(), P.nbox.as<float>(), P.rent.as<int32_t>(),
                 P.rcount.as<uint32_t>(), P.rlist.as<int32_t>(), s);
    if (P.NI > 0) {
      hipLaunchKernelGGL(k_bkeep, dim3(ceil_div(P.NI, BB)), dim3(BB), 0, s, P.rfirst.as<int32_t>(), P.rlast.as<int32_t>(),
                         mi, nm, P.NI, P.keep.as<uint32_t>());
      launch_check("k_bkeep");
      scan_excl_u32(P.keep.as<uint32_t>(), P.fidx.as<uint32_t>(), P.NI, P.stmp.as<uint32_t>(), s);
      launch_emit(P.NI, P.nsmall, P.keep.as<uint32_t>(), P.fidx.as<uint32_t>(), P.left.as<int32_t>(), P.right.as<int32_t>(),
                  P.rfirst.as<int32_t>(), P.rlast.as<int32_t>(), P.vals.as<uint32_t>(), P.blo.as<float4>(),
                  P.bhi.as<float4>(), P.nbox.as<float>(), P.nodes.as<BNode>(), nullptr, s);
      launch_depth(P.NT, P.lparent.as<int32_t>(), P.nparent.as<int32_t>(), P.keep.as<uint32_t>(),
                   P.rfirst.as<int32_t>(), P.counters.as<int32_t>() + 1, s);
    }
    hipLaunchKernelGGL(k_brecords, dim3(ceil_div(P.NT, BB)), dim3(BB), 0, s, P.verts.as<float>(), P.tris.as<int32_t>(),
                       P.vals.as<uint32_t>(), mi, nm, P.NT, P.recs.as<TriRec>());
    launch_check("k_brecords");
    launch_screen_records(P.recs.as<TriRec>(), P.NT, P.screen.as<ScreenRec>(), s);
  }
  if (mid) THG_HIP(hipEventRecord(mid, s));
  if (P.NQ > 0) {
    // queries: reuse the key/value buffers (the element order is no longer needed after the records)
    hipLaunchKernelGGL(k_bqkeys, dim3(ceil_div(P.NQ, BB)), dim3(BB), 0, s, P.queries.as<float>(), mi, nm, mb, P.mesh_bits,
                       P.NQ, P.keys.as<uint32_t>(), P.vals.as<uint32_t>());
    launch_check("k_bqkeys");
    radix_sort_pairs(P.keys.as<uint32_t>(), P.vals.as<uint32_t>(), P.tk.as<uint32_t>(), P.tv.as<uint32_t>(),
                     P.hist.as<uint32_t>(), P.NQ, 0, 32, s);
    hipLaunchKernelGGL(k_bsegs, dim3(ceil_div(nm, BB)), dim3(BB), 0, s, mi, nm, mb, P.fidx.as<uint32_t>(), P.nsmall,
                       P.nodes.as<BNode>(), P.recs.as<TriRec>(), P.screen.as<ScreenRec>(), P.queries.as<float>(),
                       P.vals.as<uint32_t>(),
                       P.out_idx.as<int32_t>(), P.out_dist.as<float>(), P.seg.as<SegDesc>());
    launch_check("k_bsegs");
    hipLaunchKernelGGL(k_bempty, dim3(ceil_div(P.NQ, BB)), dim3(BB), 0, s, mi, nm, P.NQ, P.out_idx.as<int32_t>(),
                       P.out_dist.as<float>());
    launch_check("k_bempty");
    launch_closest_multi(P.seg.as<SegDesc>(), P.nwave_meshes, P.total_waves, false, s, false, P.block_waves,
                         P.spill_build);
    launch_closest_group_multi(P.seg.as<SegDesc>() + P.nwave_meshes, nm - P.nwave_meshes, P.lane_waves, false, 64, s);
  }
}

// Sparse meshes of a plan share each query among 2-64 lanes when THG_BATCH_LANE=1 (default: every
// mesh uses the wave-cooperative kernel; the per-query kernel must reserve a 64-entry stack here
// because the depth is known only after the captured build).
static bool batch_lane_split() {
  static const bool v = [] {
    const char *e = getenv("THG_BATCH_LANE");
    return e && atoi(e) == 1;
  }();
  return v;
}

// Fails with THG_E_INVALID unless every triangle's vertex indices lie in its own mesh's vertices
// (tris: the concatenated triangles of nm meshes, device memory); synchronises s.
static void check_batch_indices(const int32_t *tris, int nm, const int64_t *nv, const int64_t *nt, hipStream_t s) {
  std::vector<MeshInfo> mi(nm);
  int64_t to = 0;
  for (int m = 0; m < nm; ++m) {
    mi[m] = MeshInfo{};
    mi[m].toff = to;
    mi[m].nv = nv[m];
    to += nt[m];
  }
  if (to == 0) return;
  AsyncBuf dmi(sizeof(MeshInfo) * nm, s), bad(sizeof(int), s);
  THG_HIP(hipMemcpyAsync(dmi.p, mi.data(), sizeof(MeshInfo) * nm, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(BB), 0, s, bad.as<int32_t>(), (int64_t)1, 0);
  hipLaunchKernelGGL(k_bcheck, dim3(ceil_div(to, BB)), dim3(BB), 0, s, tris, dmi.as<MeshInfo>(), nm, to, bad.as<int>());
  launch_check("k_bcheck");
  int hb = 0;
  THG_HIP(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  THG_HIP(hipStreamSynchronize(s));
  require(hb == 0, "triangle vertex index out of range of its mesh");
}

// Leaf sizes chosen by measurement per size class (floor(log2(triangles))), cached for the process,
// and the child orders chosen at those leaf sizes.
static std::mutex g_tune_mu;
static std::map<int, int> g_tuned_leaf, g_tuned_order;

// The plan's query launch for one mesh alone: its queries spatially sorted, one segment, the given
// visiting order and launch shape; returns the average time of `reps` launches after one warm-up.
static double time_plan_kernel(const Bvh &b, const float *q, int64_t n, int child_order, int block_waves, bool spill,
                               int reps, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  QueryScratch ws;
  ws.timed = false;
  order_queries(b, q, n, ws, s);
  AsyncBuf oi(n * 4, s), od(n * 4, s), dseg(sizeof(SegDesc), s);
  SegDesc d{};
  d.nodes = b.nodes.p;
  d.elems = b.elems.p;
  d.screen = b.screen.p;
  d.q = q;
  d.order = ws.order;
  d.idx = oi.as<int32_t>();
  d.dist = od.as<float>();
  d.n = n;
  d.best_init = INFINITY;
  d.scale = b.scale;
  d.nelem = (int32_t)b.n;
  d.child_order = child_order;
  THG_HIP(hipMemcpyAsync(dseg.p, &d, sizeof(SegDesc), hipMemcpyHostToDevice, s));
  const uint64_t waves = (uint64_t)((n + SEG_WAVE_QUERIES - 1) / SEG_WAVE_QUERIES);
  launch_closest_multi(dseg.as<SegDesc>(), 1, waves, false, s, false, block_waves, spill);
  THG_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) launch_closest_multi(dseg.as<SegDesc>(), 1, waves, false, s, false, block_waves, spill);
  THG_HIP(hipEventRecord(e1, s));
  THG_HIP(hipEventSynchronize(e1));
  float ms = 0;
  THG_HIP(hipEventElapsedTime(&ms, e0, e1));
  return ms / reps;
}

}  // namespace thg

using namespace thg;

template <typename F> static int b_guarded(F f) {
  try {
    f();
    return THG_OK;
  } catch (const Status &s) {
    set_last_error(s.msg);
    return s.code;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return THG_E_OOM;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return THG_E_DEVICE;
  }
}

extern "C" {

int thg_batch_plan_create(thg_ctx ctx, int32_t nm, const int64_t *nv, const int64_t *nt, const int64_t *nq,
                          const int32_t *leaf_sizes, thg_batch_plan *out) {
  return b_guarded([&] {
    require(ctx && out, "ctx/out is NULL");
    require(nm >= 1 && nv && nt && nq, "bad mesh counts");
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    THG_HIP(hipSetDevice(c->device));
    std::unique_ptr<BatchPlan> P(new BatchPlan());
    P->ctx = c;
    P->nm = nm;
    P->mesh_bits = 1;
    while ((1 << P->mesh_bits) < nm) ++P->mesh_bits;
    require(P->mesh_bits <= 14, "too many meshes in one batch (at most 16384)");
    P->info.resize(nm);
    int64_t vo = 0, to = 0, qo = 0, io = 0;
    uint64_t waves = 0, lwaves = 0;
    for (int m = 0; m < nm; ++m) {
      require(nv[m] >= 0 && nt[m] >= 0 && nq[m] >= 0, "negative count");
      int leaf = leaf_sizes ? leaf_sizes[m] : 4;
      require(leaf >= 1 && leaf <= 16, "leaf sizes must be in 1..16");
      MeshInfo &x = P->info[m];
      x.voff = vo; x.toff = to; x.qoff = qo; x.ioff = io;
      x.nt = nt[m]; x.nq = nq[m]; x.nv = nv[m]; x.leaf = leaf;
      x.small_slot = (nt[m] >= 1 && nt[m] <= leaf) ? P->nsmall++ : -1;
      // The single-call rule (use_wave_traversal): wave-cooperative when there is at least one query
      // per 8 leaves; a collapsed LBVH has about 1.5 * triangles / leaf_size leaves.
      const int64_t leaves = (3 * nt[m]) / (2 * leaf) + 1;
      x.lane = batch_lane_split() && 8 * nq[m] < leaves ? group_lanes_log2(nq[m], 0) : 0;
      const uint64_t w =
          nt[m] > 0 ? (uint64_t)(((nq[m] << x.lane) + SEG_WAVE_QUERIES - 1) / SEG_WAVE_QUERIES) : 0;
      if (x.lane) { x.wave_begin = lwaves; lwaves += w; }
      else { x.wave_begin = waves; waves += w; P->nwave_meshes++; }
      vo += nv[m]; to += nt[m]; qo += nq[m]; io += nt[m] >= 2 ? nt[m] - 1 : 0;
    }
    for (int m = 0, iw = 0, il = P->nwave_meshes; m < nm; ++m) P->info[m].seg_slot = P->info[m].lane ? il++ : iw++;
    P->NV = vo; P->NT = to; P->NQ = qo; P->NI = io; P->total_waves = waves; P->lane_waves = lwaves;
    require(to < (int64_t)0x7ffffff0 && qo < (int64_t)0x7ffffff0, "batch too large for 32-bit indices");
    P->nv.assign(nv, nv + nm); P->nt.assign(nt, nt + nm); P->nq.assign(nq, nq + nm);
    make_chunks(P->nt, P->tchunks);
    const int64_t K = std::max<int64_t>(std::max(to, qo), 1);
    P->verts.alloc(std::max<int64_t>(vo, 1) * 12);
    P->tris.alloc(std::max<int64_t>(to, 1) * 12);
    P->queries.alloc(std::max<int64_t>(qo, 1) * 12);
    P->d_info.alloc(sizeof(MeshInfo) * nm);
    THG_HIP(hipMemcpy(P->d_info.p, P->info.data(), sizeof(MeshInfo) * nm, hipMemcpyHostToDevice));
    P->d_tchunks.alloc(sizeof(int4) * std::max<size_t>(P->tchunks.size(), 1));
    if (!P->tchunks.empty())
      THG_HIP(hipMemcpy(P->d_tchunks.p, P->tchunks.data(), sizeof(int4) * P->tchunks.size(), hipMemcpyHostToDevice));
    P->mbounds.alloc(sizeof(int) * 12 * nm);
    const int64_t T1 = std::max<int64_t>(to, 1), I1 = std::max<int64_t>(io, 1);
    P->center.alloc(T1 * 16); P->blo.alloc(T1 * 16); P->bhi.alloc(T1 * 16);
    P->keys.alloc(K * 4); P->vals.alloc(K * 4); P->tk.alloc(K * 4); P->tv.alloc(K * 4);
    P->hist.alloc(radix_hist_words(K) * 4);
    P->left.alloc(I1 * 4); P->right.alloc(I1 * 4); P->rfirst.alloc(I1 * 4); P->rlast.alloc(I1 * 4);
    P->nparent.alloc(I1 * 4); P->lparent.alloc(T1 * 4); P->flags.alloc(I1 * 4); P->nbox.alloc(I1 * 24);
    P->keep.alloc(I1 * 4); P->fidx.alloc((I1 + 1) * 4); P->stmp.alloc(scan_tmp_words(I1) * 4);
    P->nodes.alloc(sizeof(BNode) * (size_t)(P->nsmall + io + 1));
    P->recs.alloc(sizeof(TriRec) * T1);
    P->screen.alloc(sizeof(ScreenRec) * T1);
    P->seg.alloc(sizeof(SegDesc) * nm);
    P->out_idx.alloc(std::max<int64_t>(qo, 1) * 4);
    P->out_dist.alloc(std::max<int64_t>(qo, 1) * 4);
    P->counters.alloc(16);
    P->rent.alloc(T1 * 4);
    P->rcount.alloc(refit_count_words(T1) * 4);
    P->rlist.alloc(refit_needs_list(T1) ? T1 * 4 : 4);
    P->bad.alloc(4);
    THG_HIP(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
    *out = reinterpret_cast<thg_batch_plan>(P.release());
  });
}

int thg_batch_plan_destroy(thg_batch_plan plan) {
  return b_guarded([&] { delete reinterpret_cast<BatchPlan *>(plan); });
}

int thg_batch_plan_run(thg_batch_plan plan, const float *verts, const int32_t *tris, const float *queries, int32_t *idx,
                       float *dist, void *stream, double timings_ms[2]) {
  return b_guarded([&] {
    BatchPlan *P = reinterpret_cast<BatchPlan *>(plan);
    require(P != nullptr, "plan is NULL");
    require((P->NV == 0 || verts) && (P->NT == 0 || tris) && (P->NQ == 0 || (queries && idx)), "NULL array");
    THG_HIP(hipSetDevice(P->ctx->device));
    // the caller's queued work (e.g. producing the inputs) finishes first; the plan runs on its own
    // stream and returns when the outputs are written
    THG_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    hipStream_t s = P->stream;
    if (P->NV) THG_HIP(hipMemcpyAsync(P->verts.p, verts, P->NV * 12, hipMemcpyDeviceToDevice, s));
    if (P->NT) THG_HIP(hipMemcpyAsync(P->tris.p, tris, P->NT * 12, hipMemcpyDeviceToDevice, s));
    if (P->NQ) THG_HIP(hipMemcpyAsync(P->queries.p, queries, P->NQ * 12, hipMemcpyDeviceToDevice, s));
    if (P->NT) {
      // the pipeline reads vertices through the caller's indices: reject any outside its mesh
      hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(BB), 0, s, P->bad.as<int32_t>(), (int64_t)1, 0);
      hipLaunchKernelGGL(k_bcheck, dim3(ceil_div(P->NT, BB)), dim3(BB), 0, s, P->tris.as<int32_t>(),
                         P->d_info.as<MeshInfo>(), P->nm, P->NT, P->bad.as<int>());
      launch_check("k_bcheck");
      int hb = 0;
      THG_HIP(hipMemcpyAsync(&hb, P->bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
      THG_HIP(hipStreamSynchronize(s));
      require(hb == 0, "triangle vertex index out of range of its mesh");
    }
    if (timings_ms) {
      // eager run with an event between build and queries
      hipEvent_t e0, e1, e2;
      THG_HIP(hipEventCreate(&e0));
      THG_HIP(hipEventCreate(&e1));
      THG_HIP(hipEventCreate(&e2));
      THG_HIP(hipEventRecord(e0, s));
      plan_pipeline(*P, s, e1);
      THG_HIP(hipEventRecord(e2, s));
      THG_HIP(hipEventSynchronize(e2));
      float a = 0, b = 0;
      THG_HIP(hipEventElapsedTime(&a, e0, e1));
      THG_HIP(hipEventElapsedTime(&b, e1, e2));
      timings_ms[0] = a;
      timings_ms[1] = b;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      (void)hipEventDestroy(e2);
    } else {
      if (!P->exec) {
        THG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
          plan_pipeline(*P, s);
        } catch (...) {
          hipGraph_t g;
          (void)hipStreamEndCapture(s, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        THG_HIP(hipStreamEndCapture(s, &P->graph));
        THG_HIP(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
      }
      THG_HIP(hipGraphLaunch(P->exec, s));
    }
    if (P->NQ) {
      THG_HIP(hipMemcpyAsync(idx, P->out_idx.p, P->NQ * 4, hipMemcpyDeviceToDevice, s));
      if (dist) THG_HIP(hipMemcpyAsync(dist, P->out_dist.p, P->NQ * 4, hipMemcpyDeviceToDevice, s));
    }
    int32_t cnt[2];
    THG_HIP(hipMemcpyAsync(cnt, P->counters.p, 8, hipMemcpyDeviceToHost, s));
    THG_HIP(hipStreamSynchronize(s));
    require(cnt[1] < 64, "a mesh's BVH is deeper than the query kernels' 64-entry stack; results are invalid");
  });
}

}  // extern "C"

namespace thg {

// Per size class (floor(log2 triangles)): the leaf size (2..16), then at that leaf size the child
// visiting order (CHILD_*), each the fastest for the plan's query kernel on up to two sampled
// meshes of the class (the largest query batches, up to 262,144 queries each); both cached for the
// process. Then, if `launch` is given, the launch shape of the whole plan (waves per workgroup 1, 2,
// 4 or 8; the spill-free or the 8-resident-wave build), the fastest summed over every class's sampled
// meshes at their tuned leaf sizes and orders. A leaf size whose tree is too deep for the query
// kernel's 64-entry stack on a sampled mesh is not a candidate.
static void autotune_plan(Ctx *c, int32_t nm, const int64_t *nv, const int64_t *nt, const int64_t *nq,
                          const float *verts, const int32_t *tris, const float *queries, int32_t *leaf_sizes,
                          int32_t *child_orders, int32_t *launch) {
  THG_HIP(hipSetDevice(c->device));
  std::vector<int64_t> vo(nm + 1, 0), to(nm + 1, 0), qo(nm + 1, 0);
  for (int m = 0; m < nm; ++m) { vo[m + 1] = vo[m] + nv[m]; to[m + 1] = to[m] + nt[m]; qo[m + 1] = qo[m] + nq[m]; }
  std::map<int, std::vector<int>> buckets;
  for (int m = 0; m < nm; ++m) {
    if (nt[m] <= 0) continue;
    int b = 0;
    while ((int64_t(1) << (b + 1)) <= nt[m]) ++b;
    buckets[b].push_back(m);
  }
  const int leaves[] = {2, 4, 6, 8, 12, 16};
  const int REPS = 3;
  struct TuneRes {  // stream and events released on every path, including a failed build
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~TuneRes() {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      if (s) (void)hipStreamDestroy(s);
    }
  } R;
  THG_HIP(hipStreamCreateWithFlags(&R.s, hipStreamNonBlocking));
  THG_HIP(hipEventCreate(&R.e0));
  THG_HIP(hipEventCreate(&R.e1));
  hipStream_t s = R.s;
  check_batch_indices(tris, nm, nv, nt, s);
  struct Sampled {
    int m;
    int64_t n;
    std::unique_ptr<Bvh> b;  // built at the class's tuned leaf size
    int order;
  };
  std::vector<Sampled> all;
  auto build = [&](int m, int leaf) {
    std::unique_ptr<Bvh> b(new Bvh());
    b->ctx = c;
    // vertex indices are local to their mesh: offset the vertex pointer
    build_bvh(*b, verts + 3 * vo[m], tris + 3 * to[m], nt[m], false, leaf, s);
    return b;
  };
  for (auto &kv : buckets) {
    int best_leaf = 0, best_order = -1;
    {
      std::lock_guard<std::mutex> g(g_tune_mu);
      auto it = g_tuned_leaf.find(kv.first);
      if (it != g_tuned_leaf.end()) best_leaf = it->second;
      auto io = g_tuned_order.find(kv.first);
      if (io != g_tuned_order.end()) best_order = io->second;
    }
    std::vector<int> ms = kv.second;
    ms.erase(std::remove_if(ms.begin(), ms.end(), [&](int m) { return nq[m] <= 0; }), ms.end());
    std::sort(ms.begin(), ms.end(), [&](int a, int b) { return nq[a] > nq[b]; });
    if (ms.size() > 2) ms.resize(2);
    if (best_leaf == 0) {
      double best_ms = 1e30;
      for (int leaf : leaves) {
        // a candidate counts only if it timed every sampled mesh
        double tot = 0;
        bool timed_all = !ms.empty();
        for (int m : ms) {
          std::unique_ptr<Bvh> b = build(m, leaf);
          if (b->max_depth >= 64) { timed_all = false; break; }
          tot += time_plan_kernel(*b, queries + 3 * qo[m], std::min<int64_t>(nq[m], 262144), CHILD_MAJORITY, 4, true,
                                  REPS, s, R.e0, R.e1);
        }
        if (timed_all && tot < best_ms) { best_ms = tot; best_leaf = leaf; }
      }
      if (best_leaf == 0) best_leaf = 4;
      std::lock_guard<std::mutex> g(g_tune_mu);
      g_tuned_leaf[kv.first] = best_leaf;
    }
    std::vector<Sampled> here;
    if ((child_orders && best_order < 0) || launch) {
      for (int m : ms) {
        std::unique_ptr<Bvh> b = build(m, best_leaf);
        if (b->max_depth >= 64) continue;
        here.push_back(Sampled{m, std::min<int64_t>(nq[m], 262144), std::move(b), CHILD_MAJORITY});
      }
    }
    if (child_orders && best_order < 0) {
      double best_ms = 1e30;
      for (int order : {CHILD_MAJORITY, CHILD_LEADER, CHILD_BUILD}) {
        double tot = 0;
        for (auto &x : here)
          tot += time_plan_kernel(*x.b, queries + 3 * qo[x.m], x.n, order, 4, true, REPS, s, R.e0, R.e1);
        if (tot < best_ms) { best_ms = tot; best_order = order; }
      }
      if (here.empty()) best_order = CHILD_MAJORITY;
      std::lock_guard<std::mutex> g(g_tune_mu);
      g_tuned_order[kv.first] = best_order;
    }
    if (best_order < 0) best_order = CHILD_MAJORITY;
    for (auto &x : here) x.order = best_order;
    for (auto &x : here) all.push_back(std::move(x));
    for (int m : kv.second) {
      leaf_sizes[m] = best_leaf;
      if (child_orders) child_orders[m] = best_order;
    }
  }
  for (int m = 0; m < nm; ++m)
    if (nt[m] <= 0) {
      leaf_sizes[m] = 4;
      if (child_orders) child_orders[m] = CHILD_MAJORITY;
    }
  if (launch) {
    double best_ms = 1e30;
    launch[0] = 4;
    launch[1] = 1;
    for (int spill = 0; spill < 2; ++spill)
      for (int bw : {1, 2, 4, 8}) {
        double tot = 0;
        for (auto &x : all)
          tot += time_plan_kernel(*x.b, queries + 3 * qo[x.m], x.n, x.order, bw, spill != 0, REPS, s, R.e0, R.e1);
        if (!all.empty() && tot < best_ms) {
          best_ms = tot;
          launch[0] = bw;
          launch[1] = spill;
        }
      }
  }
}

}  // namespace thg

extern "C" {

int thg_autotune_leaf(thg_ctx ctx, int32_t nm, const int64_t *nv, const int64_t *nt, const int64_t *nq,
                      const float *verts, const int32_t *tris, const float *queries, int32_t *leaf_sizes) {
  return b_guarded([&] {
    require(ctx && nv && nt && nq && leaf_sizes && nm >= 1, "bad arguments");
    autotune_plan(reinterpret_cast<Ctx *>(ctx), nm, nv, nt, nq, verts, tris, queries, leaf_sizes, nullptr, nullptr);
  });
}

int thg_autotune_plan(thg_ctx ctx, int32_t nm, const int64_t *nv, const int64_t *nt, const int64_t *nq,
                      const float *verts, const int32_t *tris, const float *queries, int32_t *leaf_sizes,
                      int32_t *child_orders, int32_t launch[2]) {
  return b_guarded([&] {
    require(ctx && nv && nt && nq && leaf_sizes && child_orders && launch && nm >= 1, "bad arguments");
    autotune_plan(reinterpret_cast<Ctx *>(ctx), nm, nv, nt, nq, verts, tris, queries, leaf_sizes, child_orders,
                  launch);
  });
}

int thg_batch_plan_set_launch(thg_batch_plan plan, const int32_t *child_orders, int32_t block_waves,
                              int32_t spill_build) {
  return b_guarded([&] {
    BatchPlan *P = reinterpret_cast<BatchPlan *>(plan);
    require(P != nullptr, "plan is NULL");
    require(block_waves == 1 || block_waves == 2 || block_waves == 4 || block_waves == 8,
            "block_waves must be 1, 2, 4 or 8");
    if (child_orders)
      for (int m = 0; m < P->nm; ++m)
        require(child_orders[m] >= CHILD_MAJORITY && child_orders[m] <= CHILD_BUILD, "child order must be 0, 1 or 2");
    THG_HIP(hipSetDevice(P->ctx->device));
    THG_HIP(hipStreamSynchronize(P->stream));
    for (int m = 0; m < P->nm; ++m) P->info[m].child_order = child_orders ? child_orders[m] : CHILD_MAJORITY;
    THG_HIP(hipMemcpy(P->d_info.p, P->info.data(), sizeof(MeshInfo) * P->nm, hipMemcpyHostToDevice));
    P->block_waves = block_waves;
    P->spill_build = spill_build != 0;
    // the captured graph holds the old launch: capture again at the next run
    if (P->exec) { (void)hipGraphExecDestroy(P->exec); P->exec = nullptr; }
    if (P->graph) { (void)hipGraphDestroy(P->graph); P->graph = nullptr; }
  });
}

int thg_batch_closest(thg_ctx ctx, int32_t nm, const int64_t *nv, const int64_t *nt, const int64_t *nq,
                      const float *verts, const int32_t *tris, const float *queries, int32_t leaf_size, int32_t *idx,
                      float *dist, void *stream, double timings_ms[2]) {
  std::vector<int32_t> leaves(nm > 0 ? nm : 1, leaf_size > 0 ? leaf_size : 4), orders(nm > 0 ? nm : 1, 0);
  int32_t launch[2] = {4, 1};
  if (leaf_size == 0 && nm > 0) {
    int rc = thg_autotune_plan(ctx, nm, nv, nt, nq, verts, tris, queries, leaves.data(), orders.data(), launch);
    if (rc != THG_OK) return rc;
  }
  thg_batch_plan plan = nullptr;
  int rc = thg_batch_plan_create(ctx, nm, nv, nt, nq, leaves.data(), &plan);
  if (rc != THG_OK) return rc;
  rc = thg_batch_plan_set_launch(plan, orders.data(), launch[0], launch[1]);
  if (rc != THG_OK) {
    thg_batch_plan_destroy(plan);
    return rc;
  }
  rc = thg_batch_plan_run(plan, verts, tris, queries, idx, dist, stream, timings_ms);
  thg_batch_plan_destroy(plan);
  return rc;
}

}  // extern "C"

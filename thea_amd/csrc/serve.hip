// Request server: resident meshes under a memory budget, continuous batching of closest-point,
// k-NN and ray requests into multi-batch launches on concurrent streams (thg_server_* in
// include/thea_gpu.h). Thea has no serving layer; its queries run on the caller's thread.
//
// Each of `num_streams` workers loops: take every queued request that fits the remaining memory
// budget (FIFO; the first one that does not fit stops the scan, so large requests are not starved),
// up to the batch limits; group them by query type; for each group write one SegDesc per request
// (BVH, inputs, outputs, bound, first wave), sort large requests' queries spatially on the worker's
// stream, copy the table to the device and launch one kernel for the whole group; wait for the
// stream; complete the requests and release their working memory. A request arriving while a
// batch runs is picked up by the next free worker, so batches overlap on the device.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <tuple>
#include <thread>
#include <unordered_map>

#include "common.hpp"
#include "kernels_common.hpp"
#include "query.hpp"

namespace thg {

using Clock = std::chrono::steady_clock;

struct SrvMesh {
  std::unique_ptr<Mesh> mesh;
  std::unique_ptr<Bvh> bvh;
  size_t bytes = 0;
};

enum ReqState { QUEUED, RUNNING, DONE };

struct SrvRequest {
  int64_t id = 0;
  thg_request r{};
  const Bvh *bvh = nullptr;  // the mesh's BVH, looked up under the lock at submit (its address is stable)
  ReqState state = QUEUED;
  int status = THG_OK;
  size_t bytes = 0;
  Clock::time_point t_submit, t_done;
};

struct Server {
  Ctx *ctx = nullptr;
  thg_server_opts opts{};
  size_t budget = 0, resident = 0, inflight = 0, peak_inflight = 0;
  // Owned meshes; the vector is read and grown only under `mu`, and workers use the Bvh pointers
  // their requests captured at submit, which stay valid until the server is destroyed.
  std::vector<std::unique_ptr<SrvMesh>> meshes;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<int64_t> queue;
  int64_t queued_points = 0;  // points of the queued requests
  std::unordered_map<int64_t, SrvRequest> reqs;
  int64_t next_id = 1;
  bool stopping = false;
  std::vector<std::thread> workers;
  std::vector<hipStream_t> streams;
  std::string worker_error;
  // statistics
  int64_t completed = 0, launches = 0, batched_requests = 0, rejected = 0, queries_done = 0;
  std::vector<double> latencies;
};

// Server-side working memory of a request: its descriptor, and the spatial-sort scratch of large
// requests (keys, values, ping-pong buffers and histograms).
static size_t request_bytes(const thg_request &r) {
  size_t b = sizeof(SegDesc);
  if (r.n >= ORDER_MIN_QUERIES) b += (size_t)r.n * 16 + radix_hist_words(r.n) * 4;
  return b;
}

constexpr int64_t FAIR_MIN_POINTS = (int64_t)1 << 18;

static void worker_loop(Server *S, int wi) {
  hipStream_t s = S->streams[wi];
  (void)hipSetDevice(S->ctx->device);
  const int maxreq = S->opts.max_batch_requests;
  // descriptor tables: pinned host staging + device copy, one region per query type
  SegDesc *host_desc = nullptr;
  DevBuf dev_desc;
  try {
    THG_HIP(hipHostMalloc((void **)&host_desc, sizeof(SegDesc) * maxreq * 3, hipHostMallocDefault));
    dev_desc.alloc(sizeof(SegDesc) * maxreq * 3);
  } catch (const Status &st) {
    std::lock_guard<std::mutex> g(S->mu);
    S->worker_error = st.msg;
    return;
  }
  while (true) {
    std::vector<int64_t> batch;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait(lk, [&] { return S->stopping || !S->queue.empty(); });
      if (S->stopping && S->queue.empty()) break;
      int64_t nq = 0;
      size_t avail = S->budget - S->resident;
      // A worker takes its share of the queued points (all of them split over the streams, at least
      // FAIR_MIN_POINTS), so a deep queue keeps every stream busy instead of landing on one.
      const int64_t share = std::max<int64_t>(S->queued_points / (int64_t)S->streams.size(), FAIR_MIN_POINTS);
      while (!S->queue.empty() && (int)batch.size() < (S->opts.coalesce < 0 ? 1 : maxreq)) {
        SrvRequest &r = S->reqs[S->queue.front()];
        if (!batch.empty() && (nq + r.r.n > S->opts.max_batch_queries || nq >= share)) break;
        if (S->inflight + r.bytes > avail) break;  // admission by memory budget
        S->inflight += r.bytes;
        S->peak_inflight = std::max(S->peak_inflight, S->inflight);
        r.state = RUNNING;
        nq += r.r.n;
        S->queued_points -= r.r.n;
        batch.push_back(r.id);
        S->queue.pop_front();
      }
      if (batch.empty()) {
        if (S->inflight == 0) {
          // cannot fit even with nothing in flight (meshes added after it was queued): fail it
          SrvRequest &r = S->reqs[S->queue.front()];
          S->queue.pop_front();
          S->queued_points -= r.r.n;
          r.state = DONE;
          r.status = THG_E_BUDGET;
          r.t_done = Clock::now();
          ++S->rejected;
          S->worker_error = "request no longer fits the memory budget";
          S->cv_done.notify_all();
          continue;
        }
        // head does not fit while other batches hold memory: wait for a completion
        S->cv_done.wait(lk);
        continue;
      }
    }
    int status = THG_OK;
    std::string err;
    std::vector<std::unique_ptr<QueryScratch>> scratch;
    try {
      // group by type, traversal (wave-cooperative for dense requests, one query per lane for
      // sparse ones, as the single-call entry points choose) and, for k-NN, k
      std::map<std::tuple<int, int, int>, std::vector<SrvRequest *>> groups;
      {
        std::lock_guard<std::mutex> g(S->mu);
        for (int64_t id : batch) {
          SrvRequest &r = S->reqs[id];
          int kk = r.r.type == THG_REQ_KNN ? (r.r.k <= 4 ? 4 : r.r.k <= 8 ? 8 : r.r.k <= 16 ? 16 : 32) : 0;
          int lane = r.r.type != THG_REQ_RAY && !use_wave_traversal(*r.bvh, r.r.n) ? 1 : 0;
          groups[{r.r.type, lane, kk}].push_back(&r);
        }
      }
      int region = 0;
      for (auto &kv : groups) {
        const int type = std::get<0>(kv.first);
        const bool lane = std::get<1>(kv.first) != 0;
        int depth1 = 0;
        for (SrvRequest *r : kv.second) depth1 = std::max(depth1, r->bvh->max_depth + 1);
        // sparse requests share each query among several lanes (group kernels)
        bool grouped = false;
        if (lane)
          for (SrvRequest *r : kv.second)
            grouped = grouped || group_lanes_log2(r->r.n, r->bvh->max_depth) > 0;
        const int stack = std::max(32, depth1);
        // k-NN segments of one launch share k; split by exact k
        std::map<int, std::vector<SrvRequest *>> byk;
        for (SrvRequest *r : kv.second) byk[type == THG_REQ_KNN ? r->r.k : 0].push_back(r);
        for (auto &kk : byk) {
          auto &list = kk.second;
          SegDesc *hd = host_desc + (size_t)(region % 3) * maxreq;
          SegDesc *dd = dev_desc.as<SegDesc>() + (size_t)(region % 3) * maxreq;
          if (region >= 3) THG_HIP(hipStreamSynchronize(s));  // the staging region is reused
          ++region;
          uint64_t waves = 0;
          int nseg = 0;
          for (SrvRequest *r : list) {
            const thg_request &q = r->r;
            const Bvh &b = *r->bvh;
            SegDesc d{};
            d.nodes = b.nodes.p;
            d.elems = b.elems.p;
            d.rays = b.rays.p;
            d.qnodes = b.qnodes.p;
            d.screen = b.screen.p;
            d.q = q.q;
            d.q2 = q.dirs;
            d.idx = q.idx;
            d.n = q.n;
            d.scale = b.scale;
            d.k = q.k;
            d.nelem = (int32_t)b.n;
            d.exclude_self = q.exclude_self;
            if (type == THG_REQ_RAY) {
              d.dist = q.dist;
              d.cp = q.cp;
              d.max_time = (float)q.bound;
              d.best_init = q.bound >= 0 ? std::nextafter((float)q.bound, -INFINITY) : INFINITY;
            } else {
              d.dist = q.dist;
              d.cp = q.cp;
              d.best_init = bound_to_best(q.bound);
            }
            if (q.n >= ORDER_MIN_QUERIES) {
              scratch.emplace_back(new QueryScratch());
              if (type == THG_REQ_RAY) order_rays(b, q.q, q.dirs, q.n, *scratch.back(), s);
              else order_queries(b, q.q, q.n, *scratch.back(), s);
              d.order = scratch.back()->order;
            }
            d.lanes_log2 = grouped ? group_lanes_log2(q.n, b.max_depth) : 0;
            d.wave_begin = waves;
            waves += (uint64_t)(((q.n << d.lanes_log2) + SEG_WAVE_QUERIES - 1) / SEG_WAVE_QUERIES);
            hd[nseg++] = d;
          }
          THG_HIP(hipMemcpyAsync(dd, hd, sizeof(SegDesc) * nseg, hipMemcpyHostToDevice, s));
          if (type == THG_REQ_CLOSEST) {
            if (grouped) launch_closest_group_multi(dd, nseg, waves, false, stack, s);
            else if (lane) launch_closest_lane_multi(dd, nseg, waves, false, stack, s);
            else launch_closest_multi(dd, nseg, waves, false, s, /*concurrent=*/true);
          } else if (type == THG_REQ_KNN) {
            if (grouped) launch_knn_group_multi(dd, nseg, waves, kk.first, false, stack, s);
            else if (lane) launch_knn_lane_multi(dd, nseg, waves, kk.first, false, stack, s);
            else launch_knn_multi(dd, nseg, waves, kk.first, false, s);
          } else {
            int q_stack = 0;
            bool q4 = true;
            for (SrvRequest *r : list) {
              q4 = q4 && r->bvh->qnodes.p != nullptr && q4_stack(*r->bvh) <= Q4_MAX_STACK;
              q_stack = std::max(q_stack, q4_stack(*r->bvh));
            }
            launch_ray_multi(dd, nseg, waves, stack, s, q4 ? q_stack : 0);
          }
        }
      }
      THG_HIP(hipStreamSynchronize(s));
    } catch (const Status &st) {
      status = st.code;
      err = st.msg;
      (void)hipStreamSynchronize(s);
    }
    scratch.clear();
    auto now = Clock::now();
    {
      std::lock_guard<std::mutex> g(S->mu);
      for (int64_t id : batch) {
        SrvRequest &r = S->reqs[id];
        r.state = DONE;
        r.status = status;
        r.t_done = now;
        S->inflight -= r.bytes;
        ++S->completed;
        S->queries_done += r.r.n;
        S->latencies.push_back(std::chrono::duration<double, std::micro>(now - r.t_submit).count());
      }
      ++S->launches;
      S->batched_requests += (int64_t)batch.size();
      if (status != THG_OK) S->worker_error = err;
    }
    S->cv_done.notify_all();
    S->cv_work.notify_all();
  }
  (void)hipHostFree(host_desc);
}

}  // namespace thg

using namespace thg;

template <typename F> static int srv_guarded(F f) {
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

int thg_server_create(thg_ctx ctx, const thg_server_opts *opts, thg_server *out) {
  return srv_guarded([&] {
    require(ctx && out, "ctx/out is NULL");
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    THG_HIP(hipSetDevice(c->device));
    std::unique_ptr<Server> S(new Server());
    S->ctx = c;
    if (opts) S->opts = *opts;
    if (S->opts.num_streams <= 0) S->opts.num_streams = 4;
    if (S->opts.max_batch_requests <= 0) S->opts.max_batch_requests = 256;
    if (S->opts.max_batch_queries <= 0) S->opts.max_batch_queries = (int64_t)16 << 20;
    if (S->opts.coalesce == 0) S->opts.coalesce = 1;
    require(S->opts.num_streams <= 16, "num_streams must be <= 16");
    if (S->opts.memory_budget_bytes > 0) {
      S->budget = (size_t)S->opts.memory_budget_bytes;
    } else {
      size_t fr = 0, tot = 0;
      THG_HIP(hipMemGetInfo(&fr, &tot));
      S->budget = (size_t)(0.9 * (double)fr);
    }
    for (int i = 0; i < S->opts.num_streams; ++i) {
      hipStream_t st;
      THG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));  // default priority
      S->streams.push_back(st);
    }
    Server *raw = S.get();
    for (int i = 0; i < S->opts.num_streams; ++i) S->workers.emplace_back(worker_loop, raw, i);
    *out = reinterpret_cast<thg_server>(S.release());
  });
}

int thg_server_destroy(thg_server srv) {
  return srv_guarded([&] {
    Server *S = reinterpret_cast<Server *>(srv);
    if (!S) return;
    {
      std::lock_guard<std::mutex> g(S->mu);
      // queued requests are cancelled; running batches finish
      for (int64_t id : S->queue) {
        SrvRequest &r = S->reqs[id];
        r.state = DONE;
        r.status = THG_E_CANCELLED;
      }
      S->queue.clear();
      S->stopping = true;
    }
    S->cv_work.notify_all();
    S->cv_done.notify_all();
    for (auto &t : S->workers) t.join();
    for (auto st : S->streams) (void)hipStreamDestroy(st);
    delete S;
  });
}

int thg_server_add_mesh(thg_server srv, const float *verts, int64_t nv, const int32_t *tris, int64_t nt, int32_t mem,
                        int32_t leaf_size, int32_t *mesh_id) {
  thg_mesh m = nullptr;
  int rc = thg_mesh_create(srv ? reinterpret_cast<thg_ctx>(reinterpret_cast<Server *>(srv)->ctx) : nullptr, verts, nv,
                           tris, nt, mem, &m);
  if (rc != THG_OK) return rc;
  return srv_guarded([&] {
    Server *S = reinterpret_cast<Server *>(srv);
    std::unique_ptr<SrvMesh> smp(new SrvMesh());
    SrvMesh &sm = *smp;
    sm.mesh.reset(reinterpret_cast<Mesh *>(m));
    require(mesh_id != nullptr, "mesh_id is NULL");
    require(nt > 0, "a served mesh needs at least one triangle");
    require(leaf_size >= 0 && leaf_size <= 16, "leaf_size must be in 0..16");
    size_t mesh_bytes = sm.mesh->verts.bytes + sm.mesh->tris.bytes;
    // records (128 + 48 + 32 B) + binary nodes (at most n - 1 of 64 B) + leaf bounds (32 B per record
    // position) + 4-wide ray nodes (the binary nodes at even depth, at most about two thirds of them)
    // bound the BVH's size before it is built
    size_t est = mesh_bytes + (size_t)nt * (sizeof(TriRec) + sizeof(RayTri) + sizeof(ScreenRec) + sizeof(BNode) + sizeof(LeafBound)) +
                 (size_t)nt * sizeof(QNode4) * 2 / 3;
    {
      std::lock_guard<std::mutex> g(S->mu);
      if (S->resident + est > S->budget)
        fail(THG_E_BUDGET, "mesh of " + std::to_string(est) + " bytes does not fit the memory budget (" +
                               std::to_string(S->budget - S->resident) + " bytes left)");
    }
    sm.bvh.reset(new Bvh());
    sm.bvh->ctx = S->ctx;
    sm.bvh->mesh = sm.mesh.get();
    hipStream_t st;
    THG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    try {
      build_bvh(*sm.bvh, sm.mesh->verts.as<float>(), sm.mesh->tris.as<int32_t>(), nt, false,
                leaf_size > 0 ? leaf_size : 4, st);
      THG_HIP(hipStreamSynchronize(st));
    } catch (...) {
      (void)hipStreamDestroy(st);
      throw;
    }
    (void)hipStreamDestroy(st);
    require(sm.bvh->max_depth < 64, "mesh BVH deeper than the serving kernels' 64-entry stack");
    sm.bytes = mesh_bytes + sm.bvh->device_bytes();
    std::lock_guard<std::mutex> g(S->mu);
    S->resident += sm.bytes;
    S->meshes.push_back(std::move(smp));
    *mesh_id = (int32_t)S->meshes.size() - 1;
  });
}

int thg_server_submit(thg_server srv, const thg_request *req, int64_t *req_id) {
  return srv_guarded([&] {
    Server *S = reinterpret_cast<Server *>(srv);
    require(S && req && req_id, "NULL argument");
    const thg_request &r = *req;
    require(r.type == THG_REQ_CLOSEST || r.type == THG_REQ_KNN || r.type == THG_REQ_RAY, "unknown request type");
    require(r.n >= 0, "negative query count");
    require(r.n == 0 || (r.q && r.idx), "q/idx is NULL");
    require(r.type != THG_REQ_RAY || r.n == 0 || r.dirs, "ray request without directions");
    require(r.type != THG_REQ_KNN || (r.k >= 1 && r.k <= 32), "k must be in 1..32");
    require(!std::isnan(r.bound), "bound is NaN");
    SrvRequest sr;
    sr.r = r;
    sr.bytes = request_bytes(r);
    sr.t_submit = Clock::now();
    {
      std::lock_guard<std::mutex> g(S->mu);
      require(r.mesh >= 0 && r.mesh < (int32_t)S->meshes.size(), "unknown mesh id");
      sr.bvh = S->meshes[r.mesh]->bvh.get();
      require(r.type != THG_REQ_KNN || !r.exclude_self || r.n <= sr.bvh->n,
              "exclude_self needs the mesh's own elements as queries");
      require(!S->stopping, "server is shutting down");
      if (sr.bytes > S->budget - S->resident) {
        ++S->rejected;
        fail(THG_E_BUDGET, "request needs " + std::to_string(sr.bytes) + " working bytes; the budget leaves " +
                               std::to_string(S->budget - S->resident));
      }
      sr.id = S->next_id++;
      *req_id = sr.id;
      S->reqs[sr.id] = sr;
      S->queue.push_back(sr.id);
      S->queued_points += sr.r.n;
    }
    S->cv_work.notify_one();
  });
}

int thg_server_wait(thg_server srv, int64_t req_id, int64_t timeout_us, double *latency_us) {
  Server *S = reinterpret_cast<Server *>(srv);
  if (!S) {
    set_last_error("srv is NULL");
    return THG_E_INVALID;
  }
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->reqs.find(req_id);
  if (it == S->reqs.end()) {
    set_last_error("unknown request id");
    return THG_E_INVALID;
  }
  auto done = [&] { return S->reqs[req_id].state == DONE; };
  if (timeout_us < 0) S->cv_done.wait(lk, done);
  else if (!S->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), done)) return THG_PENDING;
  SrvRequest r = S->reqs[req_id];
  S->reqs.erase(req_id);
  if (latency_us) *latency_us = std::chrono::duration<double, std::micro>(r.t_done - r.t_submit).count();
  if (r.status != THG_OK) set_last_error(r.status == THG_E_CANCELLED ? "request cancelled" : S->worker_error);
  return r.status;
}

int thg_server_cancel(thg_server srv, int64_t req_id) {
  Server *S = reinterpret_cast<Server *>(srv);
  if (!S) {
    set_last_error("srv is NULL");
    return THG_E_INVALID;
  }
  {
    std::lock_guard<std::mutex> g(S->mu);
    auto it = std::find(S->queue.begin(), S->queue.end(), req_id);
    if (it == S->queue.end()) {
      set_last_error("request is not queued (running, finished or unknown)");
      return THG_E_INVALID;
    }
    S->queue.erase(it);
    SrvRequest &r = S->reqs[req_id];
    S->queued_points -= r.r.n;
    r.state = DONE;
    r.status = THG_E_CANCELLED;
    r.t_done = Clock::now();
  }
  S->cv_done.notify_all();
  S->cv_work.notify_all();
  return THG_OK;
}

int thg_server_stats(thg_server srv, double out[10], int32_t reset) {
  return srv_guarded([&] {
    Server *S = reinterpret_cast<Server *>(srv);
    require(S && out, "NULL argument");
    std::lock_guard<std::mutex> g(S->mu);
    std::vector<double> lat = S->latencies;
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) {
      if (lat.empty()) return 0.0;
      size_t i = (size_t)std::ceil(p * (double)lat.size()) - 1;
      return lat[std::min(i, lat.size() - 1)];
    };
    double mean = 0;
    for (double x : lat) mean += x;
    out[0] = (double)S->completed;
    out[1] = (double)S->launches;
    out[2] = S->launches ? (double)S->batched_requests / (double)S->launches : 0.0;
    out[3] = pct(0.50);
    out[4] = pct(0.99);
    out[5] = (double)S->resident;
    out[6] = (double)S->peak_inflight;
    out[7] = (double)S->rejected;
    out[8] = (double)S->queries_done;
    out[9] = lat.empty() ? 0.0 : mean / (double)lat.size();
    if (reset) {
      S->completed = S->launches = S->batched_requests = S->rejected = S->queries_done = 0;
      S->latencies.clear();
      S->peak_inflight = S->inflight;
    }
  });
}

}  // extern "C"

// Served path: continuous batching of join requests over resident layers on one device.
//
// The reference serves one request at a time under a global GPU mutex (daemon/gpujoind.cu:35, 189-206), so the
// device idles during every host phase of a request (count copies, scans, row copies, callbacks, socket sends) and a
// small request cannot fill 256 CUs. Here a scheduler owns N lanes, each a stream with its own grow-only scratch
// pool and a worker thread. Requests queue in arrival order; the queue head is admitted when a lane is free and its
// scratch estimate fits the budget, so requests that arrive while others run are admitted as lanes and memory free
// up, and their kernels overlap on the device (HIP runs kernels of different streams concurrently). Each lane runs
// the same code as the direct entry points, so a request's rows are the same alone and in the mix.
//
// Windowed joins, distance joins and eroded intersects of the same layers and options are also batched: a lane that
// admits one takes up to max_batch - 1 more such requests of its kind from anywhere in the queue and runs them as one
// candidate search and one refinement pass (join_batch_run, dwithin_batch_run, erode_batch_run), which pays a small
// request's chain of launches and synchronisations, and its longest pair, once per batch.
#include "capi_internal.h"
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

namespace {

using Clock = std::chrono::steady_clock;
double ms_between(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); }

struct Req {
  gsj_request r{};
  uint64_t ticket = 0;
  uint64_t est = 0;   // scratch estimate, bytes
  Clock::time_point t_submit, t_start, t_end;
  int lane = -1;
  int batch = 1;   // requests run together in its pass
  bool done = false;
  std::condition_variable cv_done;   // its waiter's: a finished request wakes only the thread waiting for it
  int rc = GSJ_OK;
  std::string msg;
  gsj_join_stats stats{};
};

// Scratch estimate of a request: candidate pairs bounded by the sub-batch limit, ~120 B of pair state each
// (pairs, flags, areas, evidence, partition lists, compaction keys), plus per-probe traversal counters. The probes are
// the probe layer's, or for a window over a layer with tile boxes, the probes of the tiles meeting the window.
uint64_t estimate_scratch(const gsj_request& r) {
  uint64_t nq = r.P ? r.P->L.n : 0;
  const std::vector<float4>& tiles = r.P ? r.P->L.h_tiles : std::vector<float4>();
  if (r.opts.has_window && !tiles.empty()) {
    const double* w = r.opts.window;
    uint64_t meet = 0;
    for (const float4& b : tiles) meet += ((double)b.x <= w[2] && (double)b.z >= w[0] && (double)b.y <= w[3] && (double)b.w >= w[1]) ? 1 : 0;
    nq = std::min<uint64_t>(nq, meet * gsj::PROBE_TILE);
  }
  const uint64_t max_pairs = r.opts.max_pairs ? r.opts.max_pairs : (128ull << 20);
  const uint64_t pairs = std::min<uint64_t>(max_pairs, 8 * nq + 1024);
  return pairs * 120 + nq * 24;
}

}  // namespace

struct gsj_sched {
  gsj_ctx* owner = nullptr;
  std::vector<std::unique_ptr<gsj::Context>> lanes;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_work;
  std::deque<Req*> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
  uint64_t next_ticket = 1;
  uint64_t budget = 0, in_use = 0;
  int running = 0;
  int max_batch = 16;
  bool stop = false;

  void worker(int lane);
  void run(gsj::Context& c, Req& q);
  void run_batch(gsj::Context& c, std::vector<Req*>& group);
  static bool batchable(const Req& q) {
    if (q.r.kind == GSJ_REQ_JOIN) return gsj::join_batchable(q.r.P->L, q.r.R->L, q.r.opts);
    if (q.r.kind == GSJ_REQ_DWITHIN) return gsj::dwithin_batchable(q.r.P->L, q.r.R->L, q.r.opts);
    if (q.r.kind == GSJ_REQ_INTERSECTS_BUFFERED) return gsj::erode_batchable(q.r.P->L, q.r.R->L, q.r.opts, self_need(q.r));
    return false;
  }
  bool batch_head(const Req& q) const { return max_batch > 1 && batchable(q); }
  static bool batch_with(const Req& head, const Req& x) {
    if (x.r.kind != head.r.kind || x.r.P != head.r.P || x.r.R != head.r.R || !batchable(x)) return false;
    const gsj_join_opts &a = head.r.opts, &b = x.r.opts;
    if (x.r.kind == GSJ_REQ_JOIN) return gsj::join_compatible(a, b);
    const bool common = x.r.same == head.r.same && a.upper == b.upper && a.batch_q == b.batch_q && a.max_pairs == b.max_pairs;
    if (x.r.kind == GSJ_REQ_DWITHIN) return common && x.r.report_distance == head.r.report_distance;
    return common && a.snap_half_ulp == b.snap_half_ulp && (x.r.need_listed != 0) == (head.r.need_listed != 0);
  }
  static gsj::SelfNeedHost self_need(const gsj_request& r) {
    gsj::SelfNeedHost s;
    s.flags = r.need_self;
    s.list = r.need_list;
    s.n = r.n_need;
    s.listed = r.need_listed != 0;
    return s;
  }
};

void gsj_sched::run_batch(gsj::Context& c, std::vector<Req*>& group) {
  const gsj_request& h = group[0]->r;
  std::vector<int> rcs(group.size(), GSJ_OK);
  int rc = GSJ_OK;
  if (h.kind == GSJ_REQ_JOIN) {
    std::vector<gsj::BatchJoin> b(group.size());
    for (size_t i = 0; i < group.size(); i++) {
      b[i].o = &group[i]->r.opts;
      b[i].cb = group[i]->r.rows_cb;
      b[i].user = group[i]->r.user;
      b[i].st = &group[i]->stats;
    }
    rc = gsj::guarded_status([&] { gsj::join_batch_run(c, h.P->L, h.R->L, b); });
    for (size_t i = 0; i < group.size(); i++) rcs[i] = b[i].rc;
  } else if (h.kind == GSJ_REQ_INTERSECTS_BUFFERED) {
    std::vector<gsj::BatchErode> b(group.size());
    for (size_t i = 0; i < group.size(); i++) {
      b[i].o = &group[i]->r.opts;
      b[i].ra = group[i]->r.ra;
      b[i].rb = group[i]->r.rb;
      b[i].cb = group[i]->r.dist_cb;
      b[i].user = group[i]->r.user;
      b[i].st = &group[i]->stats;
    }
    rc = gsj::guarded_status([&] { gsj::erode_batch_run(c, h.P->L, h.R->L, h.same != 0, h.need_listed != 0, b); });
    for (size_t i = 0; i < group.size(); i++) rcs[i] = b[i].rc;
  } else {
    std::vector<gsj::BatchDist> b(group.size());
    for (size_t i = 0; i < group.size(); i++) {
      b[i].o = &group[i]->r.opts;
      b[i].d = group[i]->r.d;
      b[i].cb = group[i]->r.dist_cb;
      b[i].user = group[i]->r.user;
      b[i].st = &group[i]->stats;
    }
    rc = gsj::guarded_status([&] { gsj::dwithin_batch_run(c, h.P->L, h.R->L, h.same != 0, h.report_distance != 0, b); });
    for (size_t i = 0; i < group.size(); i++) rcs[i] = b[i].rc;
  }
  const std::string msg = rc == GSJ_OK ? std::string() : gsj::g_last_error;
  for (size_t i = 0; i < group.size(); i++) {
    Req& q = *group[i];
    q.batch = (int)group.size();
    q.rc = rc != GSJ_OK ? rc : rcs[i];
    q.msg = rc != GSJ_OK ? msg : (rcs[i] != GSJ_OK ? std::string("request cancelled by the row callback") : std::string());
  }
}

void gsj_sched::run(gsj::Context& c, Req& q) {
  const gsj_request& r = q.r;
  switch (r.kind) {
    case GSJ_REQ_JOIN:
      gsj::join_run(c, r.P->L, r.R->L, r.opts, r.rows_cb, r.user, &q.stats, nullptr);
      break;
    case GSJ_REQ_POINTS_IN_POLYGONS:
      gsj::points_join_run(c, r.P->L, r.R->L, r.opts, r.points_cb, r.user, &q.stats);
      break;
    case GSJ_REQ_CLUSTER:
      gsj::cluster_run(c, r.P->L, r.opts, r.root, r.nroot, &q.stats);
      break;
    case GSJ_REQ_CLUSTER_ROWS:
      gsj::cluster_rows_run(c, r.P->L, r.opts, r.rows_cb, r.user, &q.stats);
      break;
    case GSJ_REQ_DWITHIN:
      gsj::dwithin_run(c, r.P->L, r.R->L, r.d, r.same != 0, r.report_distance != 0, r.opts, r.dist_cb, r.user, &q.stats);
      break;
    case GSJ_REQ_POINTS_DWITHIN:
      gsj::points_dwithin_run(c, r.P->L, r.R->L, r.d, r.report_distance != 0, r.opts, r.dist_cb, r.user, &q.stats);
      break;
    case GSJ_REQ_INTERSECTS_BUFFERED:
      gsj::erode_run(c, r.P->L, r.R->L, r.ra, r.rb, r.same != 0, self_need(r), r.opts, r.dist_cb, r.user, &q.stats);
      break;
    case GSJ_REQ_POINTS_INTERSECTS_BUFFERED:
      gsj::points_erode_run(c, r.P->L, r.R->L, r.rb, r.opts, r.dist_cb, r.user, &q.stats);
      break;
    default:
      throw gsj::Error(GSJ_E_ARG, "unknown request kind " + std::to_string(r.kind));
  }
}

void gsj_sched::worker(int lane) {
  gsj::Context& c = *lanes[lane];
  (void)hipSetDevice(c.device);
  std::vector<Req*> group;
  for (;;) {
    Req* q = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] {
        if (queue.empty()) return stop;
        return running == 0 || budget == 0 || in_use + queue.front()->est <= budget;
      });
      if (queue.empty()) return;   // stop requested and nothing left
      q = queue.front();
      queue.pop_front();
      group.assign(1, q);
      if (batch_head(*q))
        for (auto it = queue.begin(); it != queue.end() && (int)group.size() < max_batch;) {
          if (batch_with(*q, **it) && (budget == 0 || in_use + q->est + (*it)->est <= budget)) {
            group.push_back(*it);
            in_use += (*it)->est;
            it = queue.erase(it);
          } else {
            ++it;
          }
        }
      in_use += q->est;
      running += (int)group.size();
      const Clock::time_point now = Clock::now();
      for (Req* x : group) { x->lane = lane; x->t_start = now; }
    }
    if (group.size() == 1) {
      q->rc = gsj::guarded_status([&] { run(c, *q); });
      q->msg = q->rc == GSJ_OK ? std::string() : gsj::g_last_error;
    } else {
      run_batch(c, group);
    }
    for (Req* x : group) {
      gsj::layer_release(const_cast<gsj_layer*>(x->r.P));
      gsj::layer_release(const_cast<gsj_layer*>(x->r.R));
      if (x->r.done_cb) x->r.done_cb(x->r.done_user);   // before `done`: its waiter cannot free it meanwhile
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      const Clock::time_point now = Clock::now();
      for (Req* x : group) {
        in_use -= x->est;
        running--;
        x->t_end = now;
        x->done = true;
        x->cv_done.notify_all();   // under the lock: its waiter frees it only after taking the lock back
      }
    }
    cv_work.notify_all();
  }
}

extern "C" {

int gsj_sched_create(gsj_ctx* ctx, int nlanes, uint64_t scratch_budget_bytes, gsj_sched** out) {
  return gsj::guarded_status([&] {
    GSJ_REQUIRE(ctx && out, GSJ_E_ARG, "null context or output pointer");
    GSJ_REQUIRE(nlanes >= 1 && nlanes <= 16, GSJ_E_ARG, "lanes must be in [1, 16]");
    *out = nullptr;
    auto s = std::make_unique<gsj_sched>();
    s->owner = ctx;
    s->budget = scratch_budget_bytes;
    for (int i = 0; i < nlanes; i++) {
      s->lanes.push_back(std::make_unique<gsj::Context>());
      gsj::context_init(*s->lanes.back(), ctx->c.device);
      s->lanes.back()->tune = ctx->c.tune;   // lanes launch with their parent context's shapes
    }
    for (int i = 0; i < nlanes; i++) s->workers.emplace_back([p = s.get(), i] { p->worker(i); });
    *out = s.release();
  });
}

void gsj_sched_destroy(gsj_sched* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (std::thread& t : s->workers) t.join();
  for (auto& l : s->lanes) {
    (void)hipSetDevice(l->device);
    (void)hipStreamSynchronize(l->stream);
  }
  delete s;
}

int gsj_sched_set_batch(gsj_sched* s, int max_requests) {
  return gsj::guarded_status([&] {
    GSJ_REQUIRE(s, GSJ_E_ARG, "null scheduler");
    GSJ_REQUIRE(max_requests >= 1 && max_requests <= (int)gsj::MAX_WINDOWS, GSJ_E_ARG, "batch size must be in [1, 32]");
    std::lock_guard<std::mutex> lk(s->mu);
    s->max_batch = max_requests;
  });
}

int gsj_submit(gsj_sched* s, const gsj_request* req, uint64_t* ticket) {
  return gsj::guarded_status([&] {
    GSJ_REQUIRE(s && req && ticket, GSJ_E_ARG, "null scheduler, request or ticket");
    const gsj_request& r = *req;
    GSJ_REQUIRE(r.P, GSJ_E_ARG, "request without a probe layer");
    GSJ_REQUIRE(r.kind == GSJ_REQ_CLUSTER || r.kind == GSJ_REQ_CLUSTER_ROWS || r.R, GSJ_E_ARG, "request without a build layer");
    GSJ_REQUIRE(r.P->L.ctx == &s->owner->c && (!r.R || r.R->L.ctx == &s->owner->c), GSJ_E_ARG,
                "layer belongs to another context than the scheduler's");
    GSJ_REQUIRE(!r.need_listed || (!r.need_self && (r.need_list || r.n_need == 0)), GSJ_E_ARG,
                "a listed self-pair rule takes need_list (or an empty list) and no need_self");
    auto q = std::make_unique<Req>();
    q->r = r;
    q->est = s->budget ? estimate_scratch(r) : 0;   // admission by estimate only under a budget
    q->t_submit = Clock::now();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      GSJ_REQUIRE(!s->stop, GSJ_E_ARG, "scheduler is stopping");
      const_cast<gsj_layer*>(r.P)->refs.fetch_add(1);   // released by the lane when the request finishes
      if (r.R) const_cast<gsj_layer*>(r.R)->refs.fetch_add(1);
      q->ticket = s->next_ticket++;
      *ticket = q->ticket;
      s->queue.push_back(q.get());
      s->reqs.emplace(q->ticket, std::move(q));
    }
    s->cv_work.notify_one();
  });
}

int gsj_wait(gsj_sched* s, uint64_t ticket, gsj_join_stats* stats, gsj_request_times* times) {
  std::unique_ptr<Req> q;
  int rc = gsj::guarded_status([&] {
    GSJ_REQUIRE(s, GSJ_E_ARG, "null scheduler");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    GSJ_REQUIRE(it != s->reqs.end(), GSJ_E_ARG, "unknown ticket " + std::to_string(ticket));
    Req* p = it->second.get();
    p->cv_done.wait(lk, [&] { return p->done; });
    q = std::move(it->second);
    s->reqs.erase(it);
  });
  if (rc != GSJ_OK) return rc;
  if (stats) *stats = q->stats;
  if (times) {
    times->queue_ms = ms_between(q->t_submit, q->t_start);
    times->run_ms = ms_between(q->t_start, q->t_end);
    times->total_ms = ms_between(q->t_submit, q->t_end);
    times->lane = q->lane;
    times->batch = q->batch;
  }
  gsj::g_last_error = q->msg;
  return q->rc;
}

int gsj_poll(gsj_sched* s, uint64_t ticket) {
  if (!s) return GSJ_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return GSJ_E_ARG;
  return it->second->done ? 1 : 0;
}

}  // extern "C"

// fl_server.cpp -- a serving loop for generate() requests on one GPU with
// continuous batching.
//
// A request is admitted (first come, first served) when the HBM it needs
// fits the server's budget.  A preparation worker then validates it,
// relabels and uploads its terrain and builds its lake-traversal visit order
// on a stream of its own.  Prepared terrains join a lane at the lane's next
// iteration boundary: each lane (one stream) iterates one union of
// terrains, the terrains of src/lem/generator.rs:140-210 side by side in the
// same launches, and a terrain leaves its lane as soon as it stops (its
// elevations are copied to the caller then).  So requests that arrive while
// others run start within an iteration of a lane, and their work shares the
// launches of the terrains already running.  An idle lane takes every
// prepared terrain; a busy one takes them when no lane is idle.  Lanes and
// preparation workers together use at most four streams (one lane, the rest
// preparing).  Per request the
// server records queueing time (submission to admission) and latency
// (submission to completion).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "fl_internal.h"

namespace {

using clk = std::chrono::steady_clock;

struct Request {
  uint64_t ticket = 0;
  fl_terrain_desc desc{};
  double* out = nullptr;
  uint64_t bytes = 0;
  clk::time_point submitted, started, finished;
  uint32_t iterations = 0;
  int32_t status = 1;  // 1 pending/running, then FL_OK or an error code
  std::string error;
};

struct Prepared {
  std::shared_ptr<Request> req;
  fl_terrain* terrain = nullptr;
  std::vector<uint32_t> new_of_old;
};

}  // namespace

struct fl_server {
  int device = 0;
  uint32_t n_lanes = 1, n_prep = 1;
  uint64_t budget = 0, in_use = 0, peak = 0;
  uint32_t running = 0, peak_running = 0;
  uint32_t idle_lanes = 0;
  fl_server_counts counts{};
  bool stop = false;
  uint64_t next_ticket = 1;
  std::deque<std::shared_ptr<Request>> queue;   // submitted, not yet admitted
  std::deque<Prepared> ready;                   // prepared, waiting for a lane
  uint32_t preparing = 0;                       // admitted, being prepared
  std::unordered_map<uint64_t, std::shared_ptr<Request>> requests;
  std::mutex mu;
  std::condition_variable cv_admit, cv_ready, cv_done;
  std::vector<std::thread> workers;
  std::string init_error;
};

namespace {

// HBM a request holds while it is served: its terrain as prepared plus its
// copy in a lane's union (both exist while it joins) and the visit order's
// scratch.
uint64_t estimate_bytes(const fl_terrain_desc& d) {
  uint64_t b = 0;
  if (fl_request_device_bytes(&d, &b) != FL_OK) return UINT64_MAX;
  return b + fl_terrain_arrays_bytes(&d);
}

// Complete request r (caller holds S->mu).
void complete(fl_server* S, const std::shared_ptr<Request>& r, int32_t rc, const std::string& err, uint32_t iters) {
  r->finished = clk::now();
  r->iterations = iters;
  r->status = rc;
  r->error = err;
  S->in_use -= r->bytes;
  S->running--;
}

// Requests that need no device work: an empty terrain runs the loop body
// once and finds nothing changed; max_iteration = 0 returns the initial
// elevations (generator.rs:140).
bool trivial(const Request& r, uint32_t* iters) {
  if (r.desc.n == 0) {
    *iters = r.desc.max_iteration > 0 ? 1 : 0;
    return true;
  }
  if (r.desc.max_iteration == 0) {
    std::memcpy(r.out, r.desc.init_elevation, r.desc.n * sizeof(double));
    *iters = 0;
    return true;
  }
  return false;
}

void prep_main(fl_server* S) {
  fl_ctx* ctx = nullptr;
  std::string ctx_err;
  if (fl_ctx_create(S->device, &ctx) != FL_OK) ctx_err = fl_last_error();
  for (;;) {
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      // stopping drains the queue: every submitted request completes
      S->cv_admit.wait(lk, [&] {
        if (S->queue.empty()) return S->stop;
        const auto& head = S->queue.front();
        return S->in_use == 0 || S->in_use + head->bytes <= S->budget;
      });
      if (S->queue.empty()) break;
      r = S->queue.front();
      S->queue.pop_front();
      S->in_use += r->bytes;
      S->peak = std::max(S->peak, S->in_use);
      S->running++;
      S->peak_running = std::max(S->peak_running, S->running);
      S->preparing++;
      r->started = clk::now();
    }
    int32_t rc = FL_OK;
    std::string err;
    Prepared p;
    p.req = r;
    uint32_t iters = 0;
    bool done_now = false;
    if (!ctx) {
      rc = FL_ERR_NO_DEVICE;
      err = "server worker has no device context: " + ctx_err;
    } else if (trivial(*r, &iters)) {
      done_now = true;
    } else {
      const auto t0 = clk::now();
      rc = fl_lane_prepare(ctx, &r->desc, &p.terrain, &p.new_of_old);
      if (rc != FL_OK) err = fl_last_error();
      const double ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
      std::lock_guard<std::mutex> lk(S->mu);
      S->counts.prep_ms += ms;
    }
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->preparing--;
      if (rc != FL_OK || done_now) {
        complete(S, r, rc, err, iters);
      } else {
        S->ready.push_back(std::move(p));
      }
    }
    S->cv_ready.notify_all();
    S->cv_done.notify_all();
    S->cv_admit.notify_all();
  }
  if (ctx) fl_ctx_destroy(ctx);
  S->cv_ready.notify_all();
}

void lane_main(fl_server* S) {
  fl_ctx* ctx = nullptr;
  std::string ctx_err;
  if (fl_ctx_create(S->device, &ctx) != FL_OK) ctx_err = fl_last_error();
  fl_lane* L = ctx ? fl_lane_create(ctx) : nullptr;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> mine;  // requests in this lane, by ticket
  auto fail_all = [&](int32_t rc, const std::string& err) {
    std::lock_guard<std::mutex> lk(S->mu);
    for (auto& kv : mine) complete(S, kv.second, rc, err, 0);
    mine.clear();
  };
  for (;;) {
    std::vector<Prepared> take;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      const bool idle = !L || fl_lane_idle(L);
      if (idle) {
        S->idle_lanes++;
        const auto t0 = clk::now();
        S->cv_ready.wait(lk, [&] {
          return !S->ready.empty() || (S->stop && S->queue.empty() && S->preparing == 0);
        });
        S->counts.idle_ms += std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        S->idle_lanes--;
        if (S->ready.empty()) break;  // stopping, nothing left to serve
        while (!S->ready.empty()) {
          take.push_back(std::move(S->ready.front()));
          S->ready.pop_front();
        }
      } else if (!S->ready.empty() && S->idle_lanes == 0) {
        // busy, and no idle lane will take them: join at the next boundary
        fl_lane_set_drain(L);
        if (fl_lane_quiet(L))
          while (!S->ready.empty()) {
            take.push_back(std::move(S->ready.front()));
            S->ready.pop_front();
          }
      }
    }
    if (!take.empty()) {
      if (!L) {
        std::lock_guard<std::mutex> lk(S->mu);
        for (auto& p : take) complete(S, p.req, FL_ERR_NO_DEVICE, "server lane has no device context: " + ctx_err, 0);
        S->cv_done.notify_all();
        S->cv_admit.notify_all();
        continue;
      }
      std::vector<std::pair<fl_terrain*, fl_lane_member>> join;
      for (auto& p : take) {
        fl_lane_member m;
        m.tag = p.req->ticket;
        m.cap = p.req->desc.max_iteration;
        m.new_of_old = std::move(p.new_of_old);
        join.push_back({p.terrain, std::move(m)});
        mine[p.req->ticket] = p.req;
      }
      {
        std::lock_guard<std::mutex> lk(S->mu);
        (fl_lane_idle(L) ? S->counts.joins_idle : S->counts.joins_busy) += join.size();
      }
      const auto t0 = clk::now();
      const int32_t jrc = fl_lane_join(L, join);
      {
        std::lock_guard<std::mutex> lk(S->mu);
        S->counts.join_ms += std::chrono::duration<double, std::milli>(clk::now() - t0).count();
      }
      if (jrc != FL_OK) {
        fail_all(FL_ERR_HIP, fl_last_error());
        S->cv_done.notify_all();
        S->cv_admit.notify_all();
        continue;
      }
    }
    if (!L || fl_lane_idle(L)) continue;
    {
      const uint32_t k = fl_lane_terrains(L);
      std::lock_guard<std::mutex> lk(S->mu);
      S->counts.max_lane_terrains = std::max(S->counts.max_lane_terrains, k);
    }
    std::vector<fl_lane_member> stopped;
    const auto t0 = clk::now();
    const int32_t src = fl_lane_step(L, stopped);
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->counts.step_ms += std::chrono::duration<double, std::milli>(clk::now() - t0).count();
      S->counts.steps++;
    }
    if (src != FL_OK) {
      const std::string err = fl_last_error();
      fl_lane_destroy(L);
      L = fl_lane_create(ctx);
      fail_all(FL_ERR_HIP, err);
      S->cv_done.notify_all();
      S->cv_admit.notify_all();
      continue;
    }
    if (!stopped.empty()) {
      std::lock_guard<std::mutex> lk(S->mu);
      for (auto& m : stopped) {
        auto it = mine.find(m.tag);
        if (it == mine.end()) continue;
        std::shared_ptr<Request> r = it->second;
        std::memcpy(r->out, m.result.data(), m.result.size() * sizeof(double));
        complete(S, r, FL_OK, "", m.iterations);
        mine.erase(it);
      }
    }
    if (!stopped.empty()) {
      S->cv_done.notify_all();
      S->cv_admit.notify_all();
    }
  }
  if (L) fl_lane_destroy(L);
  if (!mine.empty()) fail_all(FL_ERR_HIP, "server stopped with requests in a lane");
  if (ctx) fl_ctx_destroy(ctx);
}

}  // namespace

extern "C" int32_t fl_server_create(int32_t device, uint32_t slots, uint64_t hbm_budget_bytes, fl_server** out) {
  if (!out || slots == 0) {
    fl_internal_set_error("null output or zero slots");
    return FL_ERR_INVALID;
  }
  auto* S = new fl_server();
  S->device = device;
  S->budget = hbm_budget_bytes;
  // at most four streams: one lane iterates every prepared terrain in one
  // union (the most sharing of launches); preparation (host validation and
  // relabelling, upload, visit order) costs more per site than iterating, so
  // the other streams prepare
  const uint32_t streams = std::min<uint32_t>(std::max<uint32_t>(slots, 2), 4);
  S->n_lanes = 1;
  S->n_prep = streams - S->n_lanes;
  for (uint32_t i = 0; i < S->n_prep; i++) S->workers.emplace_back(prep_main, S);
  for (uint32_t i = 0; i < S->n_lanes; i++) S->workers.emplace_back(lane_main, S);
  *out = S;
  return FL_OK;
}

extern "C" int32_t fl_server_submit(fl_server* S, const fl_terrain_desc* desc, double* elevations_out,
                                    uint64_t* ticket_out) {
  if (!S || !desc || !ticket_out || (!elevations_out && desc->n)) return FL_ERR_INVALID;
  auto r = std::make_shared<Request>();
  r->desc = *desc;
  r->out = elevations_out;
  r->bytes = estimate_bytes(*desc);
  r->submitted = clk::now();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (r->bytes > S->budget) {
      fl_internal_set_error("request needs more HBM than the server's budget");
      return FL_ERR_BUDGET;
    }
    r->ticket = S->next_ticket++;
    S->requests[r->ticket] = r;
    S->queue.push_back(r);
  }
  S->cv_admit.notify_all();
  *ticket_out = r->ticket;
  return FL_OK;
}

extern "C" int32_t fl_server_wait(fl_server* S, uint64_t ticket, uint32_t* iterations_out, double* queue_ms,
                                  double* latency_ms) {
  if (!S) return FL_ERR_INVALID;
  std::shared_ptr<Request> r;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->requests.find(ticket);
  if (it == S->requests.end()) {
    fl_internal_set_error("unknown server ticket");
    return FL_ERR_INVALID;
  }
  r = it->second;
  S->cv_done.wait(lk, [&] { return r->status != 1; });
  S->requests.erase(ticket);
  if (iterations_out) *iterations_out = r->iterations;
  if (queue_ms) *queue_ms = std::chrono::duration<double, std::milli>(r->started - r->submitted).count();
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->finished - r->submitted).count();
  if (r->status != FL_OK) fl_internal_set_error(r->error.c_str());
  return r->status;
}

extern "C" int32_t fl_server_stats(fl_server* S, uint64_t* peak_bytes, uint32_t* peak_running) {
  if (!S) return FL_ERR_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  if (peak_bytes) *peak_bytes = S->peak;
  if (peak_running) *peak_running = S->peak_running;
  return FL_OK;
}

extern "C" int32_t fl_server_destroy(fl_server* S) {
  if (!S) return FL_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_admit.notify_all();
  S->cv_ready.notify_all();
  for (auto& t : S->workers) t.join();
  delete S;
  return FL_OK;
}

extern "C" int32_t fl_server_counters(fl_server* S, fl_server_counts* out) {
  if (!S || !out) return FL_ERR_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  *out = S->counts;
  out->lanes = S->n_lanes;
  out->prep_workers = S->n_prep;
  return FL_OK;
}

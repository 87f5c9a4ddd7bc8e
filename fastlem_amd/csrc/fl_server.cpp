// fl_server.cpp -- a serving loop for generate() requests on one GPU.
//
// Continuous batching: `slots` worker threads each own a context (a HIP
// stream).  A request is admitted to a free slot when the HBM it will need
// fits in the server's budget, and runs the device solve of
// src/lem/generator.rs:140-210 on that slot's stream, so the kernels of
// concurrently admitted requests overlap on the device.  A finished request
// frees its slot and memory at once and the next queued request is admitted
// (first come, first served; the queue head waits for memory).  Per request
// the server records queueing and service time.
#include <chrono>
#include <condition_variable>
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

}  // namespace

struct fl_server {
  int device = 0;
  uint64_t budget = 0, in_use = 0, peak = 0;
  uint32_t running = 0, peak_running = 0;
  bool stop = false;
  uint64_t next_ticket = 1;
  std::deque<std::shared_ptr<Request>> queue;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> requests;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::vector<std::thread> workers;
  std::string init_error;
};

namespace {

// HBM a request holds at its peak: the terrain's device arrays plus the
// visit order's scratch, from the sizes the allocations themselves use.
uint64_t estimate_bytes(const fl_terrain_desc& d) {
  uint64_t b = 0;
  return fl_request_device_bytes(&d, &b) == FL_OK ? b : UINT64_MAX;
}

void worker_main(fl_server* S) {
  fl_ctx* ctx = nullptr;
  if (fl_ctx_create(S->device, &ctx) != FL_OK) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->init_error = fl_last_error();
  }
  for (;;) {
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      // stopping drains the queue: every submitted request completes
      S->cv_work.wait(lk, [&] {
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
      r->started = clk::now();
    }
    int32_t rc;
    uint32_t iters = 0;
    std::string err;
    if (!ctx) {
      rc = FL_ERR_NO_DEVICE;
      err = "server slot has no device context";
    } else {
      rc = fl_generate_f64(ctx, &r->desc, r->out, &iters);
      if (rc != FL_OK) err = fl_last_error();
    }
    {
      std::lock_guard<std::mutex> lk(S->mu);
      r->finished = clk::now();
      r->iterations = iters;
      r->status = rc;
      r->error = err;
      S->in_use -= r->bytes;
      S->running--;
    }
    S->cv_done.notify_all();
    S->cv_work.notify_all();
  }
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
  for (uint32_t i = 0; i < slots; i++) S->workers.emplace_back(worker_main, S);
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
  *ticket_out = r->ticket;
  S->cv_work.notify_all();
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
  S->cv_work.notify_all();
  for (auto& t : S->workers) t.join();
  delete S;
  return FL_OK;
}

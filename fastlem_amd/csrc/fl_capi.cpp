// fl_capi.cpp -- C ABI of the device path: contexts, terrains resident in
// HBM, and the outer iteration loop of generate() (src/lem/generator.rs:
// 140-210) driven from the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <future>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_set>
#include <vector>

#include "fl_internal.h"
#include "fl_kernels.h"
#include "fl_tuning_table.h"

namespace {

thread_local std::string g_err;

int32_t set_err(int32_t code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(call)                                                                               \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return set_err(e_ == hipErrorOutOfMemory ? FL_ERR_OUT_OF_MEMORY : FL_ERR_HIP, "%s: %s (%s:%d)", \
                     #call, hipGetErrorString(e_), __FILE__, __LINE__);                            \
  } while (0)

#define FLCHK(call)          \
  do {                       \
    int32_t r_ = (call);     \
    if (r_ != FL_OK) return r_; \
  } while (0)

}  // namespace

struct fl_ctx;
namespace {
struct ProfScope {
  fl_profiler* saved;
  explicit ProfScope(fl_profiler* p) : saved(fl_current_profiler) { fl_current_profiler = p; }
  ~ProfScope() { fl_current_profiler = saved; }
};
}  // namespace

struct fl_ctx {
  int device = 0;
  hipMemPool_t pool = nullptr;           // the library's memory pool of the device
  uint32_t compact_min_idle = 1u << 16;  // see solve_union (the call's tuning; set per call)
  // tuning: `tune` holds the caller's settings, tune_set which of them apply
  // (1 all: fl_ctx_set_tuning; 2 pipelining; 4 compaction)
  fl_tuning tune{};
  uint32_t tune_set = 0;
  int pipelined = 1;                     // one_iteration_fast after the first iteration (2: minimal
                                         // speculation, most iterations redone)
  bool graphs = false;                   // pipelined iterations replayed from captured graphs
  bool forbid_graphs = false;            // no graph replay whatever the tuning (allocation-free calls)
  bool recompute_all = false;            // every basin in every iteration (no dirty-basin selection)
  uint32_t rehint_rounds = 0;            // sweep rounds beyond which an iteration re-hints (0: 2 log2 n + 8)
  fl_profiler* prof = nullptr;  // non-null while profiling is enabled
  hipStream_t stream = nullptr;
  uint32_t* host_ctrl = nullptr;  // pinned readback of the control block
  // pinned readbacks and completion events of pipelined iterations in flight
  uint32_t* host_slot[2] = {nullptr, nullptr};
  hipEvent_t slot_ev[2] = {nullptr, nullptr};
  std::vector<uint32_t> host_boff;
  // pinned per-terrain slots of batches (changed flags, flags of the two
  // flights, active mask), shared by the unions solved on this context (one
  // at a time); grown by new blocks, the old ones kept until the context goes
  uint8_t* pin = nullptr;
  uint32_t pin_cap = 0;
  std::vector<void*> pin_blocks;
};

struct fl_terrain {
  fl_ctx* ctx = nullptr;
  fl_dev_terrain t;
  fl_dev_work w;
  uint32_t max_iteration = FL_UNBOUNDED_ITERATIONS;
  uint32_t prev_nb = 0;      // bucket entries the previous decomposition used (cleared next)
  uint32_t jump_rounds = 4;  // pointer-jumping rounds of the previous iteration
  uint32_t last_rounds = 0;  // sweep rounds of the previous iteration
  std::vector<uint32_t> last_boff;  // its bucket offsets (boff[nb+1], bnoff[nb+1]): launch-size estimates
  uint32_t tail_split = 0;          // first round of the pipelined iterations' tail launches
  uint32_t tail_units = FL_TAIL_WAVES;  // rounds of at most this many fold units go to the tail launches
  // captured pipelined iterations, one per elevation-buffer parity, for the
  // (jump rounds, sweep rounds) they were captured with
  // elevation buffers: an iteration reads hb[i] and writes hb[(i+1) % 3], so
  // an iteration in flight never overwrites the input of the one before it
  double* hb[3] = {nullptr, nullptr, nullptr};
  // the tree of the iteration that wrote hb[i] (receivers, distances, basin
  // roots) is nxb[i], dxb[i], rtb[i]: an iteration reading hb[i] compares
  // its tree with them; prev_valid: they hold the last iteration's tree (not
  // so before a solve's first iteration or a union's first after a gather)
  uint32_t* nxb[3] = {nullptr, nullptr, nullptr};
  double* dxb[3] = {nullptr, nullptr, nullptr};
  uint32_t* rtb[3] = {nullptr, nullptr, nullptr};
  uint8_t* chgb[3] = {nullptr, nullptr, nullptr};  // per-site "elevation changed" of that iteration
  bool prev_valid = false;
  bool after_lake = true;  // the last iteration joined lakes (its tree is not raw receivers)
  // w.A holds no drainage area of a tree like the next one (a solve's
  // start, new terrains in the union, the iteration after lake joins, whose
  // drainage changes wholesale where lakes were): the next iteration runs
  // synchronously and takes subtree sizes as its heavy-child hint
  bool hint_stale = false;
  hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
  uint32_t gJ[3] = {0, 0, 0}, gR[3] = {0, 0, 0}, gS[3] = {0, 0, 0};
  int gF[3] = {0, 0, 0};
  // launch tuning a graph was captured with (workgroup size, long-chain
  // cutoff): another tuning captures afresh, so that it is the one timed
  uint32_t gB[3] = {0, 0, 0}, gL[3] = {0, 0, 0};
  uint32_t* h_tslot[2] = {nullptr, nullptr};  // pinned per-terrain changed flags of iterations in flight
  uint8_t* h_mask = nullptr;                  // pinned staging of the active mask
  bool graphs_failed = false;
  bool have_order = false;
  // statistics of the last solve
  uint32_t stat_iterations = 0, stat_lake_iterations = 0, stat_syncs = 0, stat_order_fast = 0,
           stat_compactions = 0, stat_redone = 0;
  double stat_order_ms = 0.0, stat_solve_ms = 0.0;
  double stat_order_parts[3] = {0.0, 0.0, 0.0};  // outlet phase, spanning forest, ranks
  double stat_create[3] = {0.0, 0.0, 0.0};       // validation, relabelling, allocation and upload
  uint32_t stat_create_device = 0, stat_create_clusters = 0;
  uint64_t stat_recomputed = 0;  // sites of the recomputed basins, over the solve's iterations
  uint32_t stat_rounds_first = 0, stat_rounds_max = 0;  // sweep rounds: the solve's first iteration, any
  uint32_t stat_rehinted = 0;  // iterations decomposed again with subtree sizes (too deep)
  uint64_t bytes = 0;
  // where its device arrays come from (the device pool unless a caller
  // supplies a region); region buffers are never freed one by one
  fl_arena arena;
  // outlets in resolution order (relabelled), and each outlet's row as the
  // outlet phase of the visit order needs it: its entry count o_deg[q] and
  // its neighbouring outlets o_nb[o_ptr[q] .. o_ptr[q+1]) (indices within
  // its terrain's outlet list)
  std::vector<uint32_t> h_outlets, o_deg, o_ptr, o_nb;
  // locality relabelling: device site u holds caller site old_of_new[u]
  std::vector<uint32_t> new_of_old, old_of_new;
  std::vector<void*> allocations;
  // terrains packed into this object (one for fl_terrain_create); terrain g
  // owns sites [site_off[g], site_off[g+1]) and outlets [outlet_off[g], ...)
  uint32_t nterr = 1;
  std::vector<uint32_t> site_off, outlet_off, entry_off, t_max_it, t_iters;
  std::vector<uint8_t> t_active;
  std::vector<uint8_t> t_asym;    // [nterr] some edge's two lengths differ (exact visit order)
  uint32_t* d_tid = nullptr;      // [n] terrain of each site (batches)
  uint8_t* d_active = nullptr;    // [nterr] terrains still iterating
  uint32_t* d_tchanged = nullptr; // [nterr] per-terrain changed flags
  uint32_t* h_tchanged = nullptr; // pinned copy (the context's slots)
};

namespace {

// The buffers of an iteration reading slot `in`: elevations hb[in] ->
// hb[in+1], and its tree nxb/dxb/rtb[in+1] compared with nxb/dxb/rtb[in].
void set_slots(fl_terrain* T, int in) {
  const int out = (in + 1) % 3;
  fl_dev_work& w = T->w;
  w.h = T->hb[in];
  w.h2 = T->hb[out];
  w.next = T->nxb[out];
  w.dnext = T->dxb[out];
  w.next_prev = T->nxb[in];
  w.dnext_prev = T->dxb[in];
  w.root_prev = T->rtb[in];
  w.root = T->rtb[out];
  w.chg_prev = T->chgb[in];
  w.chg = T->chgb[out];
}

template <typename T>
int32_t dev_alloc(fl_terrain* T_, T** p, uint64_t count) {
  uint64_t bytes = fl_round256(std::max<uint64_t>(count, 1) * sizeof(T));
  // pool allocations are stream-ordered on the terrain's context: no
  // device-wide synchronisation when terrains come and go while others run
  // (serving, batch compaction)
  hipError_t e = fl_arena_alloc(&T_->arena, T_->ctx ? T_->ctx->stream : nullptr, (void**)p, bytes);
  if (e != hipSuccess) {
    if (T_->arena.kind == fl_arena::REGION)
      return set_err(FL_ERR_TOO_LARGE, "device region of %llu bytes exhausted (%llu more needed)",
                     (unsigned long long)T_->arena.size, (unsigned long long)bytes);
    return set_err(e == hipErrorOutOfMemory ? FL_ERR_OUT_OF_MEMORY : FL_ERR_HIP, "hipMalloc(%llu): %s",
                   (unsigned long long)bytes, hipGetErrorString(e));
  }
  if (T_->arena.kind == fl_arena::POOL) T_->allocations.push_back((void*)*p);
  T_->bytes += bytes;
  return FL_OK;
}

// f(begin, end) over [0, count) in contiguous ranges on up to
// fl_host_threads() threads, none smaller than min_per_thread; runs on the
// calling thread when one range suffices.
template <class F>
void host_parallel(uint64_t count, uint64_t min_per_thread, F f) {
  uint64_t th = std::min<uint64_t>(fl_host_threads(), std::max<uint64_t>(count / min_per_thread, 1));
  if (th <= 1) {
    f(0, count);
    return;
  }
  std::vector<std::thread> pool;
  for (uint64_t t = 0; t < th; t++) pool.emplace_back([&, t] { f(count * t / th, count * (t + 1) / th); });
  for (auto& x : pool) x.join();
}

// f(g) for every g in [0, count), largest items first (by size(g)), on up to
// fl_host_threads() threads.
template <class S, class F>
void host_each(uint32_t count, S size, F f) {
  std::vector<uint32_t> order(count);
  for (uint32_t g = 0; g < count; g++) order[g] = g;
  std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return size(x) > size(y); });
  std::atomic<uint32_t> next{0};
  auto worker = [&] {
    for (uint32_t q; (q = next.fetch_add(1)) < count;) f(order[q]);
  };
  uint32_t th = std::min<uint32_t>(fl_host_threads(), count);
  if (th <= 1) {
    worker();
    return;
  }
  std::vector<std::thread> pool;
  for (uint32_t t = 0; t < th; t++) pool.emplace_back(worker);
  for (auto& x : pool) x.join();
}

// The first failing row of a per-row check, over host threads: the same row
// (and entry) a sequential scan would report.
struct RowError {
  uint64_t row = UINT64_MAX;
  int kind = 0;  // 0 none; 1 index out of range; 2 self-loop; 3 duplicate; 4 not symmetric
  uint32_t k = 0, j = 0;
};
template <class C>
RowError first_row_error(uint64_t n, C check_row) {
  std::mutex mu;
  RowError best;
  host_parallel(n, 65536, [&](uint64_t a, uint64_t b) {
    for (uint64_t i = a; i < b; i++) {
      RowError e = check_row(i);
      if (e.kind) {
        std::lock_guard<std::mutex> g(mu);
        if (i < best.row) best = e;
        return;
      }
    }
  });
  return best;
}

// Checks of a description that need no look at its rows: null arrays,
// sizes, and row offsets that start at 0, end at nnz and never decrease.
// *max_row receives the longest row.
int32_t check_header(const fl_terrain_desc* d, uint32_t* max_row) {
  *max_row = 0;
  if (!d) return set_err(FL_ERR_INVALID, "null terrain description");
  if (d->n > FL_MAX_SITES) return set_err(FL_ERR_TOO_LARGE, "n=%llu exceeds 2^31-1 sites", (unsigned long long)d->n);
  if (d->nnz >= 0xffffffffull) return set_err(FL_ERR_TOO_LARGE, "nnz exceeds 2^32-2 entries");
  if (d->n == 0) return FL_OK;
  if (!d->row_ptr || (d->nnz && (!d->col || !d->length)) || !d->area || !d->erodibility || !d->uplift_rate ||
      !d->init_elevation)
    return set_err(FL_ERR_INVALID, "null array in terrain description");
  if (d->n_outlets && !d->outlets) return set_err(FL_ERR_INVALID, "null outlets");
  if (d->has_max_slope && !d->max_slope) return set_err(FL_ERR_INVALID, "has_max_slope without max_slope");
  const uint64_t n = d->n;
  const uint32_t* rp = d->row_ptr;
  if (rp[0] != 0 || rp[n] != d->nnz) return set_err(FL_ERR_INVALID, "row_ptr must span [0, nnz]");
  std::mutex mu;
  uint64_t first_dec = UINT64_MAX;
  uint32_t longest = 0;
  host_parallel(n, 1 << 18, [&](uint64_t a, uint64_t b) {
    uint32_t m = 0;
    uint64_t dec = UINT64_MAX;
    for (uint64_t i = a; i < b; i++) {
      if (rp[i + 1] < rp[i]) {
        dec = i;
        break;
      }
      m = std::max(m, rp[i + 1] - rp[i]);
    }
    std::lock_guard<std::mutex> g(mu);
    first_dec = std::min(first_dec, dec);
    longest = std::max(longest, m);
  });
  if (first_dec != UINT64_MAX)
    return set_err(FL_ERR_INVALID, "row_ptr decreases at %llu", (unsigned long long)first_dec);
  *max_row = longest;
  return FL_OK;
}

// The rows of a description whose header passed: neighbour indices,
// self-loops and duplicates (the reference's BFS would visit such a site
// twice), then symmetry (terrain-graph is undirected), each reported at the
// first failing row.  *asym is set when some edge's two entries carry
// different length bits (legal input: the reference keys each heap push by
// the popped site's own entry, stream_tree.rs:231-235, which the exact
// traversal reproduces; the spanning-forest form needs one length per edge).
int32_t check_rows(const fl_terrain_desc* d, bool* asym) {
  *asym = false;
  const uint64_t n = d->n;
  if (n == 0) return FL_OK;
  const uint32_t* rp = d->row_ptr;
  const uint32_t* col = d->col;
  RowError e = first_row_error(n, [&](uint64_t i) {
    RowError r;
    const uint32_t b = rp[i], deg = rp[i + 1] - b;
    std::unordered_set<uint32_t> seen;  // rows longer than 32 entries only
    for (uint32_t k = b; k < b + deg; k++) {
      const uint32_t j = col[k];
      r.row = i;
      r.k = k;
      r.j = j;
      if (j >= n) {
        r.kind = 1;
        return r;
      }
      if (j == i) {
        r.kind = 2;
        return r;
      }
      bool dup = false;
      if (deg <= 32) {
        for (uint32_t q = b; q < k && !dup; q++) dup = col[q] == j;
      } else {
        dup = !seen.insert(j).second;
      }
      if (dup) {
        r.kind = 3;
        return r;
      }
    }
    return RowError{};
  });
  if (e.kind == 1) return set_err(FL_ERR_INVALID, "col[%u]=%u out of range", e.k, e.j);
  if (e.kind == 2) return set_err(FL_ERR_INVALID, "self-loop at site %llu", (unsigned long long)e.row);
  if (e.kind == 3) return set_err(FL_ERR_INVALID, "duplicate edge %llu-%u", (unsigned long long)e.row, e.j);
  std::atomic<bool> len_differs{false};
  e = first_row_error(n, [&](uint64_t i) {
    bool differs = false;
    for (uint32_t k = rp[i]; k < rp[i + 1]; k++) {
      const uint32_t j = col[k];
      bool found = false;
      for (uint32_t q = rp[j]; q < rp[j + 1] && !found; q++) {
        if (col[q] != i) continue;
        found = true;
        uint64_t a, b;
        std::memcpy(&a, d->length + k, 8);
        std::memcpy(&b, d->length + q, 8);
        differs |= a != b;
      }
      if (!found) {
        RowError r;
        r.row = i;
        r.kind = 4;
        r.k = k;
        r.j = j;
        return r;
      }
    }
    if (differs) len_differs.store(true, std::memory_order_relaxed);
    return RowError{};
  });
  *asym = len_differs.load();
  if (e.kind == 4) return set_err(FL_ERR_INVALID, "edge %llu-%u is not symmetric", (unsigned long long)e.row, e.j);
  return FL_OK;
}

// The message check_rows gives for its first failing entry k (terrain-local)
// of the first kind (1..3: index, self-loop, duplicate) or symmetry (4).
int32_t row_error(const fl_terrain_desc* d, uint32_t k, bool symmetry) {
  const uint32_t* rp = d->row_ptr;
  const uint64_t row = (uint64_t)(std::upper_bound(rp, rp + d->n + 1, k) - rp) - 1;
  const uint32_t j = d->col[k];
  if (symmetry) return set_err(FL_ERR_INVALID, "edge %llu-%u is not symmetric", (unsigned long long)row, j);
  if (j >= d->n) return set_err(FL_ERR_INVALID, "col[%u]=%u out of range", k, j);
  if (j == row) return set_err(FL_ERR_INVALID, "self-loop at site %llu", (unsigned long long)row);
  return set_err(FL_ERR_INVALID, "duplicate edge %llu-%u", (unsigned long long)row, j);
}

// Outlets in range and listed once; slot[i] = q + 1 for outlet q at site i
// (slot: n zeroed entries).
int32_t check_outlets(const fl_terrain_desc* d, uint32_t* slot) {
  for (uint64_t q = 0; q < d->n_outlets; q++) {
    const uint32_t o = d->outlets[q];
    if (o >= d->n) return set_err(FL_ERR_INVALID, "outlet %u out of range", o);
    if (slot[o]) return set_err(FL_ERR_INVALID, "duplicate outlet %u", o);
    slot[o] = (uint32_t)q + 1;
  }
  return FL_OK;
}

// The tuning a call on `sites` sites (a batch when terrains > 1) takes: the
// context's fixed tuning, else the table's row for the class, with the
// context's explicit pipelining and compaction settings over it.
fl_tuning effective_tuning(const fl_ctx* c, uint64_t sites, uint32_t terrains) {
  if (c->tune_set & 1) return c->tune;
  fl_tuning t{};
  const uint32_t batch = terrains > 1 ? 1 : 0;
  for (const fl_tuning_row& r : fl_tuning_table)
    if (r.batch == batch && sites <= r.max_sites) {
      t.site_block = r.site_block;
      t.long_chain = r.long_chain;
      t.tail_units = r.tail_units;
      t.pipelining = r.pipelining;
      t.compact_min_idle = r.compact_min_idle;
      break;
    }
  if (!t.site_block) t = fl_tuning{256, 6, FL_TAIL_WAVES, 1, 1u << 16, {0, 0, 0}};
  if (c->tune_set & 2) t.pipelining = c->tune.pipelining;
  if (c->tune_set & 4) t.compact_min_idle = c->tune.compact_min_idle;
  return t;
}

bool valid_tuning(const fl_tuning& t) {
  return t.site_block >= 64 && t.site_block <= 1024 && t.site_block % 64 == 0 && t.long_chain >= 1 &&
         (t.pipelining == 0 || t.pipelining == 1 || t.pipelining == 3);
}

// A call's tuning applied to the context, the calling thread and terrain T
// (and the unions derived from it), restored when the call returns.
struct TuneScope {
  fl_ctx* c;
  uint32_t block, compact;
  int pipelined;
  bool graphs;
  TuneScope(fl_ctx* c_, fl_terrain* T, const fl_tuning& t)
      : c(c_), block(fl_site_block), compact(c_->compact_min_idle), pipelined(c_->pipelined), graphs(c_->graphs) {
    fl_site_block = t.site_block;
    c->pipelined = t.pipelining == 3 ? 1 : (int)t.pipelining;
    c->graphs = t.pipelining == 3 && !c->forbid_graphs;
    c->compact_min_idle = t.compact_min_idle;
    if (T) {
      T->w.long_min = t.long_chain;
      T->tail_units = t.tail_units;
    }
  }
  ~TuneScope() {
    fl_site_block = block;
    c->compact_min_idle = compact;
    c->pipelined = pipelined;
    c->graphs = graphs;
  }
};

int32_t read_ctrl(fl_ctx* c, fl_terrain* T) {
  HIPCHK(hipMemcpyAsync(c->host_ctrl, T->w.ctrl, FL_CTRL_WORDS * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  T->stat_syncs++;
  return FL_OK;
}

inline int host_flag(const fl_ctx* c, int f) { return (int)c->host_ctrl[f]; }
constexpr uint32_t FL_NONE_ROUND = 0xffffffffu;
inline uint32_t host_counter(const fl_ctx* c, int k) { return c->host_ctrl[FL_NUM_FLAGS + k]; }

int32_t ensure_visit_order(fl_ctx* c, fl_terrain* T) {
  if (T->have_order) return FL_OK;
  auto t0 = std::chrono::steady_clock::now();
  const uint32_t n = T->t.n, G = T->nterr;
  const size_t K = T->h_outlets.size();
  hipStream_t s = c->stream;
  fl_launch_fill_u32(s, T->t.rank, n, 0xffffffffu);
  // per terrain: 0 fast form (spanning forest), 1 exact heap simulation
  // (repeated, non-positive or asymmetric lengths), 2 nothing to rank (no
  // outlets: no site is ever reached)
  std::vector<uint8_t> how(G, 0);
  uint32_t n_fast = 0, n_exact = 0;
  std::string err;
  if (K > 0) {
    std::vector<uint8_t> bad;
    int rc = fl_device_len_check(s, n, T->t.row_ptr, T->t.col, T->t.len, T->d_tid, G, T->entry_off.data(), bad,
                                 &err, &T->arena);
    if (rc != FL_OK) return set_err(rc, "length check: %s", err.c_str());
    for (uint32_t g = 0; g < G; g++) {
      how[g] = T->outlet_off[g + 1] == T->outlet_off[g] ? 2 : (bad[g] || T->t_asym[g]) ? 1 : 0;
      n_fast += how[g] == 0;
      n_exact += how[g] == 1;
    }
  }
  // the exact terrains' graphs (relabelled) from the device, at their own
  // offsets of union-sized host arrays
  std::unique_ptr<uint32_t[]> x_rp, x_col;
  std::unique_ptr<double[]> x_len;
  if (n_exact) {
    x_rp.reset(new uint32_t[(size_t)n + 1]);
    x_col.reset(new uint32_t[std::max<uint32_t>(T->t.nnz, 1)]);
    x_len.reset(new double[std::max<uint32_t>(T->t.nnz, 1)]);
    for (uint32_t g = 0; g < G; g++) {
      if (how[g] != 1) continue;
      const uint64_t s0 = T->site_off[g], s1 = T->site_off[g + 1], e0 = T->entry_off[g], e1 = T->entry_off[g + 1];
      HIPCHK(hipMemcpyAsync(x_rp.get() + s0, T->t.row_ptr + s0, (s1 - s0 + 1) * 4, hipMemcpyDeviceToHost, s));
      if (e1 > e0) {
        HIPCHK(hipMemcpyAsync(x_col.get() + e0, T->t.col + e0, (e1 - e0) * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(x_len.get() + e0, T->t.len + e0, (e1 - e0) * 8, hipMemcpyDeviceToHost, s));
      }
    }
    HIPCHK(hipStreamSynchronize(s));
  }
  // host threads: the outlet phase of the fast terrains (exact std heap tie
  // handling, from each outlet's row), and the whole traversal of the exact
  // ones
  std::vector<uint32_t> first(K), exact_rank(n_exact ? n : 0);
  std::future<int> outlet_phase = std::async(std::launch::async, [&]() {
    auto p0 = std::chrono::steady_clock::now();
    std::vector<uint32_t> pops;
    for (uint32_t g = 0; g < G; g++) {
      if (how[g] != 0) continue;
      const uint32_t o0 = T->outlet_off[g], k = T->outlet_off[g + 1] - o0;
      pops.resize(k);
      int rc = fl_outlet_pop_order_compact(k, T->o_deg.data() + o0, T->o_ptr.data() + o0, T->o_nb.data(),
                                           pops.data());
      if (rc != FL_OK) return rc;
      for (uint32_t q = 0; q < k; q++) first[o0 + q] = T->h_outlets[o0 + pops[q]];
    }
    T->stat_order_parts[0] =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - p0).count();
    return (int)FL_OK;
  });
  std::future<int> exact_phase = std::async(std::launch::async, [&]() {
    if (!n_exact) return (int)FL_OK;
    std::vector<int> rcs(G, FL_OK);
    host_each(
        G, [&](uint32_t g) { return how[g] == 1 ? T->site_off[g + 1] - T->site_off[g] : 0u; },
        [&](uint32_t g) {
          if (how[g] != 1) return;
          uint32_t s0 = T->site_off[g], o0 = T->outlet_off[g];
          rcs[g] = fl_lake_visit_order_range(s0, T->site_off[g + 1] - s0, x_rp.get(), x_col.get(), x_len.get(),
                                             T->h_outlets.data() + o0, T->outlet_off[g + 1] - o0, s0,
                                             exact_rank.data());
        });
    for (int rc : rcs)
      if (rc != FL_OK) return rc;
    return (int)FL_OK;
  });
  // the device: spanning forest and Prim ranks of the fast terrains, with
  // each terrain's outlets merged into one component (labelled by its first
  // listed outlet) and the other terrains merged whole (no forest edges)
  int rc_dev = FL_OK;
  if (n_fast) {
    auto p1 = std::chrono::steady_clock::now();
    std::vector<uint32_t> init(n), base(G);
    host_parallel(n, 1 << 18, [&](uint64_t a, uint64_t b) {
      for (uint64_t v = a; v < b; v++) init[v] = (uint32_t)v;
    });
    for (uint32_t g = 0; g < G; g++) {
      const uint32_t s0 = T->site_off[g], s1 = T->site_off[g + 1];
      base[g] = s0 + (T->outlet_off[g + 1] - T->outlet_off[g]);
      if (how[g] == 0) {
        for (uint32_t q = T->outlet_off[g]; q < T->outlet_off[g + 1]; q++)
          init[T->h_outlets[q]] = T->h_outlets[T->outlet_off[g]];
      } else {
        for (uint32_t v = s0; v < s1; v++) init[v] = s0;
      }
    }
    fl_visit_stats vs;
    rc_dev = fl_device_ranks(s, n, T->t.row_ptr, T->t.col, T->t.len, T->t.is_outlet, T->d_tid, init, base,
                             T->t.rank, &vs, &err, &T->arena);
    // the forest part includes the set-up of its labels on the host
    T->stat_order_parts[1] =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - p1).count() - vs.ranks_ms;
    T->stat_order_parts[2] = vs.ranks_ms;
  }
  int rc_outlets = outlet_phase.get();
  int rc_exact = exact_phase.get();
  if (rc_dev != FL_OK) return set_err(rc_dev, "visit order on the device: %s", err.c_str());
  if (rc_outlets != FL_OK) return set_err(rc_outlets, "outlet phase failed");
  if (rc_exact == FL_ERR_NAN_LENGTH) return set_err(FL_ERR_NAN_LENGTH, "NaN edge length reached by the lake traversal");
  if (rc_exact != FL_OK) return set_err(rc_exact, "visit order failed");
  // the outlets of the fast terrains in pop order, the exact terrains whole
  std::vector<uint32_t> osite, orank;
  osite.reserve(K);
  orank.reserve(K);
  for (uint32_t g = 0; g < G; g++) {
    if (how[g] != 0) continue;
    for (uint32_t q = T->outlet_off[g]; q < T->outlet_off[g + 1]; q++) {
      osite.push_back(first[q]);
      orank.push_back(T->site_off[g] + (q - T->outlet_off[g]));
    }
  }
  if (!osite.empty()) {
    int rc = fl_device_scatter_u32(s, (uint32_t)osite.size(), osite.data(), orank.data(), T->t.rank, &err, &T->arena);
    if (rc != FL_OK) return set_err(rc, "outlet ranks: %s", err.c_str());
  }
  for (uint32_t g = 0; g < G; g++) {
    if (how[g] != 1) continue;
    const uint32_t s0 = T->site_off[g], ng = T->site_off[g + 1] - s0;
    HIPCHK(hipMemcpyAsync(T->t.rank + s0, exact_rank.data() + s0, (uint64_t)ng * 4, hipMemcpyHostToDevice, s));
  }
  fl_device_order(s, n, T->t.rank, T->t.order);
  HIPCHK(hipStreamSynchronize(s));  // the host arrays above must outlive the copies
  T->have_order = true;
  T->stat_order_fast = (n_fast > 0 && n_exact == 0) ? 1 : 0;
  T->stat_order_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return FL_OK;
}

// Chain-end labels of the receiver forest by pointer jumping (the subroot of
// find_roots_with_lakes, stream_tree.rs:139-173) into w.sub.
int32_t label_chain_ends(fl_ctx* c, fl_terrain* T) {
  hipStream_t s = c->stream;
  fl_launch_sub_init(s, T->t, T->w);
  bool flip = false;
  for (uint32_t round = 0;; round++) {
    fl_launch_fill_u32(s, (uint32_t*)T->w.flags + FL_FLAG_JUMP, 1, 0);
    fl_launch_sub_jump(s, T->t, T->w, flip);
    flip = !flip;
    FLCHK(read_ctrl(c, T));
    if (!host_flag(c, FL_FLAG_JUMP)) break;
    if (round > 64) return set_err(FL_ERR_HIP, "chain-end labelling did not converge");
  }
  if (flip) std::swap(T->w.sub, T->w.sub2);
  return FL_OK;
}

// Lake joining (stream_tree.rs:175-243) on the receivers in w.next.
int32_t join_lakes(fl_ctx* c, fl_terrain* T) {
  FLCHK(label_chain_ends(c, T));
  FLCHK(ensure_visit_order(c, T));
  fl_launch_lake_join(c->stream, T->t, T->w);
  return FL_OK;
}

// Decomposition of the receiver forest into heavy chains laid out by light
// depth; ends with the control block on the host (bucket offsets, flags).
// `joined`: lakes were joined since the receivers marked the dirty basins
// (the joins' changes are marked too).
int32_t decompose(fl_ctx* c, fl_terrain* T, bool full, bool joined) {
  hipStream_t s = c->stream;
  fl_dev_work& w = T->w;
  uint32_t n = T->t.n;
  if (!full && joined) fl_launch_dirty(s, T->t, w);
  fl_launch_children(s, T->t, w, full, T->d_tid, T->d_active);
  uint32_t round = 0;
  uint32_t target = std::max<uint32_t>(T->jump_rounds, 1);
  uint32_t quiet = FL_NONE_ROUND;  // first round that left every pair at its fixed point
  for (;;) {
    for (; round < target; round++) fl_launch_chain_jump(s, T->t, w, round);
    FLCHK(read_ctrl(c, T));
    for (uint32_t r = 0; r < round; r++)
      if (c->host_ctrl[FL_CTRL_JUMPS + r] == 0) {
        quiet = r;
        break;
      }
    if (quiet != FL_NONE_ROUND) break;
    target = round + 2;
    if (target >= FL_MAX_JUMP_ROUNDS) return set_err(FL_ERR_HIP, "tree decomposition did not converge");
  }
  // quiet + 1 rounds suffice (the quiet round completes convergence)
  T->jump_rounds = quiet + 1;
  w.guard = w.ctrl + FL_CTRL_JUMPS + (round - 1);  // a round at or after the quiet one: zero
  fl_launch_chain_count(s, T->t, w);
  fl_launch_chain_layout(s, T->t, w, 0);
  FLCHK(read_ctrl(c, T));
  uint32_t nb = 2 * (host_counter(c, FL_CNT_MAXLD) + 1);  // short and long bucket per light depth
  if (nb > 2 * n + 2) return set_err(FL_ERR_HIP, "light depth out of range");
  T->prev_nb = nb;
  c->host_boff.resize(2 * (nb + 1));
  if (nb <= FL_CTRL_HEAD) {
    const uint32_t* hc = c->host_ctrl + FL_NUM_FLAGS + FL_NUM_COUNTERS;
    std::copy(hc, hc + nb + 1, c->host_boff.begin());
    std::copy(hc + FL_CTRL_HEAD + 1, hc + FL_CTRL_HEAD + 1 + nb + 1, c->host_boff.begin() + nb + 1);
  } else {
    HIPCHK(hipMemcpyAsync(c->host_boff.data(), w.boff, (nb + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(c->host_boff.data() + nb + 1, w.bnoff, (nb + 1) * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    T->stat_syncs++;
  }
  T->last_boff = c->host_boff;
  T->last_rounds = nb / 2;
  return FL_OK;
}

// Launch shape of round l from bucket offsets (boff[nb+1] then bnoff[nb+1]).
fl_round_grid exact_grid(const std::vector<uint32_t>& off, uint32_t l) {
  uint32_t nb = (uint32_t)off.size() / 2 - 1;
  const uint32_t* bo = off.data();
  const uint32_t* bn = bo + nb + 1;
  fl_round_grid g{};
  if (2 * l + 2 > nb) return g;
  g.positions = bn[2 * l + 2] - bn[2 * l];
  g.fold_blocks = (bo[2 * l + 2] - bo[2 * l + 1]) + (bo[2 * l + 1] - bo[2 * l] + 63) / 64;
  return g;
}

// First round of the tail launches: the rounds from it on each hold at most
// `units` fold units (tuned; the tail's FL_TAIL_WAVES wavefronts loop over
// more); `rounds` (no tail) when fewer than two rounds qualify, as one round
// gains nothing.
uint32_t tail_start(const std::vector<uint32_t>& off, uint32_t rounds, uint32_t units) {
  uint32_t l = rounds;
  while (l > 0 && exact_grid(off, l - 1).fold_blocks <= units) l--;
  return rounds - l >= 2 ? l : rounds;
}

// Chain positions of all rounds: the sites of the basins an iteration
// recomputed (0 when the offsets were not read back).
uint64_t all_positions(const std::vector<uint32_t>& off) {
  if (off.size() < 2) return 0;
  return off.back();
}

uint64_t tail_positions(const std::vector<uint32_t>& off, uint32_t from, uint32_t rounds) {
  uint64_t units = 0;
  for (uint32_t l = from; l < rounds; l++) units += exact_grid(off, l).positions;
  return units;
}

// A decomposition far deeper than subtree-size heavy children allow (light
// depth <= log2 n): the previous drainage areas described another tree (a
// river that moved takes its whole length of light edges), and the sweeps
// would run one round per light edge.
bool deep_rounds(const fl_ctx* c, uint32_t rounds, uint32_t n) {
  if (c->rehint_rounds) return rounds > c->rehint_rounds;
  uint32_t lg = 1;
  while (lg < 32 && (1u << lg) < n) lg++;
  return rounds > 2 * lg + 8;
}

// One outer iteration (generator.rs:141-209), synchronous form: the host
// confirms pointer-jumping convergence and reads the bucket offsets before
// the sweeps.  The decomposition is built speculatively before the lake test
// (lakes occur in the first iteration only in practice).  `full`: every
// basin is recomputed (no previous tree), else only the dirty ones.
// `keep`: sites whose neighbourhood the previous iteration left unchanged
// keep its receivers (its tree had no lake joins).
int32_t one_iteration(fl_ctx* c, fl_terrain* T, bool full, bool keep, bool* changed, bool* had_lake) {
  hipStream_t s = c->stream;
  fl_dev_work& w = T->w;
  uint32_t n = T->t.n;
  fl_launch_clear_iter(s, w, std::min<uint32_t>(T->prev_nb + 1, 2 * n + 2));
  fl_launch_receivers(s, T->t, w, keep, !full);
  if (T->hint_stale) {
    // the heavy children follow subtree sizes of the final tree: with the
    // sites' own areas as the hint the chains are short and the light depth
    // (sweep rounds) runs to ten times the usual
    FLCHK(read_ctrl(c, T));
    *had_lake = host_flag(c, FL_FLAG_LAKE) != 0;
    if (*had_lake) {
      FLCHK(join_lakes(c, T));
      fl_launch_clear_iter(s, w, std::min<uint32_t>(T->prev_nb + 1, 2 * n + 2));
    }
    // sizes of the whole forest: sites of basins kept from the previous tree
    // get them too (their drainage areas are read by no recomputed basin)
    std::string err;
    if (fl_device_subtree_sizes(s, n, w.next, T->d_tid, T->d_active, w.A, &err, &T->arena) != FL_OK)
      return set_err(FL_ERR_HIP, "%s", err.c_str());
    T->hint_stale = false;
    FLCHK(decompose(c, T, full, *had_lake));
  } else {
    FLCHK(decompose(c, T, full, false));
    *had_lake = host_flag(c, FL_FLAG_LAKE) != 0;
    if (*had_lake) {
      FLCHK(join_lakes(c, T));
      fl_launch_clear_iter(s, w, std::min<uint32_t>(T->prev_nb + 1, 2 * n + 2));
      FLCHK(decompose(c, T, full, true));
    }
    if (deep_rounds(c, T->last_rounds, n)) {
      // decompose again with subtree sizes as the hint
      std::string err;
      fl_launch_clear_iter(s, w, std::min<uint32_t>(T->prev_nb + 1, 2 * n + 2));
      if (fl_device_subtree_sizes(s, n, w.next, T->d_tid, T->d_active, w.A, &err, &T->arena) != FL_OK)
        return set_err(FL_ERR_HIP, "%s", err.c_str());
      T->stat_rehinted++;
      FLCHK(decompose(c, T, full, false));
    }
  }
  uint32_t nb = T->prev_nb;
  const std::vector<uint32_t>& off = c->host_boff;
  bool clamp = T->t.any_max_slope;
  uint32_t rounds = nb / 2;
  // the deep rounds, which hold few chains, go to one launch per sweep
  uint32_t split = tail_start(off, rounds, T->tail_units);
  uint64_t tail_units = tail_positions(off, split, rounds);
  if (split < rounds) fl_launch_da_tail(s, T->t, w, split, rounds - 1, tail_units, clamp);
  for (uint32_t l = split; l-- > 0;) fl_launch_da_round(s, T->t, w, l, exact_grid(off, l), clamp);
  for (uint32_t l = 0; l < split; l++) fl_launch_tau_round(s, T->t, w, l, exact_grid(off, l), clamp);
  if (split < rounds) fl_launch_tau_tail(s, T->t, w, split, rounds - 1, tail_units, clamp);
  if (T->d_tchanged) fl_launch_fill_u32(s, T->d_tchanged, T->nterr, 0);
  fl_launch_tau_scatter(s, T->t, w, clamp, T->d_tid, T->d_active, T->d_tchanged);
  if (T->d_tchanged)
    HIPCHK(hipMemcpyAsync(T->h_tchanged, T->d_tchanged, T->nterr * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  FLCHK(read_ctrl(c, T));
  *changed = host_flag(c, FL_FLAG_CHANGED) != 0;
  std::swap(w.h, w.h2);
  if (*had_lake) T->hint_stale = true;
  return FL_OK;
}

// The kernels of one pipelined iteration (J jump rounds, R sweep rounds,
// rounds S..R-1 in the tail launches), enqueued on the context's stream or
// recorded by a capture.  rec_* receive the profiler records of the launches
// (when profiling).
void enqueue_fast_body(fl_ctx* c, fl_terrain* T, uint32_t J, uint32_t R, uint32_t S, bool full, bool keep,
                       std::vector<size_t>* rec_da, std::vector<size_t>* rec_tau, size_t* rec_tail,
                       size_t* rec_scatter) {
  hipStream_t s = c->stream;
  fl_dev_work& w = T->w;
  uint32_t n = T->t.n;
  bool clamp = T->t.any_max_slope;
  // the previous iteration used 2 * (R - 1) bucket entries at most
  fl_launch_clear_iter(s, w, std::min<uint32_t>(2 * R + 1, 2 * n + 2));
  fl_launch_receivers(s, T->t, w, keep, !full);
  fl_launch_children(s, T->t, w, full, T->d_tid, T->d_active);
  for (uint32_t r = 0; r < J; r++) fl_launch_chain_jump(s, T->t, w, r);
  w.guard = w.ctrl + FL_CTRL_JUMPS + (J - 1);
  fl_launch_chain_count(s, T->t, w);
  fl_launch_chain_layout(s, T->t, w, R);
  // launch sizes from the previous iteration's rounds (grid-stride kernels:
  // any size is correct), profiler units filled in once the bounds are known
  uint32_t fold_default = std::max<uint32_t>(1, std::min<uint32_t>((n + 127) / 128, 8192));
  auto guess = [&](uint32_t l) {
    fl_round_grid g = T->last_boff.empty() ? fl_round_grid{} : exact_grid(T->last_boff, l);
    g.fold_blocks = g.fold_blocks ? std::min<uint32_t>(g.fold_blocks + g.fold_blocks / 4 + 8, fold_default) : 8;
    if (T->last_boff.empty()) g.fold_blocks = fold_default;
    g.positions = 0;
    return g;
  };
  if (S < R) {
    if (c->prof && rec_tail) rec_tail[0] = c->prof->pending.size();
    fl_launch_da_tail(s, T->t, w, S, R - 1, 0, clamp);
  }
  for (uint32_t l = S; l-- > 0;) {
    if (c->prof && rec_da) (*rec_da)[l] = c->prof->pending.size();
    fl_launch_da_round(s, T->t, w, l, guess(l), clamp);
  }
  for (uint32_t l = 0; l < S; l++) {
    if (c->prof && rec_tau) (*rec_tau)[l] = c->prof->pending.size();
    fl_launch_tau_round(s, T->t, w, l, guess(l), clamp);
  }
  if (S < R) {
    if (c->prof && rec_tail) rec_tail[1] = c->prof->pending.size();
    fl_launch_tau_tail(s, T->t, w, S, R - 1, 0, clamp);
  }
  if (T->d_tchanged) fl_launch_fill_u32(s, T->d_tchanged, T->nterr, 0);
  if (c->prof && rec_scatter) *rec_scatter = c->prof->pending.size();
  fl_launch_tau_scatter(s, T->t, w, clamp, T->d_tid, T->d_active, T->d_tchanged);
}

// The pipelined iteration for (J, R) and the current buffer parity as a
// graph, captured on first use.  False when capture is not possible (the
// caller launches directly).
bool launch_fast_graph(fl_ctx* c, fl_terrain* T, uint32_t J, uint32_t R, uint32_t S, bool full, bool keep) {
  if (T->graphs_failed) return false;
  int p = T->w.h == T->hb[0] ? 0 : (T->w.h == T->hb[1] ? 1 : 2);
  const int mode = (full ? 1 : 0) | (keep ? 2 : 0);
  if (!T->gexec[p] || T->gJ[p] != J || T->gR[p] != R || T->gS[p] != S || T->gF[p] != mode ||
      T->gB[p] != fl_site_block || T->gL[p] != T->w.long_min) {
    if (T->gexec[p]) {
      hipGraphExecDestroy(T->gexec[p]);
      T->gexec[p] = nullptr;
    }
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      T->graphs_failed = true;
      return false;
    }
    enqueue_fast_body(c, T, J, R, S, full, keep, nullptr, nullptr, nullptr, nullptr);
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&T->gexec[p], graph, nullptr, nullptr, 0);
    if (graph) hipGraphDestroy(graph);
    if (e != hipSuccess) {
      T->gexec[p] = nullptr;
      T->graphs_failed = true;
      (void)hipGetLastError();
      return false;
    }
    T->gJ[p] = J;
    T->gR[p] = R;
    T->gS[p] = S;
    T->gF[p] = mode;
    T->gB[p] = fl_site_block;
    T->gL[p] = T->w.long_min;
  }
  return hipGraphLaunch(T->gexec[p], c->stream) == hipSuccess;
}

// A pipelined iteration in flight: it reads hb[in] and writes hb[(in+1) % 3].
struct Flight {
  int in = 0;
  int slot = 0;
  bool full = false;  // every basin recomputed (no previous tree)
  bool keep = false;  // receivers kept where the neighbourhood did not change
  uint32_t J = 0, R = 0, S = 0;
  std::vector<uint8_t> mask;  // terrains iterating in it
  std::vector<size_t> rec_da, rec_tau;
  size_t rec_tail[2] = {SIZE_MAX, SIZE_MAX};
  size_t rec_scatter = SIZE_MAX;
};

// One outer iteration enqueued whole, without waiting for the previous one.
// Three things are taken from the last confirmed iteration instead of being
// read back in between: the number of pointer-jumping rounds (one more is
// launched; a round after a quiet one returns at once), the number of sweep
// rounds (one more; rounds past the deepest light depth are empty, their
// bounds read on the device) and the absence of lakes.  Its control block
// and changed flags are copied to the host slot f.slot and an event marks
// its end.
int32_t enqueue_fast(fl_ctx* c, fl_terrain* T, Flight& f) {
  hipStream_t s = c->stream;
  fl_dev_work& w = T->w;
  f.J = std::min<uint32_t>(T->jump_rounds + 1, FL_MAX_JUMP_ROUNDS);
  f.R = std::max<uint32_t>(T->last_rounds, 1) + 1;
  if (c->pipelined == 2) f.J = f.R = 1;
  // tail launches for the rounds that were small in the last confirmed
  // iteration; a split within one round of the previous one is kept (the
  // captured graphs stay valid)
  f.S = f.R;
  if (!T->last_boff.empty()) {
    uint32_t want = tail_start(T->last_boff, f.R, T->tail_units);
    uint32_t keep = T->tail_split;
    f.S = (want < f.R && keep < f.R && want + 1 >= keep && want <= keep + 1) ? keep : want;
  }
  T->tail_split = f.S;
  set_slots(T, f.in);
  f.rec_da.assign(f.R, SIZE_MAX);
  f.rec_tau.assign(f.R, SIZE_MAX);
  f.rec_tail[0] = f.rec_tail[1] = SIZE_MAX;
  bool graphed = !c->prof && c->graphs && launch_fast_graph(c, T, f.J, f.R, f.S, f.full, f.keep);
  if (!graphed)
    enqueue_fast_body(c, T, f.J, f.R, f.S, f.full, f.keep, &f.rec_da, &f.rec_tau, f.rec_tail, &f.rec_scatter);
  if (T->d_tchanged)
    HIPCHK(hipMemcpyAsync(T->h_tslot[f.slot], T->d_tchanged, T->nterr * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(c->host_slot[f.slot], w.ctrl, FL_CTRL_WORDS * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipEventRecord(c->slot_ev[f.slot], s));
  return FL_OK;
}

// Wait for flight f and check what it assumed.  *valid is false when any
// assumption failed: the decomposition had not converged (the kernels after
// it then did nothing), a lake was found, or the tree was deeper.  Its input
// hb[f.in] is untouched either way.  *changed: the changed flag; per-terrain
// flags are in T->h_tslot[f.slot].
int32_t finish_fast(fl_ctx* c, fl_terrain* T, const Flight& f, bool* valid, bool* changed) {
  HIPCHK(hipEventSynchronize(c->slot_ev[f.slot]));
  T->stat_syncs++;
  const uint32_t* hc = c->host_slot[f.slot];
  uint32_t n = T->t.n;
  bool lake = hc[FL_FLAG_LAKE] != 0;
  bool converged = hc[FL_CTRL_JUMPS + f.J - 1] == 0;
  uint32_t nb = 2 * (hc[FL_NUM_FLAGS + FL_CNT_MAXLD] + 1);
  *valid = false;
  *changed = false;
  if (!converged) {
    T->jump_rounds = std::max(T->jump_rounds, f.J + 1);
    return FL_OK;
  }
  if (nb > 2 * n + 2) return set_err(FL_ERR_HIP, "light depth out of range");
  for (uint32_t r = 0; r < f.J; r++)
    if (hc[FL_CTRL_JUMPS + r] == 0) {
      T->jump_rounds = r + 1;
      break;
    }
  if (lake || nb / 2 > f.R) return FL_OK;
  *valid = true;
  T->prev_nb = nb;
  T->last_rounds = nb / 2;
  // the next iterations would speculate as many rounds: they wait for a
  // synchronous one with subtree sizes (the flights already enqueued land)
  if (deep_rounds(c, T->last_rounds, n)) {
    T->hint_stale = true;
    T->stat_rehinted++;
  }
  if (nb <= FL_CTRL_HEAD) {
    const uint32_t* head = hc + FL_NUM_FLAGS + FL_NUM_COUNTERS;
    T->last_boff.assign(head, head + nb + 1);
    T->last_boff.insert(T->last_boff.end(), head + FL_CTRL_HEAD + 1, head + FL_CTRL_HEAD + 1 + nb + 1);
    if (c->prof) {
      for (uint32_t l = 0; l < f.R; l++) {
        uint64_t units = exact_grid(T->last_boff, l).positions;
        for (size_t q : {f.rec_da[l], f.rec_tau[l]})
          if (q < c->prof->pending.size()) c->prof->pending[q].units = units;
      }
      uint64_t tail_units = tail_positions(T->last_boff, f.S, f.R);
      for (size_t q : f.rec_tail)
        if (q < c->prof->pending.size()) c->prof->pending[q].units = tail_units;
    }
  } else {
    T->last_boff.clear();
  }
  *changed = hc[FL_FLAG_CHANGED] != 0;
  return FL_OK;
}

// Sweep rounds of iteration `it` of a solve into its statistics.
void note_rounds(fl_terrain* T, uint32_t rounds, uint32_t it) {
  if (it == 0) T->stat_rounds_first = rounds;
  T->stat_rounds_max = std::max(T->stat_rounds_max, rounds);
}

int32_t reset_state(fl_ctx* c, fl_terrain* T) {
  uint32_t n = T->t.n;
  T->prev_valid = false;
  T->after_lake = true;
  T->hint_stale = true;
  HIPCHK(hipMemcpyAsync(T->w.h, T->t.init_h, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(T->w.A, T->t.area, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  return FL_OK;
}

}  // namespace

// ------------------------------------------------------------------ API ---

extern "C" const char* fl_last_error(void) { return g_err.c_str(); }

void fl_internal_set_error(const char* msg) { g_err = msg ? msg : ""; }

hipError_t fl_arena_alloc(fl_arena* a, hipStream_t s, void** p, uint64_t bytes) {
  bytes = fl_round256(bytes);
  if (!a || a->kind == fl_arena::POOL)
    return a && a->pool ? hipMallocFromPoolAsync(p, bytes, a->pool, s) : hipMallocAsync(p, bytes, s);
  if (a->kind == fl_arena::REGION && a->used + bytes > a->size) {
    *p = nullptr;
    return hipErrorOutOfMemory;
  }
  *p = a->kind == fl_arena::REGION ? (void*)(a->base + a->used) : nullptr;
  a->used += bytes;
  a->peak = std::max(a->peak, a->used);
  return hipSuccess;
}

extern "C" const char* fl_version(void) { return "fastlem_hip 0.1.0 (gfx950, f64)"; }

namespace {

// The library's own memory pool per device.  Freed device memory stays in it
// for the next terrain (unbounded release threshold), without changing the
// device's default pool that other HIP code in the process uses; it is
// trimmed when the device's last context is destroyed.  The registry is
// never destroyed (contexts may be destroyed during process exit).
struct PoolRegistry {
  std::mutex mu;
  std::vector<std::pair<hipMemPool_t, int>> dev;  // per device: pool, live contexts
};
PoolRegistry& pools() {
  static PoolRegistry* r = new PoolRegistry();
  return *r;
}

int32_t acquire_pool(int device, hipMemPool_t* out) {
  PoolRegistry& r = pools();
  std::lock_guard<std::mutex> g(r.mu);
  if ((int)r.dev.size() <= device) r.dev.resize(device + 1, {nullptr, 0});
  auto& e = r.dev[device];
  if (!e.first) {
    hipMemPoolProps props{};
    props.allocType = hipMemAllocationTypePinned;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = device;
    HIPCHK(hipMemPoolCreate(&e.first, &props));
    uint64_t keep = UINT64_MAX;
    HIPCHK(hipMemPoolSetAttribute(e.first, hipMemPoolAttrReleaseThreshold, &keep));
  }
  e.second++;
  *out = e.first;
  return FL_OK;
}

void release_pool(int device) {
  PoolRegistry& r = pools();
  std::lock_guard<std::mutex> g(r.mu);
  if ((int)r.dev.size() <= device || !r.dev[device].first) return;
  auto& e = r.dev[device];
  if (--e.second == 0) (void)hipMemPoolTrimTo(e.first, 0);
}

}  // namespace

extern "C" int32_t fl_ctx_create(int32_t device, fl_ctx** out) {
  if (!out) return set_err(FL_ERR_INVALID, "null output pointer");
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return set_err(FL_ERR_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= count) return set_err(FL_ERR_INVALID, "device %d out of range", device);
  // a failure below releases what was set up so far
  std::unique_ptr<fl_ctx, int32_t (*)(fl_ctx*)> c(new fl_ctx(), fl_ctx_destroy);
  c->device = device;
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  FLCHK(acquire_pool(device, &c->pool));
  FLCHK(fl_noalloc_scratch_init(device));  // the static scratch's address on this device
  HIPCHK(hipHostMalloc((void**)&c->host_ctrl, FL_CTRL_WORDS * sizeof(uint32_t), hipHostMallocDefault));
  for (int q = 0; q < 2; q++) {
    HIPCHK(hipHostMalloc((void**)&c->host_slot[q], FL_CTRL_WORDS * sizeof(uint32_t), hipHostMallocDefault));
    HIPCHK(hipEventCreateWithFlags(&c->slot_ev[q], hipEventDisableTiming));
  }
  *out = c.release();
  return FL_OK;
}

int32_t fl_ctx_stream(fl_ctx* c, hipStream_t* stream, int* device) {
  if (!c) return set_err(FL_ERR_INVALID, "null context");
  *stream = c->stream;
  *device = c->device;
  return FL_OK;
}

extern "C" int32_t fl_ctx_destroy(fl_ctx* c) {
  if (!c) return FL_OK;
  hipSetDevice(c->device);
  delete c->prof;
  if (c->stream) hipStreamDestroy(c->stream);
  if (c->host_ctrl) hipHostFree(c->host_ctrl);
  for (int q = 0; q < 2; q++) {
    if (c->host_slot[q]) hipHostFree(c->host_slot[q]);
    if (c->slot_ev[q]) hipEventDestroy(c->slot_ev[q]);
  }
  for (void* b : c->pin_blocks) hipHostFree(b);
  if (c->pool) release_pool(c->device);
  delete c;
  return FL_OK;
}

extern "C" int32_t fl_terrain_destroy(fl_terrain* T) {
  if (!T) return FL_OK;
  if (T->ctx) hipSetDevice(T->ctx->device);
  for (hipGraphExec_t& e : T->gexec)
    if (e) hipGraphExecDestroy(e);
  // every call that used the terrain has returned (they end synchronised),
  // so the frees are ordered after its last work on any stream; the pool
  // reuses the memory in stream order
  if (T->ctx) {
    for (void* p : T->allocations) (void)hipFreeAsync(p, T->ctx->stream);
  } else {
    for (void* p : T->allocations) hipFree(p);
  }
  delete T;
  return FL_OK;
}

namespace {

// Breadth-first (Cuthill-McKee style) site order from the graph alone:
// neighbours of a site get nearby labels, so the gathers of every kernel
// touch few cache lines.  Row contents keep their order, so every slot-order
// tie-break of the reference is preserved.
void locality_order(uint32_t n, const uint32_t* rp, const uint32_t* col, std::vector<uint32_t>& old_of_new,
                    std::vector<uint32_t>& new_of_old) {
  old_of_new.assign(n, 0);
  new_of_old.assign(n, 0xffffffffu);
  uint32_t head = 0, tail = 0;
  for (uint32_t start = 0; start < n; start++) {
    if (new_of_old[start] != 0xffffffffu) continue;
    new_of_old[start] = tail;
    old_of_new[tail++] = start;
    while (head < tail) {
      uint32_t u = old_of_new[head++];
      for (uint32_t k = rp[u]; k < rp[u + 1]; k++) {
        uint32_t v = col[k];
        if (new_of_old[v] == 0xffffffffu) {
          new_of_old[v] = tail;
          old_of_new[tail++] = v;
        }
      }
    }
  }
}

template <typename V>
void gather_perm(const V* src, const std::vector<uint32_t>& old_of_new, std::vector<V>& dst) {
  dst.resize(old_of_new.size());
  for (size_t u = 0; u < old_of_new.size(); u++) dst[u] = src[old_of_new[u]];
}

}  // namespace

namespace {

// The device arrays of a union of `count` terrains with n sites and nnz
// entries, from P->arena (with a COUNT arena: their sizes only).
int32_t layout_device(fl_terrain* P, uint32_t n, uint32_t nnz, uint32_t count, bool clamp) {
  fl_dev_terrain& t = P->t;
  fl_dev_work& w = P->w;
  t.n = n;
  t.nnz = nnz;
  FLCHK(dev_alloc(P, &t.row_ptr, (uint64_t)n + 1));
  FLCHK(dev_alloc(P, &t.col, nnz));
  FLCHK(dev_alloc(P, &t.len, nnz));
  FLCHK(dev_alloc(P, &t.area, n));
  FLCHK(dev_alloc(P, &t.erod, n));
  FLCHK(dev_alloc(P, &t.uplift, n));
  FLCHK(dev_alloc(P, &t.tan_ms, n));
  FLCHK(dev_alloc(P, &t.init_h, n));
  FLCHK(dev_alloc(P, &t.is_outlet, n));
  FLCHK(dev_alloc(P, &t.rank, n));
  FLCHK(dev_alloc(P, &t.order, n));
  t.any_max_slope = clamp;
  for (double** p : {&w.h, &w.h2, &w.A, &w.tau, &w.xp, &w.taup, &w.parea, &w.pK, &w.pdn}) FLCHK(dev_alloc(P, p, n));
  for (int k = 0; k < 3; k++) {
    FLCHK(dev_alloc(P, &P->dxb[k], n));
    FLCHK(dev_alloc(P, &P->nxb[k], n));
    FLCHK(dev_alloc(P, &P->rtb[k], n));
    FLCHK(dev_alloc(P, &P->chgb[k], n));
  }
  FLCHK(dev_alloc(P, &w.dirty, n));
  FLCHK(dev_alloc(P, &w.act, n));
  FLCHK(dev_alloc(P, &w.cval, 4 * (uint64_t)n));
  FLCHK(dev_alloc(P, &P->hb[2], n));
  P->hb[0] = w.h;
  P->hb[1] = w.h2;
  set_slots(P, 0);
  if (clamp)
    for (double** p : {&w.hnp, &w.up, &w.tanp, &w.dp, &w.hop}) FLCHK(dev_alloc(P, p, n));
  for (uint32_t** p : {&w.sub, &w.sub2, &w.hc, &w.pcnt, &w.pdest, &w.head_chain, &w.head_base, &w.perm})
    FLCHK(dev_alloc(P, p, n));
  FLCHK(dev_alloc(P, &w.cmeta, n));
  FLCHK(dev_alloc(P, &w.cp, n));
  FLCHK(dev_alloc(P, &w.rt, n));
  FLCHK(dev_alloc(P, &w.cmask, n));
  for (uint32_t** p : {&w.bcnt, &w.bnodes, &w.boff, &w.bnoff, &w.bcur, &w.bncur})
    FLCHK(dev_alloc(P, p, 2 * (uint64_t)n + 4));
  FLCHK(dev_alloc(P, &w.lake_key, n));
  FLCHK(dev_alloc(P, &w.ctrl, FL_CTRL_WORDS));
  w.flags = (int*)w.ctrl;
  w.counters = w.ctrl + FL_NUM_FLAGS;
  if (count > 1) {
    FLCHK(dev_alloc(P, &P->d_tid, n));
    FLCHK(dev_alloc(P, &P->d_active, count));
    FLCHK(dev_alloc(P, &P->d_tchanged, count));
  }
  return FL_OK;
}

// HBM of a union of `count` terrains (n sites, nnz entries) as
// layout_device takes it.
uint64_t union_device_bytes(uint64_t n, uint64_t nnz, uint32_t count, bool clamp) {
  fl_terrain P;
  P.arena.kind = fl_arena::COUNT;
  if (layout_device(&P, (uint32_t)n, (uint32_t)nnz, count, clamp) != FL_OK) return UINT64_MAX;
  return P.bytes;
}

// Device arrays of a union of `count` terrains with n sites and nnz entries
// (layout_device), pinned host slots for a batch; the bucket histograms
// start at zero.
int32_t alloc_device(fl_ctx* c, fl_terrain* P, uint32_t n, uint32_t nnz, uint32_t count, bool clamp) {
  if (P->arena.kind == fl_arena::POOL) {
    // one pool block for all the arrays (one allocation and one free: the
    // pool's cost grows with the number of its blocks), cut as a region
    const uint64_t bytes = union_device_bytes(n, nnz, count, clamp);
    void* block = nullptr;
    hipError_t e = fl_arena_alloc(&P->arena, c->stream, &block, bytes);
    if (e != hipSuccess)
      return set_err(e == hipErrorOutOfMemory ? FL_ERR_OUT_OF_MEMORY : FL_ERR_HIP, "device arrays (%llu bytes): %s",
                     (unsigned long long)bytes, hipGetErrorString(e));
    P->allocations.push_back(block);
    const fl_arena pool = P->arena;
    P->arena = fl_arena{};
    P->arena.kind = fl_arena::REGION;
    P->arena.base = (uint8_t*)block;
    P->arena.size = fl_round256(bytes);
    const int32_t rc = layout_device(P, n, nnz, count, clamp);
    P->arena = pool;
    FLCHK(rc);
  } else {
    FLCHK(layout_device(P, n, nnz, count, clamp));
  }
  fl_dev_work& w = P->w;
  if (count > 1) {
    if (count > c->pin_cap) {
      const uint32_t cap = std::max<uint32_t>({count, 2 * c->pin_cap, 1024});
      void* block = nullptr;
      HIPCHK(hipHostMalloc(&block, (size_t)cap * 13, hipHostMallocDefault));
      c->pin_blocks.push_back(block);
      c->pin = (uint8_t*)block;
      c->pin_cap = cap;
    }
    P->h_tchanged = (uint32_t*)c->pin;
    P->h_tslot[0] = P->h_tchanged + c->pin_cap;
    P->h_tslot[1] = P->h_tslot[0] + c->pin_cap;
    P->h_mask = (uint8_t*)(P->h_tslot[1] + c->pin_cap);
  }
  hipStream_t s = c->stream;
  // basin roots are site indices at all times (dirty-basin marks index by them)
  for (int k = 0; k < 3; k++) fl_launch_fill_u32(s, P->rtb[k], n, 0);
  P->prev_valid = false;
  fl_launch_fill_u32(s, w.bcnt, 2 * (uint64_t)n + 4, 0);
  fl_launch_fill_u32(s, w.bnodes, 2 * (uint64_t)n + 4, 0);
  fl_launch_fill_u32(s, w.bcur, 2 * (uint64_t)n + 4, 0);
  fl_launch_fill_u32(s, w.bncur, 2 * (uint64_t)n + 4, 0);
  P->prev_nb = 0;
  return FL_OK;
}

// Each outlet's row as the outlet phase needs it (fl_terrain::o_deg, o_ptr,
// o_nb) from the caller's graphs, which passed their checks; slot: union site
// -> 1 + its index in its terrain's outlet list (0: not an outlet).
void outlet_rows(fl_terrain* T, const fl_terrain_desc* descs, uint32_t count, const uint32_t* slot) {
  const uint32_t K = T->outlet_off[count];
  T->o_deg.assign(K, 0);
  T->o_ptr.assign((size_t)K + 1, 0);
  // f(g, q, row of outlet q, its terrain's first union site) for outlets
  // [a, b) of the union, terrain by terrain
  auto each = [&](auto f) {
    host_parallel(K, 1 << 14, [&](uint64_t a, uint64_t b) {
      uint32_t g = (uint32_t)(std::upper_bound(T->outlet_off.begin(), T->outlet_off.end(), (uint32_t)a) -
                              T->outlet_off.begin()) - 1;
      for (uint64_t q = a; q < b; q++) {
        while (T->outlet_off[g + 1] <= q) g++;
        const fl_terrain_desc* d = &descs[g];
        const uint32_t o = d->outlets[q - T->outlet_off[g]];
        f(q, d->row_ptr[o], d->row_ptr[o + 1], d->col, T->site_off[g]);
      }
    });
  };
  each([&](uint64_t q, uint32_t ka, uint32_t kb, const uint32_t* col, uint32_t s0) {
    uint32_t m = 0;
    for (uint32_t k = ka; k < kb; k++) m += slot[s0 + col[k]] != 0;
    T->o_deg[q] = kb - ka;
    T->o_ptr[q + 1] = m;
  });
  for (uint32_t q = 0; q < K; q++) T->o_ptr[q + 1] += T->o_ptr[q];
  T->o_nb.resize(T->o_ptr[K]);
  each([&](uint64_t q, uint32_t ka, uint32_t kb, const uint32_t* col, uint32_t s0) {
    uint32_t e = T->o_ptr[q];
    for (uint32_t k = ka; k < kb; k++)
      if (const uint32_t z = slot[s0 + col[k]]) T->o_nb[e++] = z - 1;
  });
}

// Staging of the device path of create_union: the caller's arrays as given
// and the relabelling maps (the per-site arrays stage in work arrays).
struct Staging {
  uint32_t *meta = nullptr, *rp = nullptr, *col = nullptr, *idx = nullptr, *o2n = nullptr, *n2o = nullptr;
  double* len = nullptr;
};
hipError_t alloc_staging(fl_scratch& S, uint64_t n, uint64_t nnz, uint64_t K, uint32_t G, Staging& st) {
  hipError_t e;
  if ((e = S.alloc(&st.meta, 2ull * (G + 1))) != hipSuccess) return e;
  if ((e = S.alloc(&st.rp, n + G)) != hipSuccess) return e;
  if ((e = S.alloc(&st.col, nnz)) != hipSuccess) return e;
  if ((e = S.alloc(&st.len, nnz)) != hipSuccess) return e;
  if ((e = S.alloc(&st.idx, K)) != hipSuccess) return e;
  if ((e = S.alloc(&st.o2n, n)) != hipSuccess) return e;
  return S.alloc(&st.n2o, n);
}

uint64_t staging_bytes(uint64_t n, uint64_t nnz, uint64_t K, uint32_t G) {
  fl_arena a;
  a.kind = fl_arena::COUNT;
  {
    fl_scratch S(nullptr, &a);
    Staging st;
    alloc_staging(S, n, nnz, K, G, st);
  }
  return a.peak;
}

// Device bytes beyond a union's arrays that its creation and first visit
// order need at their peak.
uint64_t scratch_bytes(uint64_t n, uint64_t nnz, uint64_t K, uint32_t G) {
  const uint64_t create = staging_bytes(n, nnz, K, G) + fl_create_scratch_bytes((uint32_t)n, (uint32_t)nnz, G);
  return std::max({create, fl_msf_scratch_bytes((uint32_t)n, (uint32_t)nnz), fl_subtree_scratch_bytes((uint32_t)n)});
}

// Upload `count` terrains as one union graph: terrain g's sites occupy
// [site_off[g], site_off[g+1]), each relabelled for locality inside its range;
// no edge joins two terrains.  The graphs are checked and relabelled on the
// device when every row holds at most FL_DEVICE_MAX_ROW entries, else on host
// threads; both report the same first error.
int32_t create_union(fl_ctx* c, const fl_terrain_desc* descs, uint32_t count, fl_terrain** out,
                     const fl_arena* region = nullptr) {
  if (!c || !out || (!descs && count)) return set_err(FL_ERR_INVALID, "null context, description or output");
  if (count == 0) return set_err(FL_ERR_INVALID, "no terrain");
  using ms_ = std::chrono::duration<double, std::milli>;
  const auto c0 = std::chrono::steady_clock::now();
  // a terrain's error, named as the first failing terrain of a batch
  auto fail = [&](uint32_t g, int32_t rc, const std::string& msg) {
    return count > 1 ? set_err(rc, "terrain %u: %s", g, msg.c_str()) : set_err(rc, "%s", msg.c_str());
  };
  // checks that need no look at the rows, on host threads
  std::vector<int32_t> rc(count, FL_OK);
  std::vector<std::string> msg(count);
  std::vector<uint32_t> longest(count, 0);
  auto each_terrain = [&](auto f) {
    if (count == 1)
      f(0u);
    else
      host_each(count, [&](uint32_t g) { return descs[g].nnz; }, f);
  };
  each_terrain([&](uint32_t g) {
    rc[g] = check_header(&descs[g], &longest[g]);
    if (rc[g] != FL_OK) msg[g] = g_err;
  });
  bool device_path = true;
  for (uint32_t g = 0; g < count; g++) device_path &= rc[g] == FL_OK && longest[g] <= FL_DEVICE_MAX_ROW;
  // host path: every check of every terrain on host threads
  std::vector<uint8_t> asym(count, 0);
  std::unique_ptr<uint32_t[]> slot;
  if (!device_path) {
    std::vector<std::unique_ptr<uint32_t[]>> tslot(count);
    each_terrain([&](uint32_t g) {
      if (rc[g] != FL_OK) return;
      bool a = false;
      rc[g] = check_rows(&descs[g], &a);
      asym[g] = a;
      if (rc[g] == FL_OK) {
        tslot[g].reset(new uint32_t[descs[g].n + 1]());
        rc[g] = check_outlets(&descs[g], tslot[g].get());
      }
      if (rc[g] != FL_OK) msg[g] = g_err;
    });
    for (uint32_t g = 0; g < count; g++)
      if (rc[g] != FL_OK) return fail(g, rc[g], msg[g]);
    uint64_t N = 0;
    for (uint32_t g = 0; g < count; g++) N += descs[g].n;
    if (N < 0xffffffffull) {
      slot.reset(new uint32_t[N + 1]);
      for (uint32_t g = 0, s0 = 0; g < count; s0 += (uint32_t)descs[g].n, g++)
        std::copy(tslot[g].get(), tslot[g].get() + descs[g].n, slot.get() + s0);
    }
  }
  uint64_t N = 0, NNZ = 0, K = 0;
  for (uint32_t g = 0; g < count; g++) {
    N += descs[g].n;
    NNZ += descs[g].nnz;
    K += descs[g].n_outlets;
  }
  if (N > FL_MAX_SITES || NNZ >= 0xffffffffull)
    return set_err(FL_ERR_TOO_LARGE, "batch exceeds 2^31-1 sites or 2^32-2 entries");
  HIPCHK(hipSetDevice(c->device));
  std::unique_ptr<fl_terrain, int32_t (*)(fl_terrain*)> T(new fl_terrain(), fl_terrain_destroy);
  T->ctx = c;
  T->arena.pool = c->pool;
  if (region) T->arena = *region;
  T->nterr = count;
  T->max_iteration = descs[0].max_iteration;
  uint32_t n = (uint32_t)N, nnz = (uint32_t)NNZ;
  T->t.n = n;
  T->t.nnz = nnz;
  T->site_off.assign(count + 1, 0);
  T->outlet_off.assign(count + 1, 0);
  T->entry_off.assign(count + 1, 0);
  T->t_max_it.resize(count);
  T->t_iters.assign(count, 0);
  T->t_active.assign(count, 0);
  T->t_asym = asym;
  for (uint32_t g = 0; g < count; g++) {
    T->site_off[g + 1] = T->site_off[g] + (uint32_t)descs[g].n;
    T->outlet_off[g + 1] = T->outlet_off[g] + (uint32_t)descs[g].n_outlets;
    T->entry_off[g + 1] = T->entry_off[g] + (uint32_t)descs[g].nnz;
    T->t_max_it[g] = descs[g].max_iteration;
  }
  if (n == 0) {
    *out = T.release();
    return FL_OK;
  }
  if (device_path) {
    // outlets checked on host threads (reported after the terrain's row errors)
    slot.reset(new uint32_t[n]());
    each_terrain([&](uint32_t g) {
      rc[g] = check_outlets(&descs[g], slot.get() + T->site_off[g]);
      if (rc[g] != FL_OK) msg[g] = g_err;
    });
  }
  // tan(max slope) on the host, with the reference's libm; NaN: no clamp
  std::unique_ptr<double[]> tan_ms;
  bool clamp = false;
  for (uint32_t g = 0; g < count; g++) clamp |= descs[g].n > 0 && descs[g].has_max_slope != nullptr;
  if (clamp) {
    tan_ms.reset(new double[n]);
    std::atomic<bool> any{false};
    each_terrain([&](uint32_t g) {
      const fl_terrain_desc* d = &descs[g];
      double* dst = tan_ms.get() + T->site_off[g];
      auto run = [&](uint64_t a, uint64_t b) {
        bool some = false;
        for (uint64_t i = a; i < b; i++) {
          dst[i] = (d->has_max_slope && d->has_max_slope[i]) ? std::tan(d->max_slope[i]) : std::nan("");
          some |= !std::isnan(dst[i]);
        }
        if (some) any.store(true, std::memory_order_relaxed);
      };
      if (count == 1)
        host_parallel(d->n, 1 << 16, run);
      else
        run(0, d->n);
    });
    clamp = any.load();
  }
  fl_terrain* P = T.get();
  fl_dev_terrain& t = T->t;
  hipStream_t s = c->stream;
  if (!device_path) {
    // host path: breadth-first order and copies per terrain (rows and site
    // values over host threads for one terrain)
    const auto c1 = std::chrono::steady_clock::now();
    T->old_of_new.resize(n);
    T->new_of_old.resize(n);
    T->h_outlets.resize(K);
    std::vector<uint32_t> h_row_ptr((size_t)n + 1, 0), h_col(nnz);
    std::vector<double> h_len(nnz), area(n), erod(n), uplift(n), init(n);
    std::vector<uint8_t> is_out(n, 0);
    auto relabel = [&](uint32_t g, bool inner) {
      const fl_terrain_desc* d = &descs[g];
      const uint32_t s0 = T->site_off[g], ng = (uint32_t)d->n;
      if (ng == 0) return;
      std::vector<uint32_t> o2, n2;
      locality_order(ng, d->row_ptr, d->col, o2, n2);
      uint32_t q = T->entry_off[g];
      for (uint32_t u = 0; u < ng; u++) {
        const uint32_t i = o2[u];
        q += d->row_ptr[i + 1] - d->row_ptr[i];
        h_row_ptr[s0 + u + 1] = q;
      }
      auto fill = [&](uint64_t lo, uint64_t hi) {
        for (uint64_t u = lo; u < hi; u++) {
          const uint32_t i = o2[u];
          T->old_of_new[s0 + u] = s0 + i;
          T->new_of_old[s0 + i] = s0 + (uint32_t)u;
          // (the start of the terrain's first row is its own offset: the
          // entry before it belongs to the previous terrain)
          uint32_t e = u == 0 ? T->entry_off[g] : h_row_ptr[s0 + u];
          for (uint32_t k = d->row_ptr[i]; k < d->row_ptr[i + 1]; k++, e++) {
            h_col[e] = s0 + n2[d->col[k]];
            h_len[e] = d->length[k];
          }
          area[s0 + u] = d->area[i];
          erod[s0 + u] = d->erodibility[i];
          uplift[s0 + u] = d->uplift_rate[i];
          init[s0 + u] = d->init_elevation[i];
        }
      };
      if (inner)
        host_parallel(ng, 65536, fill);
      else
        fill(0, ng);
      for (uint64_t k = 0; k < d->n_outlets; k++) {
        uint32_t o = s0 + n2[d->outlets[k]];
        T->h_outlets[T->outlet_off[g] + k] = o;
        is_out[o] = 1;
      }
    };
    if (count == 1)
      relabel(0, true);
    else
      host_each(
          count, [&](uint32_t g) { return descs[g].nnz; }, [&](uint32_t g) { relabel(g, false); });
    std::vector<double> tan_new;
    if (clamp) {
      tan_new.resize(n);
      for (uint32_t u = 0; u < n; u++) tan_new[u] = tan_ms[T->old_of_new[u]];
    }
    outlet_rows(P, descs, count, slot.get());
    const auto c2 = std::chrono::steady_clock::now();
    FLCHK(alloc_device(c, P, n, nnz, count, clamp));
    HIPCHK(hipMemcpyAsync(t.row_ptr, h_row_ptr.data(), ((uint64_t)n + 1) * 4, hipMemcpyHostToDevice, s));
    if (nnz) {
      HIPCHK(hipMemcpyAsync(t.col, h_col.data(), (uint64_t)nnz * 4, hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(t.len, h_len.data(), (uint64_t)nnz * 8, hipMemcpyHostToDevice, s));
    }
    HIPCHK(hipMemcpyAsync(t.area, area.data(), (uint64_t)n * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(t.erod, erod.data(), (uint64_t)n * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(t.uplift, uplift.data(), (uint64_t)n * 8, hipMemcpyHostToDevice, s));
    if (clamp)
      HIPCHK(hipMemcpyAsync(t.tan_ms, tan_new.data(), (uint64_t)n * 8, hipMemcpyHostToDevice, s));
    else
      fl_launch_fill_u32(s, (uint32_t*)t.tan_ms, 2ull * n, 0xffffffffu);  // NaN
    HIPCHK(hipMemcpyAsync(t.init_h, init.data(), (uint64_t)n * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(t.is_outlet, is_out.data(), n, hipMemcpyHostToDevice, s));
    {
      fl_scratch S(s, &T->arena);
      uint32_t* d_off = nullptr;
      HIPCHK(S.alloc(&d_off, (uint64_t)count + 1));
      HIPCHK(hipMemcpyAsync(d_off, T->site_off.data(), (count + 1ull) * 4, hipMemcpyHostToDevice, s));
      if (count > 1) fl_device_fill_tid(s, n, count, d_off, T->d_tid);
      FLCHK(reset_state(c, P));
      // the pageable staging vectors above must outlive the copies
      HIPCHK(hipStreamSynchronize(s));
    }
    T->stat_create[0] = ms_(c1 - c0).count();
    T->stat_create[1] = ms_(c2 - c1).count();
    T->stat_create[2] = ms_(std::chrono::steady_clock::now() - c2).count();
    *out = T.release();
    return FL_OK;
  }
  // device path: the caller's arrays uploaded as given, checked, ordered and
  // relabelled on the device
  const auto c1 = std::chrono::steady_clock::now();
  FLCHK(alloc_device(c, P, n, nnz, count, clamp));
  fl_dev_work& w = T->w;
  std::string err;
  double t_check = 0.0, t_relabel = 0.0;
  {
    fl_scratch S(s, &T->arena);
    Staging st;
    hipError_t e = S.reserve(staging_bytes(n, nnz, K, count));
    if (e == hipSuccess) e = alloc_staging(S, n, nnz, K, count, st);
    if (e != hipSuccess) {
      if (T->arena.kind == fl_arena::REGION)
        return set_err(FL_ERR_TOO_LARGE, "device region of %llu bytes exhausted", (unsigned long long)T->arena.size);
      return set_err(e == hipErrorOutOfMemory ? FL_ERR_OUT_OF_MEMORY : FL_ERR_HIP, "staging: %s", hipGetErrorString(e));
    }
    std::vector<uint32_t> meta(2ull * (count + 1));
    std::copy(T->site_off.begin(), T->site_off.end(), meta.begin());
    std::copy(T->entry_off.begin(), T->entry_off.end(), meta.begin() + count + 1);
    HIPCHK(hipMemcpyAsync(st.meta, meta.data(), meta.size() * 4, hipMemcpyHostToDevice, s));
    // per-site arrays stage in work arrays, which the solve writes before it
    // reads them
    double* raw[5] = {w.xp, w.taup, w.tau, T->dxb[0], w.parea};
    for (uint32_t g = 0; g < count; g++) {
      const fl_terrain_desc* d = &descs[g];
      const uint64_t s0 = T->site_off[g], e0 = T->entry_off[g], ng = d->n, eg = d->nnz;
      if (ng == 0) continue;
      HIPCHK(hipMemcpyAsync(st.rp + s0 + g, d->row_ptr, (ng + 1) * 4, hipMemcpyHostToDevice, s));
      if (eg) {
        HIPCHK(hipMemcpyAsync(st.col + e0, d->col, eg * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(st.len + e0, d->length, eg * 8, hipMemcpyHostToDevice, s));
      }
      HIPCHK(hipMemcpyAsync(raw[0] + s0, d->area, ng * 8, hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(raw[1] + s0, d->erodibility, ng * 8, hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(raw[2] + s0, d->uplift_rate, ng * 8, hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(raw[3] + s0, d->init_elevation, ng * 8, hipMemcpyHostToDevice, s));
    }
    if (clamp) HIPCHK(hipMemcpyAsync(raw[4], tan_ms.get(), (uint64_t)n * 8, hipMemcpyHostToDevice, s));
    const uint32_t* d_site_off = st.meta;
    if (count > 1) fl_device_fill_tid(s, n, count, d_site_off, T->d_tid);
    fl_raw_union R;
    R.n = n;
    R.nnz = nnz;
    R.G = count;
    R.site_off = d_site_off;
    R.entry_off = st.meta + count + 1;
    R.tid = count > 1 ? T->d_tid : nullptr;
    R.rp = st.rp;
    R.col = st.col;
    const auto c2 = std::chrono::steady_clock::now();
    std::vector<uint32_t> bad, unpaired;
    std::vector<uint8_t> differs;
    int r = fl_device_check_graph(s, R, st.len, bad, unpaired, differs, &err, &T->arena);
    if (r != FL_OK) return set_err(r, "graph checks: %s", err.c_str());
    for (uint32_t g = 0; g < count; g++) {
      if (bad[g] != 0xffffffffu || unpaired[g] != 0xffffffffu) {
        row_error(&descs[g], bad[g] != 0xffffffffu ? bad[g] : unpaired[g], bad[g] == 0xffffffffu);
        return fail(g, FL_ERR_INVALID, g_err);
      }
      if (rc[g] != FL_OK) return fail(g, rc[g], msg[g]);
      T->t_asym[g] = differs[g];
    }
    const auto c3 = std::chrono::steady_clock::now();
    t_check = ms_(c3 - c2).count();
    // the outlets' rows on a host thread while the device orders and relabels
    std::future<void> rows = std::async(std::launch::async, [&] { outlet_rows(P, descs, count, slot.get()); });
    r = fl_device_locality(s, R, st.o2n, st.n2o, &T->stat_create_clusters, &err, &T->arena);
    if (r == FL_OK) {
      const double* src[5] = {raw[0], raw[1], raw[2], raw[3], clamp ? raw[4] : nullptr};
      double* const dst[5] = {t.area, t.erod, t.uplift, t.init_h, t.tan_ms};
      r = fl_device_relabel(s, R, st.len, st.o2n, st.n2o, t.row_ptr, t.col, t.len, src, dst, &err, &T->arena);
    }
    rows.get();
    if (r != FL_OK) return set_err(r, "relabelling: %s", err.c_str());
    T->old_of_new.resize(n);
    T->new_of_old.resize(n);
    HIPCHK(hipMemcpyAsync(T->old_of_new.data(), st.o2n, (uint64_t)n * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(T->new_of_old.data(), st.n2o, (uint64_t)n * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    T->h_outlets.resize(K);
    for (uint32_t g = 0; g < count; g++)
      for (uint64_t q = 0; q < descs[g].n_outlets; q++)
        T->h_outlets[T->outlet_off[g] + q] = T->new_of_old[T->site_off[g] + descs[g].outlets[q]];
    if (K) HIPCHK(hipMemcpyAsync(st.idx, T->h_outlets.data(), K * 4, hipMemcpyHostToDevice, s));
    fl_device_mark_u8(s, n, (uint32_t)K, st.idx, t.is_outlet);
    FLCHK(reset_state(c, P));
    // the host arrays above must outlive the copies
    HIPCHK(hipStreamSynchronize(s));
    t_relabel = ms_(std::chrono::steady_clock::now() - c3).count();
    T->stat_create[2] = ms_(c2 - c1).count();
  }
  T->stat_create[0] = ms_(c1 - c0).count() + t_check;
  T->stat_create[1] = t_relabel;
  T->stat_create_device = 1;
  *out = T.release();
  return FL_OK;
}

// A batch keeps sweeping the sites of terrains that stopped iterating; once
// the running terrains hold less than half of the union (and at least
// fl_ctx::compact_min_idle sites are idle), they are gathered into a smaller
// union.

// The running terrains of `cur` (a union derived from T, or T itself) gathered
// on the device into a new union; orig maps cur's terrains to T's.  The
// stopped terrains' elevations are copied to their ranges of T->w.h first.
int32_t compact_union(fl_ctx* c, fl_terrain* T, fl_terrain* cur, const uint32_t* caps,
                      const std::vector<uint32_t>& orig, fl_terrain** out, std::vector<uint32_t>& new_orig) {
  hipStream_t s = c->stream;
  std::vector<uint32_t> keep;
  for (uint32_t g = 0; g < cur->nterr; g++) {
    if (cur->t_active[g]) {
      keep.push_back(g);
    } else if (cur != T) {
      uint32_t s0 = cur->site_off[g], ng = cur->site_off[g + 1] - s0;
      if (ng)
        HIPCHK(hipMemcpyAsync(T->w.h + T->site_off[orig[g]], cur->w.h + s0, (uint64_t)ng * 8,
                              hipMemcpyDeviceToDevice, s));
    }
  }
  uint32_t G = (uint32_t)keep.size();
  std::unique_ptr<fl_terrain, int32_t (*)(fl_terrain*)> N(new fl_terrain(), fl_terrain_destroy);
  N->ctx = c;
  N->arena.pool = c->pool;
  N->nterr = G;
  N->site_off.assign(G + 1, 0);
  N->entry_off.assign(G + 1, 0);
  N->outlet_off.assign(G + 1, 0);
  N->t_max_it.resize(G);
  N->t_iters.resize(G);
  N->t_active.assign(G, 1);
  N->t_asym.assign(G, 0);
  new_orig.resize(G);
  std::vector<uint32_t> map(4 * (size_t)G + 1);  // new_off[G+1], old_off[G], new_eoff[G], old_eoff[G]
  for (uint32_t q = 0; q < G; q++) {
    uint32_t g = keep[q];
    N->t_asym[q] = cur->t_asym[g];
    N->site_off[q + 1] = N->site_off[q] + (cur->site_off[g + 1] - cur->site_off[g]);
    N->entry_off[q + 1] = N->entry_off[q] + (cur->entry_off[g + 1] - cur->entry_off[g]);
    N->outlet_off[q + 1] = N->outlet_off[q] + (cur->outlet_off[g + 1] - cur->outlet_off[g]);
    N->t_max_it[q] = caps[g];
    N->t_iters[q] = cur->t_iters[g];
    new_orig[q] = orig[g];
  }
  for (uint32_t q = 0; q <= G; q++) map[q] = N->site_off[q];
  for (uint32_t q = 0; q < G; q++) {
    map[G + 1 + q] = cur->site_off[keep[q]];
    map[2 * G + 1 + q] = N->entry_off[q];
    map[3 * G + 1 + q] = cur->entry_off[keep[q]];
  }
  uint32_t n = N->site_off[G], nnz = N->entry_off[G];
  FLCHK(alloc_device(c, N.get(), n, nnz, G > 1 ? G : 2, cur->t.any_max_slope));
  uint32_t* d_map = nullptr;
  FLCHK(dev_alloc(N.get(), &d_map, map.size()));
  HIPCHK(hipMemcpyAsync(d_map, map.data(), map.size() * 4, hipMemcpyHostToDevice, s));
  fl_compact_map m{G, d_map, d_map + G + 1, d_map + 2 * G + 1, d_map + 3 * G + 1};
  fl_launch_compact(s, m, cur->t, cur->w, N->t, N->w, N->d_tid, 0, n, 0);
  HIPCHK(hipMemcpyAsync(N->d_active, N->t_active.data(), G, hipMemcpyHostToDevice, s));
  N->have_order = true;
  N->jump_rounds = cur->jump_rounds;
  N->last_rounds = cur->last_rounds;
  N->w.long_min = cur->w.long_min;
  N->tail_units = cur->tail_units;
  HIPCHK(hipStreamSynchronize(s));  // `map` is pageable
  *out = N.release();
  return FL_OK;
}

// Iterate every terrain of T from its initial elevations: terrain g runs
// until an iteration changes none of its elevations or it reached max_it[g];
// it is then frozen (its elevations no longer change) while the others go
// on, so each terrain's result and count equal its solo run.
int32_t solve_union(fl_ctx* c, fl_terrain* T, const uint32_t* max_it) {
  uint32_t G = T->nterr;
  auto t0 = std::chrono::steady_clock::now();
  T->stat_syncs = 0;
  T->stat_order_ms = 0.0;
  T->stat_recomputed = 0;
  T->stat_rounds_first = 0;
  T->stat_rounds_max = 0;
  T->stat_rehinted = 0;
  T->stat_lake_iterations = 0;
  T->stat_compactions = 0;
  T->stat_redone = 0;
  uint32_t active = 0;
  uint64_t active_sites = 0;
  for (uint32_t g = 0; g < G; g++) {
    T->t_iters[g] = 0;
    bool empty = T->site_off[g + 1] == T->site_off[g];
    // an empty terrain runs the loop body once and finds nothing changed
    if (empty && max_it[g] > 0) T->t_iters[g] = 1;
    T->t_active[g] = (!empty && max_it[g] > 0) ? 1 : 0;
    active += T->t_active[g];
    if (T->t_active[g]) active_sites += T->site_off[g + 1] - T->site_off[g];
  }
  if (T->t.n == 0) return FL_OK;
  set_slots(T, 0);
  FLCHK(reset_state(c, T));  // elevations start from the initial state even if nothing iterates
  if (active == 0) {
    HIPCHK(hipStreamSynchronize(c->stream));
    return FL_OK;
  }
  if (T->d_active)
    HIPCHK(hipMemcpyAsync(T->d_active, T->t_active.data(), G, hipMemcpyHostToDevice, c->stream));
  fl_terrain* cur = T;
  const uint32_t* caps = max_it;  // iteration caps of cur's terrains
  std::unique_ptr<fl_terrain, int32_t (*)(fl_terrain*)> owned(nullptr, fl_terrain_destroy);
  std::vector<uint32_t> orig(G);
  for (uint32_t g = 0; g < G; g++) orig[g] = g;
  hipStream_t s = c->stream;
  uint32_t it = 0;
  int hin = 0;  // cur->hb[hin] holds the last confirmed elevations (reset_state wrote hb[0])
  std::deque<Flight> q;
  std::vector<uint32_t> enq(cur->t_iters);  // iterations enqueued per terrain (confirmed + in flight)
  std::vector<uint8_t> dev_mask(cur->t_active);  // the mask the device holds
  int next_slot = 0;
  bool drain = false;  // stop enqueueing: a compaction waits for the flights to land
  auto upload_mask = [&](const std::vector<uint8_t>& m) -> int32_t {
    if (!cur->d_active || m == dev_mask) return FL_OK;
    std::copy(m.begin(), m.end(), cur->h_mask);
    HIPCHK(hipMemcpyAsync(cur->d_active, cur->h_mask, cur->nterr, hipMemcpyHostToDevice, s));
    dev_mask = m;
    return FL_OK;
  };
  // account one confirmed iteration of the terrains in `mask`
  auto confirm = [&](const std::vector<uint8_t>& mask, const uint32_t* tflags, bool changed) {
    for (uint32_t g = 0; g < cur->nterr; g++) {
      if (!mask[g] || !cur->t_active[g]) continue;
      cur->t_iters[g]++;
      T->t_iters[orig[g]] = cur->t_iters[g];
      bool ch = tflags ? tflags[g] != 0 : changed;
      if (!ch || cur->t_iters[g] >= caps[g]) {
        cur->t_active[g] = 0;
        active--;
        active_sites -= cur->site_off[g + 1] - cur->site_off[g];
      }
    }
  };
  while (active) {
    if (cur != T) {
      cur->stat_syncs = 0;
      cur->stat_rehinted = 0;
    }
    // the first iteration, those whose heavy-child hint is stale (after lake
    // joins: a deep decomposition there would also make a poor speculation
    // for the iterations after it) and the unpipelined form
    if (it == 0 || (cur->hint_stale && q.empty()) || !c->pipelined) {
      bool changed = false, lake = false;
      std::vector<uint8_t> mask(cur->t_active);
      FLCHK(upload_mask(mask));
      set_slots(cur, hin);
      const bool full = !cur->prev_valid || c->recompute_all;
      const bool keep = !full && !cur->after_lake;
      cur->prev_valid = true;
      FLCHK(one_iteration(c, cur, full, keep, &changed, &lake));
      cur->after_lake = lake;
      T->stat_lake_iterations += lake ? 1 : 0;
      for (uint32_t g = 0; g < cur->nterr; g++) enq[g] += mask[g];
      confirm(mask, cur->d_tchanged ? cur->h_tchanged : nullptr, changed);
      T->stat_recomputed += all_positions(cur->last_boff);
      note_rounds(T, cur->last_rounds, it);
      hin = (hin + 1) % 3;
      it++;
    } else {
      // keep two iterations in flight: the host enqueues the next one while
      // the device runs the current one
      while (!drain && !cur->hint_stale && q.size() < 2) {
        std::vector<uint8_t> mask(cur->nterr);
        bool any = false;
        for (uint32_t g = 0; g < cur->nterr; g++) {
          mask[g] = cur->t_active[g] && enq[g] < caps[g];
          any |= mask[g] != 0;
        }
        if (!any) break;
        FLCHK(upload_mask(mask));
        Flight f;
        f.in = q.empty() ? hin : (q.back().in + 1) % 3;
        f.slot = next_slot;
        next_slot ^= 1;
        f.mask = mask;
        f.full = !cur->prev_valid || c->recompute_all;
        f.keep = !f.full && !cur->after_lake;
        cur->prev_valid = true;
        cur->after_lake = false;  // a flight that finds lakes is redone synchronously
        FLCHK(enqueue_fast(c, cur, f));
        for (uint32_t g = 0; g < cur->nterr; g++) enq[g] += mask[g];
        q.push_back(std::move(f));
      }
      if (q.empty()) break;  // cannot happen while a terrain is active
      Flight f = std::move(q.front());
      q.pop_front();
      bool valid = false, changed = false;
      FLCHK(finish_fast(c, cur, f, &valid, &changed));
      const uint32_t* tflags = cur->d_tchanged ? cur->h_tslot[f.slot] : nullptr;
      if (!valid) {
        // the flights after it started from a wrong input: let them land,
        // then redo this iteration synchronously on its own input and mask
        HIPCHK(hipStreamSynchronize(s));
        for (const Flight& x : q)
          for (uint32_t g = 0; g < cur->nterr; g++) enq[g] -= x.mask[g];
        q.clear();
        T->stat_redone++;
        bool lake = false;
        FLCHK(upload_mask(f.mask));
        set_slots(cur, f.in);
        FLCHK(one_iteration(c, cur, f.full, f.keep, &changed, &lake));
        cur->after_lake = lake;
        T->stat_lake_iterations += lake ? 1 : 0;
        tflags = cur->d_tchanged ? cur->h_tchanged : nullptr;
      }
      confirm(f.mask, tflags, changed);
      T->stat_recomputed += all_positions(cur->last_boff);
      note_rounds(T, cur->last_rounds, it);
      hin = (f.in + 1) % 3;
      it++;
    }
    if (cur != T) {
      T->stat_syncs += cur->stat_syncs;
      T->stat_rehinted += cur->stat_rehinted;
    }
    if (!active) break;
    bool want = cur->nterr > 1 && active_sites * 2 < cur->t.n && cur->t.n - active_sites >= c->compact_min_idle;
    if (want) drain = true;
    if (drain && q.empty()) {
      set_slots(cur, hin);
      fl_terrain* next_union = nullptr;
      std::vector<uint32_t> next_orig;
      FLCHK(compact_union(c, T, cur, caps, orig, &next_union, next_orig));
      owned.reset(next_union);  // destroys the previous derived union, if any
      cur = next_union;
      caps = cur->t_max_it.data();
      orig.swap(next_orig);
      hin = 0;
      enq = cur->t_iters;
      dev_mask = cur->t_active;
      drain = false;
      T->stat_compactions++;
    }
  }
  // iterations still in flight after the last one that counted recompute a
  // fixed point (or masked terrains): wait for them, their results are unused
  HIPCHK(hipStreamSynchronize(s));
  set_slots(cur, hin);
  if (cur != T) {
    for (uint32_t g = 0; g < cur->nterr; g++) {
      uint32_t s0 = cur->site_off[g], ng = cur->site_off[g + 1] - s0;
      if (ng)
        HIPCHK(hipMemcpyAsync(T->w.h + T->site_off[orig[g]], cur->w.h + s0, (uint64_t)ng * 8,
                              hipMemcpyDeviceToDevice, s));
    }
    HIPCHK(hipStreamSynchronize(s));
  }
  T->stat_iterations = it;
  T->stat_solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return FL_OK;
}

}  // namespace

// ---------------------------------------------------------------- lanes ---
//
// A lane iterates a union whose terrains change at iteration boundaries (the
// serving loop's continuous batching).  Prepared terrains join it: the
// terrains still running and the joining ones are gathered into a new union
// on the device, and the first iteration after a join runs synchronously
// (the joining terrains' first iteration joins lakes).  A terrain that stops
// (no elevation changed, or its max_iteration reached) has its elevations
// copied out at once and leaves the union at the next gather, which also
// happens when stopped terrains hold half of the union.  Each terrain sees
// exactly the iterations of its own solve, so its result equals its solo run.

struct fl_lane {
  fl_ctx* c = nullptr;
  fl_terrain* cur = nullptr;  // owned; null when idle
  std::vector<fl_lane_member> mem;  // per terrain of cur
  int hin = 0;
  std::deque<Flight> q;
  std::vector<uint32_t> enq;
  std::vector<uint8_t> dev_mask;
  int next_slot = 0;
  uint32_t sync_next = 1;  // iterations to run synchronously (after a join: its first)
  bool drain = false;          // enqueue nothing ahead: the lane is to become quiet
  bool compact_wanted = false;  // gather (drop the stopped terrains) once quiet
  uint32_t active = 0;
  uint64_t active_sites = 0;
  uint32_t gathers = 0;
};

namespace {

int32_t lane_upload_mask(fl_lane* L, const std::vector<uint8_t>& m) {
  fl_terrain* cur = L->cur;
  if (m == L->dev_mask) return FL_OK;
  std::copy(m.begin(), m.end(), cur->h_mask);
  HIPCHK(hipMemcpyAsync(cur->d_active, cur->h_mask, cur->nterr, hipMemcpyHostToDevice, L->c->stream));
  L->dev_mask = m;
  return FL_OK;
}

// The lane's running terrains and the joining ones gathered into a new
// union (device to device); the old union and the joining terrains' own
// objects are released.
int32_t lane_gather(fl_lane* L, std::vector<std::pair<fl_terrain*, fl_lane_member>>& join) {
  fl_ctx* c = L->c;
  hipStream_t s = c->stream;
  fl_terrain* cur = L->cur;
  struct Part {
    fl_terrain* src;
    uint32_t g;  // terrain of src
  };
  std::vector<Part> parts;
  std::vector<fl_lane_member> mem;
  if (cur) {
    set_slots(cur, L->hin);  // the last confirmed elevations
    for (uint32_t g = 0; g < cur->nterr; g++)
      if (cur->t_active[g]) {
        parts.push_back({cur, g});
        mem.push_back(std::move(L->mem[g]));
        mem.back().iterations = cur->t_iters[g];
      }
  }
  for (auto& j : join) {
    parts.push_back({j.first, 0});
    mem.push_back(std::move(j.second));
    mem.back().iterations = 0;
  }
  const uint32_t G = (uint32_t)parts.size();
  std::unique_ptr<fl_terrain, int32_t (*)(fl_terrain*)> N(new fl_terrain(), fl_terrain_destroy);
  N->ctx = c;
  N->arena.pool = c->pool;
  N->nterr = G;
  N->site_off.assign(G + 1, 0);
  N->entry_off.assign(G + 1, 0);
  N->outlet_off.assign(G + 1, 0);
  N->t_max_it.resize(G);
  N->t_iters.resize(G);
  N->t_active.assign(G, 1);
  N->t_asym.assign(G, 0);
  bool clamp = false;
  uint64_t total_sites = 0, total_entries = 0;
  for (const Part& p : parts) {
    total_sites += p.src->site_off[p.g + 1] - p.src->site_off[p.g];
    total_entries += p.src->entry_off[p.g + 1] - p.src->entry_off[p.g];
  }
  if (total_sites > FL_MAX_SITES || total_entries >= 0xffffffffull)
    return set_err(FL_ERR_TOO_LARGE, "lane union exceeds 2^31-1 sites or 2^32-2 entries");
  for (uint32_t q = 0; q < G; q++) {
    const Part& p = parts[q];
    fl_terrain* S = p.src;
    N->site_off[q + 1] = N->site_off[q] + (S->site_off[p.g + 1] - S->site_off[p.g]);
    N->entry_off[q + 1] = N->entry_off[q] + (S->entry_off[p.g + 1] - S->entry_off[p.g]);
    N->outlet_off[q + 1] = N->outlet_off[q] + (S->outlet_off[p.g + 1] - S->outlet_off[p.g]);
    N->t_max_it[q] = mem[q].cap;
    N->t_iters[q] = mem[q].iterations;
    N->t_asym[q] = S->t_asym.empty() ? 0 : S->t_asym[p.g];
    clamp |= S->t.any_max_slope;
  }
  const uint32_t n = N->site_off[G], nnz = N->entry_off[G];
  if (G == 0 || n == 0) {
    // nothing left to iterate (or only empty terrains, which stop at once)
    if (cur) fl_terrain_destroy(cur);
    L->cur = nullptr;
    for (auto& j : join) fl_terrain_destroy(j.first);
    L->mem.clear();
    L->active = 0;
    L->active_sites = 0;
    return FL_OK;
  }
  FLCHK(alloc_device(c, N.get(), n, nnz, G > 1 ? G : 2, clamp));
  // one map per source union: its parts are contiguous in the new union
  std::vector<uint32_t> maps;
  std::vector<std::pair<size_t, std::pair<uint32_t, uint32_t>>> launches;  // map offset, part range [q0, q1)
  for (uint32_t q0 = 0; q0 < G;) {
    uint32_t q1 = q0 + 1;
    while (q1 < G && parts[q1].src == parts[q0].src) q1++;
    const uint32_t k = q1 - q0;
    const size_t off = maps.size();
    maps.resize(off + 4 * (size_t)k + 1);
    uint32_t* m = maps.data() + off;
    fl_terrain* S = parts[q0].src;
    for (uint32_t i = 0; i <= k; i++) m[i] = N->site_off[q0 + i];
    for (uint32_t i = 0; i < k; i++) {
      m[k + 1 + i] = S->site_off[parts[q0 + i].g];
      m[2 * k + 1 + i] = N->entry_off[q0 + i];
      m[3 * k + 1 + i] = S->entry_off[parts[q0 + i].g];
    }
    launches.push_back({off, {q0, q1}});
    q0 = q1;
  }
  uint32_t* d_map = nullptr;
  FLCHK(dev_alloc(N.get(), &d_map, maps.size()));
  HIPCHK(hipMemcpyAsync(d_map, maps.data(), maps.size() * 4, hipMemcpyHostToDevice, s));
  for (const auto& l : launches) {
    const uint32_t q0 = l.second.first, q1 = l.second.second, k = q1 - q0;
    uint32_t* m = d_map + l.first;
    fl_compact_map cm{k, m, m + k + 1, m + 2 * k + 1, m + 3 * k + 1};
    fl_terrain* S = parts[q0].src;
    if (N->site_off[q1] > N->site_off[q0])
      fl_launch_compact(s, cm, S->t, S->w, N->t, N->w, N->d_tid, N->site_off[q0], N->site_off[q1], q0);
  }
  HIPCHK(hipMemcpyAsync(N->d_active, N->t_active.data(), G, hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));  // `maps` is pageable; the sources are released below
  N->have_order = true;
  N->jump_rounds = cur ? cur->jump_rounds : 4;
  N->last_rounds = cur ? cur->last_rounds : 0;
  if (cur) fl_terrain_destroy(cur);
  for (auto& j : join) {
    j.first->ctx = c;  // its buffers were last used on its own (synchronised) stream: free them on this one
    fl_terrain_destroy(j.first);
  }
  L->cur = N.release();
  L->mem = std::move(mem);
  L->hin = 0;
  L->enq = L->cur->t_iters;
  L->dev_mask = L->cur->t_active;
  L->sync_next = join.empty() ? 0u : 1u;
  L->cur->hint_stale = !join.empty();
  L->drain = false;
  L->compact_wanted = false;
  L->active = G;
  L->active_sites = n;
  L->gathers++;
  return FL_OK;
}

// Terrain g of the lane's union stopped: its elevations (device buffer
// `buf`) into member g's host copy, in the caller's site order.
int32_t lane_copy_out(fl_lane* L, uint32_t g, const double* buf) {
  fl_terrain* cur = L->cur;
  fl_lane_member& m = L->mem[g];
  const uint32_t s0 = cur->site_off[g], ng = cur->site_off[g + 1] - s0;
  m.result.resize(ng);
  if (!ng) return FL_OK;
  std::vector<double> tmp(ng);
  HIPCHK(hipMemcpyAsync(tmp.data(), buf + s0, (uint64_t)ng * 8, hipMemcpyDeviceToHost, L->c->stream));
  HIPCHK(hipStreamSynchronize(L->c->stream));
  for (uint32_t i = 0; i < ng; i++) m.result[i] = tmp[m.new_of_old[i]];
  return FL_OK;
}

}  // namespace

fl_lane* fl_lane_create(fl_ctx* c) {
  fl_lane* L = new fl_lane();
  L->c = c;
  return L;
}

void fl_lane_destroy(fl_lane* L) {
  if (!L) return;
  if (L->cur) {
    hipStreamSynchronize(L->c->stream);
    fl_terrain_destroy(L->cur);
  }
  delete L;
}

int32_t fl_lane_prepare(fl_ctx* c, const fl_terrain_desc* d, fl_terrain** out, std::vector<uint32_t>* new_of_old) {
  *out = nullptr;
  HIPCHK(hipSetDevice(c->device));
  fl_terrain* T = nullptr;
  FLCHK(create_union(c, d, 1, &T));
  int32_t rc = T->t.n ? ensure_visit_order(c, T) : FL_OK;
  if (rc != FL_OK) {
    fl_terrain_destroy(T);
    return rc;
  }
  *new_of_old = std::move(T->new_of_old);
  *out = T;
  return FL_OK;
}

uint64_t fl_lane_active_sites(const fl_lane* L) { return L->active_sites; }
uint32_t fl_lane_terrains(const fl_lane* L) { return L->cur ? L->cur->nterr : 0; }
bool fl_lane_idle(const fl_lane* L) { return L->active == 0; }
bool fl_lane_quiet(const fl_lane* L) { return L->q.empty(); }
void fl_lane_set_drain(fl_lane* L) { L->drain = true; }

int32_t fl_lane_join(fl_lane* L, std::vector<std::pair<fl_terrain*, fl_lane_member>>& join) {
  HIPCHK(hipSetDevice(L->c->device));
  if (!L->q.empty()) return set_err(FL_ERR_INVALID, "a lane joins terrains only between iterations");
  return lane_gather(L, join);
}

int32_t fl_lane_step(fl_lane* L, std::vector<fl_lane_member>& stopped) {
  fl_ctx* c = L->c;
  hipStream_t s = c->stream;
  HIPCHK(hipSetDevice(c->device));
  if (!L->cur || L->active == 0) return FL_OK;
  fl_terrain* cur = L->cur;
  // a lane keeps its pipelining and graph settings (the mode of the
  // iterations in flight), the rest follows the union's class
  fl_tuning tt = effective_tuning(c, cur->t.n, std::max<uint32_t>(cur->nterr, 2));
  tt.pipelining = c->graphs ? 3 : (uint32_t)c->pipelined;
  TuneScope tune(c, cur, tt);
  std::vector<uint32_t> stop_now;
  // account one confirmed iteration of the terrains in `mask`
  auto confirm = [&](const std::vector<uint8_t>& mask, const uint32_t* tflags) {
    for (uint32_t g = 0; g < cur->nterr; g++) {
      if (!mask[g] || !cur->t_active[g]) continue;
      cur->t_iters[g]++;
      const bool ch = tflags[g] != 0;
      if (!ch || cur->t_iters[g] >= cur->t_max_it[g]) {
        cur->t_active[g] = 0;
        L->active--;
        L->active_sites -= cur->site_off[g + 1] - cur->site_off[g];
        stop_now.push_back(g);
      }
    }
  };
  const double* result = nullptr;  // the buffer holding the confirmed iteration's elevations
  if (L->sync_next || (cur->hint_stale && L->q.empty()) || !c->pipelined) {
    std::vector<uint8_t> mask(cur->t_active);
    FLCHK(lane_upload_mask(L, mask));
    set_slots(cur, L->hin);
    const bool full = !cur->prev_valid || c->recompute_all;
    const bool keep = !full && !cur->after_lake;
    cur->prev_valid = true;
    bool changed = false, lake = false;
    FLCHK(one_iteration(c, cur, full, keep, &changed, &lake));
    cur->after_lake = lake;
    for (uint32_t g = 0; g < cur->nterr; g++) L->enq[g] += mask[g];
    confirm(mask, cur->h_tchanged);
    L->hin = (L->hin + 1) % 3;
    result = cur->hb[L->hin];
    if (L->sync_next) L->sync_next--;
  } else {
    // a drain for a join whose terrains went to another lane: go on
    if (L->drain && L->q.empty() && !L->compact_wanted) L->drain = false;
    while (!L->drain && !cur->hint_stale && L->q.size() < 2) {
      std::vector<uint8_t> mask(cur->nterr);
      bool any = false;
      for (uint32_t g = 0; g < cur->nterr; g++) {
        mask[g] = cur->t_active[g] && L->enq[g] < cur->t_max_it[g];
        any |= mask[g] != 0;
      }
      if (!any) break;
      FLCHK(lane_upload_mask(L, mask));
      Flight f;
      f.in = L->q.empty() ? L->hin : (L->q.back().in + 1) % 3;
      f.slot = L->next_slot;
      L->next_slot ^= 1;
      f.mask = mask;
      f.full = !cur->prev_valid || c->recompute_all;
      f.keep = !f.full && !cur->after_lake;
      cur->prev_valid = true;
      cur->after_lake = false;  // a flight that finds lakes is redone synchronously
      FLCHK(enqueue_fast(c, cur, f));
      for (uint32_t g = 0; g < cur->nterr; g++) L->enq[g] += mask[g];
      L->q.push_back(std::move(f));
    }
    if (L->q.empty()) return FL_OK;
    Flight f = std::move(L->q.front());
    L->q.pop_front();
    bool valid = false, changed = false;
    FLCHK(finish_fast(c, cur, f, &valid, &changed));
    const uint32_t* tflags = cur->h_tslot[f.slot];
    if (!valid) {
      // the flights after it started from a wrong input: let them land, then
      // redo this iteration synchronously on its own input and mask
      HIPCHK(hipStreamSynchronize(s));
      for (const Flight& x : L->q)
        for (uint32_t g = 0; g < cur->nterr; g++) L->enq[g] -= x.mask[g];
      L->q.clear();
      cur->stat_redone++;
      bool lake = false;
      FLCHK(lane_upload_mask(L, f.mask));
      set_slots(cur, f.in);
      FLCHK(one_iteration(c, cur, f.full, f.keep, &changed, &lake));
      cur->after_lake = lake;
      tflags = cur->h_tchanged;
    }
    confirm(f.mask, tflags);
    L->hin = (f.in + 1) % 3;
    result = cur->hb[L->hin];
  }
  // stopped terrains: their elevations out now (the buffer of the confirmed
  // iteration is rewritten only two enqueues later)
  for (uint32_t g : stop_now) {
    FLCHK(lane_copy_out(L, g, result));
    fl_lane_member m = std::move(L->mem[g]);
    L->mem[g].tag = m.tag;  // the slot stays until the next gather
    m.iterations = cur->t_iters[g];
    stopped.push_back(std::move(m));
  }
  // gather when the stopped terrains hold half of the union
  if (L->active && L->active_sites * 2 < cur->t.n && cur->t.n - L->active_sites >= c->compact_min_idle) {
    L->drain = true;
    L->compact_wanted = true;
  }
  if (L->active == 0) {
    // nothing left: release the union once the last flights have landed
    HIPCHK(hipStreamSynchronize(s));
    L->q.clear();
    fl_terrain_destroy(cur);
    L->cur = nullptr;
    L->mem.clear();
    L->active_sites = 0;
  } else if (L->compact_wanted && L->q.empty()) {
    std::vector<std::pair<fl_terrain*, fl_lane_member>> none;
    FLCHK(lane_gather(L, none));
  }
  return FL_OK;
}

uint64_t fl_terrain_arrays_bytes(const fl_terrain_desc* d) {
  bool clamp = false;
  if (d->has_max_slope)
    for (uint64_t i = 0; i < d->n && !clamp; i++) clamp = d->has_max_slope[i] != 0;
  return union_device_bytes(d->n, d->nnz, 2, clamp);
}

extern "C" int32_t fl_terrain_create(fl_ctx* c, const fl_terrain_desc* d, fl_terrain** out) {
  if (!d) return set_err(FL_ERR_INVALID, "null terrain description");
  return create_union(c, d, 1, out);
}

extern "C" int32_t fl_batch_create(fl_ctx* c, const fl_terrain_desc* descs, uint32_t count, fl_terrain** out) {
  return create_union(c, descs, count, out);
}

extern "C" int32_t fl_batch_solve(fl_ctx* c, fl_terrain* T, uint32_t* iterations_out) {
  if (!c || !T) return set_err(FL_ERR_INVALID, "null context or batch");
  ProfScope prof_scope(c->prof);
  TuneScope tune(c, T, effective_tuning(c, T->t.n, T->nterr));
  HIPCHK(hipSetDevice(c->device));
  FLCHK(solve_union(c, T, T->t_max_it.data()));
  if (iterations_out) std::copy(T->t_iters.begin(), T->t_iters.end(), iterations_out);
  return FL_OK;
}

extern "C" int32_t fl_batch_count(const fl_terrain* T, uint32_t* count) {
  if (!T || !count) return set_err(FL_ERR_INVALID, "null argument");
  *count = T->nterr;
  return FL_OK;
}

extern "C" int32_t fl_terrain_solve(fl_ctx* c, fl_terrain* T, uint32_t max_iteration, uint32_t* iterations_out) {
  if (!c || !T) return set_err(FL_ERR_INVALID, "null context or terrain");
  ProfScope prof_scope(c->prof);
  TuneScope tune(c, T, effective_tuning(c, T->t.n, T->nterr));
  HIPCHK(hipSetDevice(c->device));
  std::vector<uint32_t> caps(T->nterr, max_iteration);
  FLCHK(solve_union(c, T, caps.data()));
  if (iterations_out) *iterations_out = T->t_iters[0];
  return FL_OK;
}

namespace {
// Copy a device f64[n] array (device labels) into caller order.
int32_t download_f64_old(fl_ctx* c, fl_terrain* T, double* dst, const double* src) {
  if (!dst) return FL_OK;
  uint32_t n = T->t.n;
  std::vector<double> tmp(n);
  HIPCHK(hipMemcpyAsync(tmp.data(), src, (uint64_t)n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (uint32_t i = 0; i < n; i++) dst[i] = tmp[T->new_of_old[i]];
  return FL_OK;
}
// Copy a device u32[n] array of site indices into caller order and labels.
int32_t download_idx_old(fl_ctx* c, fl_terrain* T, uint32_t* dst, const uint32_t* src) {
  if (!dst) return FL_OK;
  uint32_t n = T->t.n;
  std::vector<uint32_t> tmp(n);
  HIPCHK(hipMemcpyAsync(tmp.data(), src, (uint64_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (uint32_t i = 0; i < n; i++) dst[i] = T->old_of_new[tmp[T->new_of_old[i]]];
  return FL_OK;
}
int32_t upload_h(fl_ctx* c, fl_terrain* T, const double* elev) {
  uint32_t n = T->t.n;
  std::vector<double> tmp(n);
  for (uint32_t u = 0; u < n; u++) tmp[u] = elev[T->old_of_new[u]];
  HIPCHK(hipMemcpyAsync(T->w.h, tmp.data(), (uint64_t)n * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  T->prev_valid = false;  // given elevations: no previous tree
  T->after_lake = true;
  return FL_OK;
}
}  // namespace

extern "C" int32_t fl_terrain_download(fl_ctx* c, fl_terrain* T, double* out) {
  if (!c || !T || (!out && T->t.n)) return set_err(FL_ERR_INVALID, "null argument");
  if (T->t.n == 0) return FL_OK;
  HIPCHK(hipSetDevice(c->device));
  return download_f64_old(c, T, out, T->w.h);
}

extern "C" int32_t fl_batch_download(fl_ctx* c, fl_terrain* T, uint32_t index, double* out) {
  if (!c || !T || index >= T->nterr) return set_err(FL_ERR_INVALID, "bad batch index or null argument");
  uint32_t s0 = T->site_off[index], ng = T->site_off[index + 1] - s0;
  if (ng == 0) return FL_OK;
  if (!out) return set_err(FL_ERR_INVALID, "null output");
  HIPCHK(hipSetDevice(c->device));
  // a terrain's relabelled sites stay inside its own range
  std::vector<double> tmp(ng);
  HIPCHK(hipMemcpyAsync(tmp.data(), T->w.h + s0, (uint64_t)ng * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (uint32_t i = 0; i < ng; i++) out[i] = tmp[T->new_of_old[s0 + i] - s0];
  return FL_OK;
}

extern "C" int32_t fl_terrain_device_elevations(fl_terrain* T, double** dev_ptr) {
  if (!T || !dev_ptr) return set_err(FL_ERR_INVALID, "null argument");
  *dev_ptr = T->w.h;
  return FL_OK;
}

extern "C" int32_t fl_terrain_device_bytes(const fl_terrain* T, uint64_t* bytes) {
  if (!T || !bytes) return set_err(FL_ERR_INVALID, "null argument");
  *bytes = T->bytes;
  return FL_OK;
}

extern "C" int32_t fl_generate_f64(fl_ctx* c, const fl_terrain_desc* d, double* elevations_out,
                                   uint32_t* iterations_out) {
  if (!c || !d) return set_err(FL_ERR_INVALID, "null context or description");
  fl_terrain* T = nullptr;
  FLCHK(fl_terrain_create(c, d, &T));
  int32_t rc = fl_terrain_solve(c, T, d->max_iteration, iterations_out);
  if (rc == FL_OK) rc = fl_terrain_download(c, T, elevations_out);
  fl_terrain_destroy(T);
  return rc;
}

extern "C" int32_t fl_generate_f64_noalloc(fl_ctx* c, const fl_terrain_desc* d, double* elevations_out,
                                           uint32_t* iterations_out) {
  if (!c || !d) return set_err(FL_ERR_INVALID, "null context or description");
  if (d->n > FL_MAX_SITES || d->nnz >= 0xffffffffull)
    return set_err(FL_ERR_TOO_LARGE, "n=%llu, nnz=%llu exceed 2^31-1 sites or 2^32-2 entries", (unsigned long long)d->n,
                   (unsigned long long)d->nnz);
  uint8_t* base = nullptr;
  uint64_t size = 0;
  std::mutex* call_mu = nullptr;
  FLCHK(fl_noalloc_scratch(c->device, &base, &size, &call_mu));
  bool clamp = false;
  if (d->has_max_slope)
    for (uint64_t i = 0; i < d->n && !clamp; i++) clamp = d->has_max_slope[i] != 0;
  const uint64_t need = union_device_bytes(d->n, d->nnz, 1, clamp) + scratch_bytes(d->n, d->nnz, d->n_outlets, 1);
  if (need > size)
    return set_err(FL_ERR_TOO_LARGE,
                   "fl_generate_f64_noalloc: this terrain (%llu sites, %llu entries) needs %llu bytes of device "
                   "scratch; the static scratch holds %llu",
                   (unsigned long long)d->n, (unsigned long long)d->nnz, (unsigned long long)need,
                   (unsigned long long)size);
  // the array is the device's, shared by every call on it
  std::lock_guard<std::mutex> serial(*call_mu);
  fl_arena region;
  region.kind = fl_arena::REGION;
  region.base = base;
  region.size = size;
  // launches only: a captured graph's instantiation may allocate
  c->forbid_graphs = true;
  fl_terrain* T = nullptr;
  int32_t rc = create_union(c, d, 1, &T, &region);
  if (rc == FL_OK) rc = fl_terrain_solve(c, T, d->max_iteration, iterations_out);
  if (rc == FL_OK) rc = fl_terrain_download(c, T, elevations_out);
  if (T) fl_terrain_destroy(T);
  c->forbid_graphs = false;
  return rc;
}

// ------------------------------------------------------------- stages ---

extern "C" int32_t fl_stage_receivers(fl_ctx* c, fl_terrain* T, const double* elev, uint32_t* next_out) {
  if (!c || !T || !elev || !next_out) return set_err(FL_ERR_INVALID, "null argument");
  ProfScope prof_scope(c->prof);
  TuneScope tune(c, T, effective_tuning(c, T->t.n, T->nterr));
  if (T->t.n == 0) return FL_OK;
  HIPCHK(hipSetDevice(c->device));
  FLCHK(upload_h(c, T, elev));
  fl_launch_clear_iter(c->stream, T->w, std::min<uint32_t>(T->prev_nb + 1, 2 * T->t.n + 2));
  fl_launch_receivers(c->stream, T->t, T->w, false, false);
  return download_idx_old(c, T, next_out, T->w.next);
}

extern "C" int32_t fl_stage_stream_tree(fl_ctx* c, fl_terrain* T, const double* elev, uint32_t* next_out,
                                        uint32_t* subroot_out, int32_t* flag) {
  if (!c || !T || !elev) return set_err(FL_ERR_INVALID, "null argument");
  ProfScope prof_scope(c->prof);
  TuneScope tune(c, T, effective_tuning(c, T->t.n, T->nterr));
  if (T->t.n == 0) return FL_OK;
  HIPCHK(hipSetDevice(c->device));
  FLCHK(upload_h(c, T, elev));
  fl_launch_clear_iter(c->stream, T->w, std::min<uint32_t>(T->prev_nb + 1, 2 * T->t.n + 2));
  fl_launch_receivers(c->stream, T->t, T->w, false, false);
  FLCHK(read_ctrl(c, T));
  bool lake = host_flag(c, FL_FLAG_LAKE) != 0;
  FLCHK(label_chain_ends(c, T));
  if (lake) {
    FLCHK(ensure_visit_order(c, T));
    fl_launch_lake_join(c->stream, T->t, T->w);
  }
  FLCHK(download_idx_old(c, T, next_out, T->w.next));
  FLCHK(download_idx_old(c, T, subroot_out, T->w.sub));
  if (flag) *flag = lake ? 1 : 0;
  return FL_OK;
}

extern "C" int32_t fl_stage_iteration(fl_ctx* c, fl_terrain* T, double* elev, uint32_t* next_out,
                                      double* area_out, double* response_out, int32_t* flag) {
  if (!c || !T || !elev) return set_err(FL_ERR_INVALID, "null argument");
  ProfScope prof_scope(c->prof);
  TuneScope tune(c, T, effective_tuning(c, T->t.n, T->nterr));
  if (T->t.n == 0) return FL_OK;
  HIPCHK(hipSetDevice(c->device));
  FLCHK(upload_h(c, T, elev));
  bool changed = false, lake = false;
  FLCHK(one_iteration(c, T, true, false, &changed, &lake));
  FLCHK(download_f64_old(c, T, elev, T->w.h));
  FLCHK(download_idx_old(c, T, next_out, T->w.next));
  FLCHK(download_f64_old(c, T, area_out, T->w.A));
  FLCHK(download_f64_old(c, T, response_out, T->w.tau));
  if (flag) *flag = changed ? 1 : 0;
  return FL_OK;
}

extern "C" int32_t fl_ctx_set_batch_compaction(fl_ctx* c, uint32_t min_idle_sites) {
  if (!c) return set_err(FL_ERR_INVALID, "null context");
  c->tune.compact_min_idle = min_idle_sites;
  c->tune_set |= 4;
  return FL_OK;
}

extern "C" int32_t fl_ctx_set_pipelining(fl_ctx* c, int32_t enable) {
  if (!c) return set_err(FL_ERR_INVALID, "null context");
  if (enable < 0 || enable > 3) return set_err(FL_ERR_INVALID, "pipelining mode must be 0, 1, 2 or 3");
  // mode 2 (minimal speculation, for tests) is held in the context's field
  c->tune.pipelining = (uint32_t)enable;
  c->tune_set |= 2;
  return FL_OK;
}

extern "C" int32_t fl_ctx_set_recompute_all(fl_ctx* c, int32_t all) {
  if (!c) return set_err(FL_ERR_INVALID, "null context");
  c->recompute_all = all != 0;
  return FL_OK;
}

extern "C" int32_t fl_ctx_set_rehint_rounds(fl_ctx* c, uint32_t rounds) {
  if (!c) return set_err(FL_ERR_INVALID, "null context");
  c->rehint_rounds = rounds;
  return FL_OK;
}

extern "C" int32_t fl_ctx_set_tuning(fl_ctx* c, const fl_tuning* t) {
  if (!c) return set_err(FL_ERR_INVALID, "null context");
  if (!t) {
    c->tune_set &= ~1u;
    return FL_OK;
  }
  if (!valid_tuning(*t)) return set_err(FL_ERR_INVALID, "tuning out of range");
  c->tune = *t;
  c->tune_set |= 1;
  return FL_OK;
}

extern "C" int32_t fl_autotune(fl_ctx* c, fl_terrain* T, uint32_t repeats, fl_tuning* best, double* best_ms,
                               double* table_ms, uint32_t* candidates) {
  if (!c || !T || !best) return set_err(FL_ERR_INVALID, "null argument");
  if (T->t.n == 0) return set_err(FL_ERR_INVALID, "empty terrain");
  HIPCHK(hipSetDevice(c->device));
  if (repeats == 0) repeats = 1;
  const fl_tuning saved = c->tune;
  const uint32_t saved_set = c->tune_set;
  const uint32_t G = T->nterr;
  const uint64_t n = T->t.n;
  std::vector<uint64_t> ref_bits, bits(n);
  std::vector<uint32_t> ref_iters;
  uint32_t tried = 0;
  // the fastest of `repeats` solves with tuning t, its results checked
  // against the first tuning's bit for bit
  auto run = [&](const fl_tuning& t, double* ms) -> int32_t {
    double fastest = 1e300;
    for (uint32_t r = 0; r < repeats; r++) {
      const auto t0 = std::chrono::steady_clock::now();
      {
        TuneScope scope(c, T, t);
        FLCHK(solve_union(c, T, T->t_max_it.data()));
      }
      fastest = std::min(fastest, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    HIPCHK(hipMemcpyAsync(bits.data(), T->w.h, n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (ref_bits.empty()) {
      ref_bits = bits;
      ref_iters = T->t_iters;
    } else if (bits != ref_bits || T->t_iters != ref_iters) {
      return set_err(FL_ERR_HIP, "tuning (block %u, long chains %u, tail %u, pipelining %u) changed the results",
                     t.site_block, t.long_chain, t.tail_units, t.pipelining);
    }
    *ms = fastest;
    tried++;
    return FL_OK;
  };
  const fl_tuning table = effective_tuning(c, n, G);
  double ms = 0.0, cur_ms = 0.0;
  int32_t rc = run(table, &ms);  // warm-up: the visit order is computed and kept
  if (rc == FL_OK) rc = run(table, &cur_ms);
  fl_tuning cur = table;
  const double table_time = cur_ms;
  // coordinate descent, two passes; a candidate replaces the current one
  // only when it is faster by more than 1% (timing noise)
  const std::vector<uint32_t> chains = {3, 4, 6, 8, 12, 16, 24}, tails = {0, 8, 16, 32, 64},
                              blocks = {128, 256, 512, 1024}, pipes = {1, 3},
                              compacts = {16384, 65536, 262144, 0xffffffffu};
  for (int pass = 0; pass < 2 && rc == FL_OK; pass++) {
    auto sweep = [&](uint32_t fl_tuning::*field, const std::vector<uint32_t>& values) {
      for (uint32_t v : values) {
        if (rc != FL_OK || v == cur.*field) continue;
        fl_tuning t = cur;
        t.*field = v;
        rc = run(t, &ms);
        if (rc == FL_OK && ms < 0.99 * cur_ms) {
          cur = t;
          cur_ms = ms;
        }
      }
    };
    sweep(&fl_tuning::long_chain, chains);
    sweep(&fl_tuning::tail_units, tails);
    sweep(&fl_tuning::site_block, blocks);
    sweep(&fl_tuning::pipelining, pipes);
    if (G > 1) sweep(&fl_tuning::compact_min_idle, compacts);
  }
  c->tune = saved;
  c->tune_set = saved_set;
  if (rc != FL_OK) return rc;
  *best = cur;
  if (best_ms) *best_ms = cur_ms;
  if (table_ms) *table_ms = table_time;
  if (candidates) *candidates = tried;
  return FL_OK;
}

extern "C" int32_t fl_ctx_tuning_for(fl_ctx* c, uint64_t sites, uint32_t terrains, fl_tuning* out) {
  if (!c || !out) return set_err(FL_ERR_INVALID, "null argument");
  *out = effective_tuning(c, sites, terrains);
  return FL_OK;
}

extern "C" int32_t fl_terrain_invalidate_cache(fl_terrain* T) {
  if (!T) return set_err(FL_ERR_INVALID, "null terrain");
  T->have_order = false;
  return FL_OK;
}

extern "C" int32_t fl_ctx_set_profiling(fl_ctx* c, int32_t enable) {
  if (!c) return set_err(FL_ERR_INVALID, "null context");
  if (enable && !c->prof) c->prof = new fl_profiler();
  if (!enable && c->prof) {
    delete c->prof;
    c->prof = nullptr;
  }
  return FL_OK;
}

extern "C" int32_t fl_ctx_kernel_time(fl_ctx* c, const char* name, uint64_t* launches, double* total_ms,
                                      uint64_t* units) {
  if (!c || !name || !launches || !total_ms) return set_err(FL_ERR_INVALID, "null argument");
  *launches = 0;
  *total_ms = 0.0;
  if (units) *units = 0;
  if (!c->prof) return FL_OK;
  c->prof->flush();
  auto it = c->prof->totals.find(name);
  if (it != c->prof->totals.end()) {
    *launches = it->second.launches;
    *total_ms = it->second.ms;
    if (units) *units = it->second.units;
  }
  return FL_OK;
}

extern "C" int32_t fl_ctx_reset_profile(fl_ctx* c) {
  if (!c) return set_err(FL_ERR_INVALID, "null context");
  if (c->prof) c->prof->reset();
  return FL_OK;
}

extern "C" int32_t fl_terrain_last_stats(const fl_terrain* T, fl_solve_stats* out) {
  if (!T || !out) return set_err(FL_ERR_INVALID, "null argument");
  out->iterations = T->stat_iterations;
  out->lake_iterations = T->stat_lake_iterations;
  out->host_syncs = T->stat_syncs;
  out->order_from_forest = T->stat_order_fast;
  out->visit_order_ms = T->stat_order_ms;
  out->outlet_phase_ms = T->stat_order_parts[0];
  out->forest_ms = T->stat_order_parts[1];
  out->ranks_ms = T->stat_order_parts[2];
  out->solve_ms = T->stat_solve_ms;
  out->compactions = T->stat_compactions;
  out->redone = T->stat_redone;
  out->create_validate_ms = T->stat_create[0];
  out->create_relabel_ms = T->stat_create[1];
  out->create_upload_ms = T->stat_create[2];
  out->create_on_device = T->stat_create_device;
  out->create_clusters = T->stat_create_clusters;
  out->recomputed_sites = T->stat_recomputed;
  out->first_sweep_rounds = T->stat_rounds_first;
  out->max_sweep_rounds = T->stat_rounds_max;
  out->rehinted = T->stat_rehinted;
  return FL_OK;
}

extern "C" int32_t fl_stage_visit_order(fl_ctx* c, fl_terrain* T, uint32_t* rank_out) {
  if (!c || !T || (!rank_out && T->t.n)) return set_err(FL_ERR_INVALID, "null argument");
  ProfScope prof_scope(c->prof);
  TuneScope tune(c, T, effective_tuning(c, T->t.n, T->nterr));
  const uint32_t n = T->t.n;
  if (n == 0) return FL_OK;
  HIPCHK(hipSetDevice(c->device));
  FLCHK(ensure_visit_order(c, T));
  std::vector<uint32_t> tmp(n);
  HIPCHK(hipMemcpyAsync(tmp.data(), T->t.rank, (uint64_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  // ranks of terrain g start at site_off[g]: report positions within it
  for (uint32_t g = 0; g < T->nterr; g++)
    for (uint32_t i = T->site_off[g]; i < T->site_off[g + 1]; i++) {
      const uint32_t r = tmp[T->new_of_old[i]];
      rank_out[i] = r == 0xffffffffu ? r : r - T->site_off[g];
    }
  return FL_OK;
}

extern "C" int32_t fl_request_device_bytes(const fl_terrain_desc* d, uint64_t* bytes) {
  if (!d || !bytes) return set_err(FL_ERR_INVALID, "null argument");
  if (d->n > FL_MAX_SITES || d->nnz >= 0xffffffffull) return set_err(FL_ERR_TOO_LARGE, "terrain too large");
  bool clamp = false;
  if (d->has_max_slope)
    for (uint64_t i = 0; i < d->n && !clamp; i++) clamp = d->has_max_slope[i] != 0;
  *bytes = union_device_bytes(d->n, d->nnz, 1, clamp) + scratch_bytes(d->n, d->nnz, d->n_outlets, 1);
  return FL_OK;
}

extern "C" int32_t fl_stage_basin_order(fl_ctx* c, fl_terrain* T, const double* elev, uint32_t outlet,
                                        uint32_t* order_out, uint32_t* len_out) {
  if (!c || !T || !elev || !order_out || !len_out) return set_err(FL_ERR_INVALID, "null argument");
  if (outlet >= T->t.n) return set_err(FL_ERR_INVALID, "outlet out of range");
  ProfScope prof_scope(c->prof);
  TuneScope tune(c, T, effective_tuning(c, T->t.n, T->nterr));
  HIPCHK(hipSetDevice(c->device));
  uint32_t n = T->t.n;
  FLCHK(upload_h(c, T, elev));
  hipStream_t s = c->stream;
  fl_launch_clear_iter(s, T->w, std::min<uint32_t>(T->prev_nb + 1, 2 * n + 2));
  fl_launch_receivers(s, T->t, T->w, false, false);
  FLCHK(read_ctrl(c, T));
  if (host_flag(c, FL_FLAG_LAKE)) FLCHK(join_lakes(c, T));
  fl_launch_children(s, T->t, T->w, true, nullptr, nullptr);  // child masks of the final tree
  // scratch: the chain-major permutation buffer and one counter
  uint32_t* d_order = T->w.perm;
  uint32_t* d_len = T->w.ctrl + FL_CTRL_JUMPS;
  fl_launch_basin_order(s, T->t, T->w, T->new_of_old[outlet], d_order, d_len);
  uint32_t len = 0;
  HIPCHK(hipMemcpyAsync(&len, d_len, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  std::vector<uint32_t> tmp(len);
  if (len) HIPCHK(hipMemcpyAsync(tmp.data(), d_order, (uint64_t)len * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (uint32_t q = 0; q < len; q++) order_out[q] = T->old_of_new[tmp[q]];
  *len_out = len;
  return FL_OK;
}

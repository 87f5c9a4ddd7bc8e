// fl_capi.cpp -- C ABI of the device path: contexts, terrains resident in
// HBM, and the outer iteration loop of generate() (src/lem/generator.rs:
// 140-210) driven from the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fl_internal.h"
#include "fl_kernels.h"

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
  fl_profiler* prof = nullptr;  // non-null while profiling is enabled
  hipStream_t stream = nullptr;
  int* host_flags = nullptr;        // pinned readback of flags
  uint32_t* host_counters = nullptr;
  std::vector<uint32_t> host_boff;
};

struct fl_terrain {
  fl_ctx* ctx = nullptr;
  fl_dev_terrain t;
  fl_dev_work w;
  uint32_t max_iteration = FL_UNBOUNDED_ITERATIONS;
  uint32_t prev_nb = 0;
  bool have_order = false;
  uint64_t bytes = 0;
  // host copies needed for the one-time visit order
  std::vector<uint32_t> h_row_ptr, h_col, h_outlets;
  std::vector<double> h_len;
  std::vector<void*> allocations;
};

namespace {

template <typename T>
int32_t dev_alloc(fl_terrain* T_, T** p, uint64_t count) {
  size_t bytes = std::max<uint64_t>(count, 1) * sizeof(T);
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e != hipSuccess)
    return set_err(e == hipErrorOutOfMemory ? FL_ERR_OUT_OF_MEMORY : FL_ERR_HIP, "hipMalloc(%zu): %s", bytes,
                   hipGetErrorString(e));
  T_->allocations.push_back((void*)*p);
  T_->bytes += bytes;
  return FL_OK;
}

int32_t validate(const fl_terrain_desc* d) {
  if (!d) return set_err(FL_ERR_INVALID, "null terrain description");
  if (d->n >= 0xffffffffull) return set_err(FL_ERR_TOO_LARGE, "n=%llu exceeds 2^32-2 sites", (unsigned long long)d->n);
  if (d->nnz >= 0xffffffffull) return set_err(FL_ERR_TOO_LARGE, "nnz exceeds 2^32-2 entries");
  if (d->n == 0) return FL_OK;
  if (!d->row_ptr || (d->nnz && (!d->col || !d->length)) || !d->area || !d->erodibility || !d->uplift_rate ||
      !d->init_elevation)
    return set_err(FL_ERR_INVALID, "null array in terrain description");
  if (d->n_outlets && !d->outlets) return set_err(FL_ERR_INVALID, "null outlets");
  if (d->has_max_slope && !d->max_slope) return set_err(FL_ERR_INVALID, "has_max_slope without max_slope");
  uint64_t n = d->n;
  if (d->row_ptr[0] != 0 || d->row_ptr[n] != d->nnz) return set_err(FL_ERR_INVALID, "row_ptr must span [0, nnz]");
  for (uint64_t i = 0; i < n; i++)
    if (d->row_ptr[i + 1] < d->row_ptr[i]) return set_err(FL_ERR_INVALID, "row_ptr decreases at %llu",
                                                          (unsigned long long)i);
  // neighbour indices, self-loops, duplicates (the reference's BFS would
  // visit such a site twice), symmetry (terrain-graph is undirected)
  std::vector<uint32_t> mark(n, 0xffffffffu);
  for (uint64_t i = 0; i < n; i++) {
    for (uint32_t k = d->row_ptr[i]; k < d->row_ptr[i + 1]; k++) {
      uint32_t j = d->col[k];
      if (j >= n) return set_err(FL_ERR_INVALID, "col[%u]=%u out of range", k, j);
      if (j == i) return set_err(FL_ERR_INVALID, "self-loop at site %llu", (unsigned long long)i);
      if (mark[j] == (uint32_t)i) return set_err(FL_ERR_INVALID, "duplicate edge %llu-%u", (unsigned long long)i, j);
      mark[j] = (uint32_t)i;
    }
  }
  for (uint64_t i = 0; i < n; i++) {
    for (uint32_t k = d->row_ptr[i]; k < d->row_ptr[i + 1]; k++) {
      uint32_t j = d->col[k];
      bool found = false;
      for (uint32_t q = d->row_ptr[j]; q < d->row_ptr[j + 1]; q++)
        if (d->col[q] == i) {
          found = true;
          break;
        }
      if (!found) return set_err(FL_ERR_INVALID, "edge %llu-%u is not symmetric", (unsigned long long)i, j);
    }
  }
  std::vector<uint8_t> seen(n, 0);
  for (uint64_t k = 0; k < d->n_outlets; k++) {
    if (d->outlets[k] >= n) return set_err(FL_ERR_INVALID, "outlet %u out of range", d->outlets[k]);
    if (seen[d->outlets[k]]) return set_err(FL_ERR_INVALID, "duplicate outlet %u", d->outlets[k]);
    seen[d->outlets[k]] = 1;
  }
  return FL_OK;
}

int32_t read_flags(fl_ctx* c, const fl_terrain* T) {
  HIPCHK(hipMemcpyAsync(c->host_flags, T->w.flags, FL_NUM_FLAGS * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return FL_OK;
}

int32_t clear_flags(fl_ctx* c, fl_terrain* T) {
  fl_launch_fill_u32(c->stream, (uint32_t*)T->w.flags, FL_NUM_FLAGS, 0);
  return FL_OK;
}

int32_t ensure_visit_order(fl_ctx* c, fl_terrain* T) {
  if (T->have_order) return FL_OK;
  uint32_t n = T->t.n;
  std::vector<uint32_t> rank(n), order(n, 0xffffffffu);
  uint64_t cnt = 0;
  int rc = fl_lake_visit_order(n, T->h_row_ptr.data(), T->h_col.data(), T->h_len.data(), T->h_outlets.data(),
                               T->h_outlets.size(), rank.data(), &cnt);
  if (rc == FL_ERR_NAN_LENGTH) return set_err(FL_ERR_NAN_LENGTH, "NaN edge length reached by the lake traversal");
  if (rc != FL_OK) return set_err(rc, "visit order failed");
  for (uint32_t i = 0; i < n; i++)
    if (rank[i] != 0xffffffffu) order[rank[i]] = i;
  HIPCHK(hipMemcpyAsync(T->t.rank, rank.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(T->t.order, order.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  T->have_order = true;
  return FL_OK;
}

// Receivers and, if any lake exists, lake joining (StreamTree::construct,
// stream_tree.rs:72-99).  want_sub forces the chain-end labels (for the
// stage API) even without lakes.
int32_t stream_tree(fl_ctx* c, fl_terrain* T, bool want_sub, bool* had_lake) {
  hipStream_t s = c->stream;
  FLCHK(clear_flags(c, T));
  fl_launch_receivers(s, T->t, T->w);
  FLCHK(read_flags(c, T));
  bool lake = c->host_flags[FL_FLAG_LAKE] != 0;
  *had_lake = lake;
  if (!lake && !want_sub) return FL_OK;
  fl_launch_sub_init(s, T->t, T->w);
  bool flip = false;
  for (uint32_t round = 0;; round++) {
    fl_launch_fill_u32(s, (uint32_t*)T->w.flags + FL_FLAG_JUMP, 1, 0);
    fl_launch_sub_jump(s, T->t, T->w, flip);
    flip = !flip;
    FLCHK(read_flags(c, T));
    if (!c->host_flags[FL_FLAG_JUMP]) break;
    if (round > 64) return set_err(FL_ERR_HIP, "chain-end labelling did not converge");
  }
  if (flip) std::swap(T->w.sub, T->w.sub2);
  if (!lake) return FL_OK;
  FLCHK(ensure_visit_order(c, T));
  fl_launch_lake_join(s, T->t, T->w);
  return FL_OK;
}

// Everything after the stream tree: decomposition and the two sweeps.
int32_t tree_sweeps(fl_ctx* c, fl_terrain* T, bool* changed) {
  hipStream_t s = c->stream;
  fl_dev_work& w = T->w;
  uint32_t n = T->t.n;
  fl_launch_fill_u32(s, w.counters, FL_NUM_COUNTERS, 0);
  fl_launch_children(s, T->t, w);
  bool flip = false;
  for (uint32_t round = 0;; round++) {
    fl_launch_fill_u32(s, (uint32_t*)w.flags + FL_FLAG_JUMP, 1, 0);
    fl_launch_chain_jump(s, T->t, w, flip);
    flip = !flip;
    if (round % 4 == 3) {
      FLCHK(read_flags(c, T));
      if (!c->host_flags[FL_FLAG_JUMP]) break;
    }
    if (round > 80) return set_err(FL_ERR_HIP, "tree decomposition did not converge");
  }
  if (flip) {
    std::swap(w.jc, w.jc2);
    std::swap(w.pc, w.pc2);
    std::swap(w.jn, w.jn2);
    std::swap(w.lw, w.lw2);
  }
  uint32_t clear = std::min<uint32_t>(T->prev_nb + 1, n + 1);
  fl_launch_fill_u32(s, w.bcnt, clear, 0);
  fl_launch_fill_u32(s, w.bnodes, clear, 0);
  fl_launch_fill_u32(s, w.bcur, clear, 0);
  fl_launch_fill_u32(s, w.bncur, clear, 0);
  fl_launch_chain_count(s, T->t, w);
  HIPCHK(hipMemcpyAsync(c->host_counters, w.counters, FL_NUM_COUNTERS * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  uint32_t nb = c->host_counters[FL_CNT_MAXLD] + 1;
  if (nb > n + 1) return set_err(FL_ERR_HIP, "light depth out of range");
  T->prev_nb = nb;
  fl_launch_chain_layout(s, T->t, w, nb);
  c->host_boff.resize(nb + 1);
  HIPCHK(hipMemcpyAsync(c->host_boff.data(), w.boff, (nb + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const std::vector<uint32_t>& bo = c->host_boff;
  for (uint32_t l = nb; l-- > 0;) fl_launch_da_round(s, T->t, w, bo[l], bo[l + 1]);
  fl_launch_celerity(s, T->t, w);
  for (uint32_t l = 0; l < nb; l++) fl_launch_tau_round(s, T->t, w, bo[l], bo[l + 1]);
  FLCHK(read_flags(c, T));
  *changed = c->host_flags[FL_FLAG_CHANGED] != 0;
  return FL_OK;
}

int32_t one_iteration(fl_ctx* c, fl_terrain* T, bool* changed, bool* had_lake) {
  FLCHK(stream_tree(c, T, false, had_lake));
  return tree_sweeps(c, T, changed);
}

int32_t reset_state(fl_ctx* c, fl_terrain* T) {
  uint32_t n = T->t.n;
  HIPCHK(hipMemcpyAsync(T->w.h, T->t.init_h, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(T->w.A, T->t.area, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  return FL_OK;
}

}  // namespace

// ------------------------------------------------------------------ API ---

extern "C" const char* fl_last_error(void) { return g_err.c_str(); }

extern "C" const char* fl_version(void) { return "fastlem_hip 0.1.0 (gfx950, f64)"; }

extern "C" int32_t fl_ctx_create(int32_t device, fl_ctx** out) {
  if (!out) return set_err(FL_ERR_INVALID, "null output pointer");
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return set_err(FL_ERR_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= count) return set_err(FL_ERR_INVALID, "device %d out of range", device);
  std::unique_ptr<fl_ctx> c(new fl_ctx());
  c->device = device;
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPCHK(hipHostMalloc((void**)&c->host_flags, 64 * sizeof(int), hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void**)&c->host_counters, 64 * sizeof(uint32_t), hipHostMallocDefault));
  *out = c.release();
  return FL_OK;
}

extern "C" int32_t fl_ctx_destroy(fl_ctx* c) {
  if (!c) return FL_OK;
  hipSetDevice(c->device);
  delete c->prof;
  if (c->stream) hipStreamDestroy(c->stream);
  if (c->host_flags) hipHostFree(c->host_flags);
  if (c->host_counters) hipHostFree(c->host_counters);
  delete c;
  return FL_OK;
}

extern "C" int32_t fl_terrain_destroy(fl_terrain* T) {
  if (!T) return FL_OK;
  if (T->ctx) hipSetDevice(T->ctx->device);
  for (void* p : T->allocations) hipFree(p);
  delete T;
  return FL_OK;
}

extern "C" int32_t fl_terrain_create(fl_ctx* c, const fl_terrain_desc* d, fl_terrain** out) {
  if (!c || !out) return set_err(FL_ERR_INVALID, "null context or output");
  FLCHK(validate(d));
  HIPCHK(hipSetDevice(c->device));
  std::unique_ptr<fl_terrain, int32_t (*)(fl_terrain*)> T(new fl_terrain(), fl_terrain_destroy);
  T->ctx = c;
  T->max_iteration = d->max_iteration;
  uint32_t n = (uint32_t)d->n, nnz = (uint32_t)d->nnz;
  T->t.n = n;
  T->t.nnz = nnz;
  if (n == 0) {
    *out = T.release();
    return FL_OK;
  }
  T->h_row_ptr.assign(d->row_ptr, d->row_ptr + n + 1);
  T->h_col.assign(d->col, d->col + nnz);
  T->h_len.assign(d->length, d->length + nnz);
  T->h_outlets.assign(d->outlets, d->outlets + d->n_outlets);
  fl_dev_terrain& t = T->t;
  fl_dev_work& w = T->w;
  fl_terrain* P = T.get();
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
  for (double** p : {&w.h, &w.A, &w.tau, &w.x, &w.dnext}) FLCHK(dev_alloc(P, p, n));
  for (uint32_t** p : {&w.next, &w.sub, &w.sub2, &w.coff, &w.ccnt, &w.clist, &w.hc, &w.jc, &w.pc, &w.jn, &w.lw,
                       &w.jc2, &w.pc2, &w.jn2, &w.lw2, &w.clen, &w.chain_base, &w.chain_len, &w.head_base, &w.perm})
    FLCHK(dev_alloc(P, p, n));
  for (uint32_t** p : {&w.bcnt, &w.bnodes, &w.boff, &w.bnoff, &w.bcur, &w.bncur}) FLCHK(dev_alloc(P, p, (uint64_t)n + 2));
  FLCHK(dev_alloc(P, &w.lake_key, n));
  FLCHK(dev_alloc(P, &w.flags, FL_NUM_FLAGS));
  FLCHK(dev_alloc(P, &w.counters, FL_NUM_COUNTERS));

  std::vector<double> tan_ms(n);
  std::vector<uint8_t> is_out(n, 0);
  for (uint32_t i = 0; i < n; i++)
    tan_ms[i] = (d->has_max_slope && d->has_max_slope[i]) ? std::tan(d->max_slope[i]) : std::nan("");
  for (uint64_t k = 0; k < d->n_outlets; k++) is_out[d->outlets[k]] = 1;
  hipStream_t s = c->stream;
  HIPCHK(hipMemcpyAsync(t.row_ptr, d->row_ptr, ((uint64_t)n + 1) * 4, hipMemcpyHostToDevice, s));
  if (nnz) {
    HIPCHK(hipMemcpyAsync(t.col, d->col, (uint64_t)nnz * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(t.len, d->length, (uint64_t)nnz * 8, hipMemcpyHostToDevice, s));
  }
  HIPCHK(hipMemcpyAsync(t.area, d->area, (uint64_t)n * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(t.erod, d->erodibility, (uint64_t)n * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(t.uplift, d->uplift_rate, (uint64_t)n * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(t.tan_ms, tan_ms.data(), (uint64_t)n * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(t.init_h, d->init_elevation, (uint64_t)n * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(t.is_outlet, is_out.data(), n, hipMemcpyHostToDevice, s));
  fl_launch_fill_u32(s, w.bcnt, (uint64_t)n + 2, 0);
  fl_launch_fill_u32(s, w.bnodes, (uint64_t)n + 2, 0);
  fl_launch_fill_u32(s, w.bcur, (uint64_t)n + 2, 0);
  fl_launch_fill_u32(s, w.bncur, (uint64_t)n + 2, 0);
  T->prev_nb = 0;
  FLCHK(reset_state(c, P));
  HIPCHK(hipStreamSynchronize(s));
  *out = T.release();
  return FL_OK;
}

extern "C" int32_t fl_terrain_solve(fl_ctx* c, fl_terrain* T, uint32_t max_iteration, uint32_t* iterations_out) {
  if (!c || !T) return set_err(FL_ERR_INVALID, "null context or terrain");
  ProfScope prof_scope(c->prof);
  HIPCHK(hipSetDevice(c->device));
  uint32_t it = 0;
  if (T->t.n == 0) {
    // the loop body runs once and finds nothing changed (generator.rs:140-209)
    if (iterations_out) *iterations_out = max_iteration > 0 ? 1 : 0;
    return FL_OK;
  }
  FLCHK(reset_state(c, T));
  for (uint32_t step = 0; step < max_iteration; step++) {
    bool changed = false, lake = false;
    FLCHK(one_iteration(c, T, &changed, &lake));
    it++;
    if (!changed) break;
  }
  if (iterations_out) *iterations_out = it;
  return FL_OK;
}

extern "C" int32_t fl_terrain_download(fl_ctx* c, fl_terrain* T, double* out) {
  if (!c || !T || (!out && T->t.n)) return set_err(FL_ERR_INVALID, "null argument");
  if (T->t.n == 0) return FL_OK;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(out, T->w.h, (uint64_t)T->t.n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
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

// ------------------------------------------------------------- stages ---

namespace {
int32_t upload_h(fl_ctx* c, fl_terrain* T, const double* elev) {
  HIPCHK(hipMemcpyAsync(T->w.h, elev, (uint64_t)T->t.n * 8, hipMemcpyHostToDevice, c->stream));
  return FL_OK;
}
int32_t download_u32(fl_ctx* c, uint32_t* dst, const uint32_t* src, uint32_t n) {
  if (!dst) return FL_OK;
  HIPCHK(hipMemcpyAsync(dst, src, (uint64_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  return FL_OK;
}
int32_t download_f64(fl_ctx* c, double* dst, const double* src, uint32_t n) {
  if (!dst) return FL_OK;
  HIPCHK(hipMemcpyAsync(dst, src, (uint64_t)n * 8, hipMemcpyDeviceToHost, c->stream));
  return FL_OK;
}
}  // namespace

extern "C" int32_t fl_stage_receivers(fl_ctx* c, fl_terrain* T, const double* elev, uint32_t* next_out) {
  if (!c || !T || !elev || !next_out) return set_err(FL_ERR_INVALID, "null argument");
  ProfScope prof_scope(c->prof);
  if (T->t.n == 0) return FL_OK;
  HIPCHK(hipSetDevice(c->device));
  FLCHK(upload_h(c, T, elev));
  FLCHK(clear_flags(c, T));
  fl_launch_receivers(c->stream, T->t, T->w);
  FLCHK(download_u32(c, next_out, T->w.next, T->t.n));
  HIPCHK(hipStreamSynchronize(c->stream));
  return FL_OK;
}

extern "C" int32_t fl_stage_stream_tree(fl_ctx* c, fl_terrain* T, const double* elev, uint32_t* next_out,
                                        uint32_t* subroot_out, int32_t* flag) {
  if (!c || !T || !elev) return set_err(FL_ERR_INVALID, "null argument");
  ProfScope prof_scope(c->prof);
  if (T->t.n == 0) return FL_OK;
  HIPCHK(hipSetDevice(c->device));
  FLCHK(upload_h(c, T, elev));
  bool lake = false;
  FLCHK(stream_tree(c, T, true, &lake));
  FLCHK(download_u32(c, next_out, T->w.next, T->t.n));
  FLCHK(download_u32(c, subroot_out, T->w.sub, T->t.n));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (flag) *flag = lake ? 1 : 0;
  return FL_OK;
}

extern "C" int32_t fl_stage_iteration(fl_ctx* c, fl_terrain* T, double* elev, uint32_t* next_out,
                                      double* area_out, double* response_out, int32_t* flag) {
  if (!c || !T || !elev) return set_err(FL_ERR_INVALID, "null argument");
  ProfScope prof_scope(c->prof);
  if (T->t.n == 0) return FL_OK;
  HIPCHK(hipSetDevice(c->device));
  FLCHK(upload_h(c, T, elev));
  bool changed = false, lake = false;
  FLCHK(one_iteration(c, T, &changed, &lake));
  FLCHK(download_f64(c, elev, T->w.h, T->t.n));
  FLCHK(download_u32(c, next_out, T->w.next, T->t.n));
  FLCHK(download_f64(c, area_out, T->w.A, T->t.n));
  FLCHK(download_f64(c, response_out, T->w.tau, T->t.n));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (flag) *flag = changed ? 1 : 0;
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

extern "C" int32_t fl_ctx_kernel_time(fl_ctx* c, const char* name, uint64_t* launches, double* total_ms) {
  if (!c || !name || !launches || !total_ms) return set_err(FL_ERR_INVALID, "null argument");
  *launches = 0;
  *total_ms = 0.0;
  if (!c->prof) return FL_OK;
  c->prof->flush();
  auto it = c->prof->totals.find(name);
  if (it != c->prof->totals.end()) {
    *launches = it->second.first;
    *total_ms = it->second.second;
  }
  return FL_OK;
}

extern "C" int32_t fl_ctx_reset_profile(fl_ctx* c) {
  if (!c) return set_err(FL_ERR_INVALID, "null context");
  if (c->prof) c->prof->reset();
  return FL_OK;
}

// fl_kernels.hip -- gfx950 kernels of the uplift-erosion iteration.
//
// One outer iteration of fastlem's generate() loop (src/lem/generator.rs:
// 140-210) runs as:
//
//   k_receivers         stream_tree.rs:109-137  steepest-descent receiver,
//                                               first-match edge length
//   lake join (only when a non-outlet site has no receiver):
//     k_sub_*           stream_tree.rs:139-173  chain ends by pointer jumping
//     k_lake_keys       stream_tree.rs:175-243  per lake: 64-bit atomicMin on
//                                               (visit rank << 32 | slot)
//     k_lake_flip                               reverse the lake's flow path
//   tree decomposition (heavy paths, rebuilt every iteration):
//     k_children        children of each site in row (slot) order
//     k_chain_init/jump chain head, position in chain, light depth, root
//     k_chain_*         chains bucketed by light depth, laid out contiguously
//   sweeps
//     k_da_round        generator.rs:154-159  drainage area, exact fold order
//     k_celerity        generator.rs:172      1/(K*A^0.5) * distance
//     k_tau_round       generator.rs:162-203  response time, elevation,
//                                             max-slope clamp, changed flag
//
// Exactness.  Every floating-point expression is evaluated in the order the
// reference writes it; the library is built with -ffp-contract=off so no
// multiply-add is fused.  No floating-point atomics are used: the drainage
// area of site u is folded by one lane as
//     A_u = ((a_u + A_{c_m}) + A_{c_{m-1}}) + ... + A_{c_1}
// (children c_1..c_m in row order), which is the association the reference's
// reverse-BFS accumulation produces; response time is a sequential
// root-to-leaf prefix per chain.  Results are therefore bit-identical to the
// reference for the same inputs.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fl_kernels.h"

#define FL_NONE 0xffffffffu

namespace {

__device__ __forceinline__ uint32_t gid() { return blockIdx.x * blockDim.x + threadIdx.x; }

// f64::powf(x, 0.5) as lowered in optimised Rust builds.
__device__ __forceinline__ double rust_pow_half(double x) {
  return (x == -__builtin_inf()) ? __builtin_inf() : __builtin_fabs(__builtin_sqrt(x));
}

// f64::max: the non-NaN operand wins.
__device__ __forceinline__ double rust_max(double a, double b) { return __builtin_fmax(a, b); }

// terrain-graph has_edge(a, b): first entry of row a equal to b; 1.0 when
// absent (generator.rs:164-171).
__device__ __forceinline__ double first_len(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                                            const double* __restrict__ len, uint32_t a, uint32_t b) {
  uint32_t e1 = rp[a + 1];
  for (uint32_t k = rp[a]; k < e1; ++k)
    if (col[k] == b) return len[k];
  return 1.0;
}

}  // namespace

// ------------------------------------------------------------ receivers ---

// next[i]: the neighbour with the strictly largest positive slope
// (h_i - h_j)/len, earliest slot on ties; i itself for outlets and pits.
// flags[FL_FLAG_LAKE] is raised when a non-outlet site has no receiver.
__global__ void k_receivers(uint32_t n, const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                            const double* __restrict__ len, const double* __restrict__ h,
                            const uint8_t* __restrict__ is_outlet, uint32_t* __restrict__ next,
                            double* __restrict__ dnext, int* __restrict__ flags) {
  uint32_t i = gid();
  if (i >= n) return;
  if (is_outlet[i]) {
    next[i] = i;
    dnext[i] = 1.0;
    return;
  }
  const double hi = h[i];
  double best = 0.0;
  uint32_t bj = i;
  const uint32_t b = rp[i], e = rp[i + 1];
  for (uint32_t k = b; k < e; ++k) {
    uint32_t j = col[k];
    double hj = h[j];
    if (hi > hj) {
      double s = (hi - hj) / len[k];
      if (s > best) {
        best = s;
        bj = j;
      }
    }
  }
  double d = 1.0;
  if (bj != i) {
    for (uint32_t k = b; k < e; ++k)
      if (col[k] == bj) {
        d = len[k];
        break;
      }
  } else {
    flags[FL_FLAG_LAKE] = 1;
  }
  next[i] = bj;
  dnext[i] = d;
}

// ------------------------------------------------------------ lake join ---

__global__ void k_copy_u32(uint32_t n, const uint32_t* __restrict__ src, uint32_t* __restrict__ dst) {
  uint32_t i = gid();
  if (i < n) dst[i] = src[i];
}

__global__ void k_fill_u64(uint32_t n, uint64_t* __restrict__ dst, uint64_t v) {
  uint32_t i = gid();
  if (i < n) dst[i] = v;
}

__global__ void k_fill_u32(uint32_t n, uint32_t* __restrict__ dst, uint32_t v) {
  uint32_t i = gid();
  if (i < n) dst[i] = v;
}

// One pointer-jumping round: out[i] = in[in[i]]; raises *flag if anything moved.
__global__ void k_sub_jump(uint32_t n, const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                           int* __restrict__ flag) {
  uint32_t i = gid();
  if (i >= n) return;
  uint32_t s = in[i];
  uint32_t ss = in[s];
  out[i] = ss;
  if (ss != s) *flag = 1;
}

// For every reached site i and every neighbour j lying in a different
// chain whose end is a non-outlet pit (a lake), offer the key
// (visit rank of i) << 32 | slot to that lake.  The minimum key names the
// site i* and the slot of j* through which the reference's traversal first
// touches the lake (stream_tree.rs:204-230).
__global__ void k_lake_keys(uint32_t n, const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                            const uint32_t* __restrict__ rank, const uint32_t* __restrict__ sub,
                            const uint8_t* __restrict__ is_outlet, unsigned long long* __restrict__ key) {
  uint32_t i = gid();
  if (i >= n) return;
  uint32_t r = rank[i];
  if (r == FL_NONE) return;
  uint32_t si = sub[i];
  uint32_t b = rp[i], e = rp[i + 1];
  for (uint32_t k = b; k < e; ++k) {
    uint32_t s = sub[col[k]];
    if (s != si && !is_outlet[s]) {
      unsigned long long kv = ((unsigned long long)r << 32) | (unsigned long long)(k - b);
      atomicMin(key + s, kv);
    }
  }
}

// One thread per lake (indexed by its pit s): reverse the flow path from j*
// to the pit so that j* drains into i* (stream_tree.rs:214-229) and refresh
// the receiver distances of the reversed sites.  Lakes are disjoint, so all
// flips run concurrently without conflicts.
__global__ void k_lake_flip(uint32_t n, const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                            const double* __restrict__ len, const uint32_t* __restrict__ order,
                            const unsigned long long* __restrict__ key, const uint8_t* __restrict__ is_outlet,
                            uint32_t* __restrict__ next, double* __restrict__ dnext) {
  uint32_t s = gid();
  if (s >= n) return;
  if (is_outlet[s] || next[s] != s) return;
  unsigned long long kv = key[s];
  if (kv == ~0ull) return;  // never reached: stays a pit, as in the reference
  uint32_t istar = order[(uint32_t)(kv >> 32)];
  uint32_t jstar = col[rp[istar] + (uint32_t)(kv & 0xffffffffu)];
  uint32_t k = jstar, nk = istar;
  for (uint32_t guard = 0; guard <= n; ++guard) {
    uint32_t nx = next[k];
    if (nx == k) break;
    next[k] = nk;
    dnext[k] = first_len(rp, col, len, k, nk);
    nk = k;
    k = nx;
  }
  next[k] = nk;
  dnext[k] = first_len(rp, col, len, k, nk);
}

// ------------------------------------------------------- decomposition ----

// Children of u are the neighbours v with next[v] == u, in row order (the
// order DrainageBasin::construct appends them, drainage_basin.rs:23-27).
// The heavy child (first with the largest previous drainage area) continues
// u's chain; the choice only shapes the work, never the results.  Child
// segments are allocated with one atomic per wavefront.
__global__ void k_children(uint32_t n, const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                           const uint32_t* __restrict__ next, const double* __restrict__ prev_area,
                           uint32_t* __restrict__ coff, uint32_t* __restrict__ ccnt, uint32_t* __restrict__ clist,
                           uint32_t* __restrict__ hc, uint32_t* __restrict__ total) {
  uint32_t u = gid();
  uint32_t cnt = 0, heavy = FL_NONE;
  uint32_t b = 0, e = 0;
  if (u < n) {
    b = rp[u];
    e = rp[u + 1];
    double best = 0.0;
    for (uint32_t k = b; k < e; ++k) {
      uint32_t v = col[k];
      if (v != u && next[v] == u) {
        double av = prev_area[v];
        if (heavy == FL_NONE || av > best) {
          best = av;
          heavy = v;
        }
        ++cnt;
      }
    }
  }
  // wavefront-aggregated allocation of the child segments
  uint32_t lane = threadIdx.x & 63u;
  uint32_t incl = cnt;
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t t = __shfl_up(incl, off, 64);
    if ((int)lane >= off) incl += t;
  }
  uint32_t wave_total = __shfl(incl, 63, 64);
  uint32_t base = 0;
  if (lane == 63 && wave_total) base = atomicAdd(total, wave_total);
  base = __shfl(base, 63, 64);
  if (u >= n) return;
  uint32_t off = base + incl - cnt;
  coff[u] = off;
  ccnt[u] = cnt;
  hc[u] = heavy;
  if (cnt) {
    for (uint32_t k = b; k < e; ++k) {
      uint32_t v = col[k];
      if (v != u && next[v] == u) clist[off++] = v;
    }
  }
}

// Chain and root pointers for pointer jumping.
//   jc: chain parent (self at a chain head), pc: 1 (0 at a head)
//   jn: receiver, lw: 1 at a head that is not a root (a light edge)
__global__ void k_chain_init(uint32_t n, const uint32_t* __restrict__ next, const uint32_t* __restrict__ hc,
                             uint32_t* __restrict__ jc, uint32_t* __restrict__ pc, uint32_t* __restrict__ jn,
                             uint32_t* __restrict__ lw) {
  uint32_t v = gid();
  if (v >= n) return;
  uint32_t p = next[v];
  bool root = (p == v);
  bool head = root || hc[p] != v;
  jc[v] = head ? v : p;
  pc[v] = head ? 0u : 1u;
  jn[v] = p;
  lw[v] = (head && !root) ? 1u : 0u;
}

// One doubling round on both pointer systems.  After convergence
// jc[v] = chain head, pc[v] = distance to it, jn[v] = root of v's tree,
// lw[v] = light edges between v and its root (= light depth of v's chain).
__global__ void k_chain_jump(uint32_t n, const uint32_t* __restrict__ jc, const uint32_t* __restrict__ pc,
                             const uint32_t* __restrict__ jn, const uint32_t* __restrict__ lw,
                             uint32_t* __restrict__ jc2, uint32_t* __restrict__ pc2, uint32_t* __restrict__ jn2,
                             uint32_t* __restrict__ lw2, int* __restrict__ flag) {
  uint32_t v = gid();
  if (v >= n) return;
  uint32_t a = jc[v], b = jn[v];
  uint32_t aa = jc[a], bb = jn[b];
  jc2[v] = aa;
  pc2[v] = pc[v] + pc[a];
  jn2[v] = bb;
  lw2[v] = lw[v] + lw[b];
  if (aa != a || bb != b) *flag = 1;
}

// Per chain head inside an outlet's basin: chain length (from its tail, the
// member without a heavy child) and light-depth histogram.
__global__ void k_chain_len(uint32_t n, const uint32_t* __restrict__ hc, const uint32_t* __restrict__ jc,
                            const uint32_t* __restrict__ pc, const uint32_t* __restrict__ jn,
                            const uint8_t* __restrict__ is_outlet, uint32_t* __restrict__ clen) {
  uint32_t v = gid();
  if (v >= n) return;
  if (hc[v] == FL_NONE && is_outlet[jn[v]]) clen[jc[v]] = pc[v] + 1;
}

__global__ void k_bucket_count(uint32_t n, const uint32_t* __restrict__ jc, const uint32_t* __restrict__ jn,
                               const uint32_t* __restrict__ lw, const uint8_t* __restrict__ is_outlet,
                               const uint32_t* __restrict__ clen, uint32_t* __restrict__ bcnt,
                               uint32_t* __restrict__ bnodes, uint32_t* __restrict__ maxld) {
  uint32_t v = gid();
  if (v >= n) return;
  if (jc[v] != v || !is_outlet[jn[v]]) return;
  uint32_t l = lw[v];
  atomicAdd(bcnt + l, 1u);
  atomicAdd(bnodes + l, clen[v]);
  atomicMax(maxld, l);
}

// Exclusive scans of the (small) bucket histograms, one workgroup.
__global__ void k_bucket_scan(uint32_t nb, const uint32_t* __restrict__ bcnt, const uint32_t* __restrict__ bnodes,
                              uint32_t* __restrict__ boff, uint32_t* __restrict__ bnoff) {
  __shared__ uint32_t s_a[1024], s_b[1024];
  __shared__ uint32_t carry_a, carry_b;
  if (threadIdx.x == 0) {
    carry_a = 0;
    carry_b = 0;
  }
  __syncthreads();
  for (uint32_t base = 0; base < nb; base += blockDim.x) {
    uint32_t i = base + threadIdx.x;
    uint32_t a = i < nb ? bcnt[i] : 0, b = i < nb ? bnodes[i] : 0;
    s_a[threadIdx.x] = a;
    s_b[threadIdx.x] = b;
    __syncthreads();
    for (uint32_t off = 1; off < blockDim.x; off <<= 1) {
      uint32_t ta = threadIdx.x >= off ? s_a[threadIdx.x - off] : 0;
      uint32_t tb = threadIdx.x >= off ? s_b[threadIdx.x - off] : 0;
      __syncthreads();
      s_a[threadIdx.x] += ta;
      s_b[threadIdx.x] += tb;
      __syncthreads();
    }
    if (i < nb) {
      boff[i] = carry_a + s_a[threadIdx.x] - a;
      bnoff[i] = carry_b + s_b[threadIdx.x] - b;
    }
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) {
      carry_a += s_a[threadIdx.x];
      carry_b += s_b[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    boff[nb] = carry_a;
    bnoff[nb] = carry_b;
  }
}

// Give every chain a slot in its light-depth bucket and a contiguous node
// range.  Order inside a bucket is immaterial to the results.
__global__ void k_bucket_place(uint32_t n, const uint32_t* __restrict__ jc, const uint32_t* __restrict__ jn,
                               const uint32_t* __restrict__ lw, const uint8_t* __restrict__ is_outlet,
                               const uint32_t* __restrict__ clen, const uint32_t* __restrict__ boff,
                               const uint32_t* __restrict__ bnoff, uint32_t* __restrict__ bcur,
                               uint32_t* __restrict__ bncur, uint32_t* __restrict__ chain_base,
                               uint32_t* __restrict__ chain_len, uint32_t* __restrict__ head_base) {
  uint32_t v = gid();
  if (v >= n) return;
  if (jc[v] != v || !is_outlet[jn[v]]) return;
  uint32_t l = lw[v];
  uint32_t len = clen[v];
  uint32_t c = boff[l] + atomicAdd(bcur + l, 1u);
  uint32_t base = bnoff[l] + atomicAdd(bncur + l, len);
  chain_base[c] = base;
  chain_len[c] = len;
  head_base[v] = base;
}

__global__ void k_chain_perm(uint32_t n, const uint32_t* __restrict__ jc, const uint32_t* __restrict__ pc,
                             const uint32_t* __restrict__ jn, const uint8_t* __restrict__ is_outlet,
                             const uint32_t* __restrict__ head_base, uint32_t* __restrict__ perm) {
  uint32_t v = gid();
  if (v >= n) return;
  if (!is_outlet[jn[v]]) return;
  perm[head_base[jc[v]] + pc[v]] = v;
}

// --------------------------------------------------------------- sweeps ---

// Drainage area for the chains of one light depth, one lane per chain,
// tail to head.  Light children were finished in the previous (deeper)
// round; the heavy child's value is the running register.
__global__ void k_da_round(uint32_t c0, uint32_t c1, const uint32_t* __restrict__ chain_base,
                           const uint32_t* __restrict__ chain_len, const uint32_t* __restrict__ perm,
                           const uint32_t* __restrict__ coff, const uint32_t* __restrict__ ccnt,
                           const uint32_t* __restrict__ clist, const uint32_t* __restrict__ hc,
                           const double* __restrict__ area, double* __restrict__ A) {
  uint32_t c = c0 + gid();
  if (c >= c1) return;
  uint32_t base = chain_base[c], len = chain_len[c];
  double running = 0.0;
  for (uint32_t t = len; t-- > 0;) {
    uint32_t k = perm[base + t];
    double acc = area[k];
    uint32_t o = coff[k], m = ccnt[k], hk = hc[k];
    for (uint32_t q = m; q-- > 0;) {
      uint32_t ch = clist[o + q];
      double val = (ch == hk) ? running : A[ch];
      acc = acc + val;
    }
    A[k] = acc;
    running = acc;
  }
}

// x_k = 1.0 / (K_k * A_k^0.5) * d_k for every site in an outlet's basin
// (generator.rs:172-173), evaluated in the reference's association.  Sites
// outside every basin keep drainage area = own area and response time 0
// (generator.rs:144-145), as the reference leaves them.
__global__ void k_celerity(uint32_t n, const uint32_t* __restrict__ jn, const uint8_t* __restrict__ is_outlet,
                           const double* __restrict__ K, const double* __restrict__ area, double* __restrict__ A,
                           const double* __restrict__ dnext, double* __restrict__ x, double* __restrict__ tau) {
  uint32_t v = gid();
  if (v >= n) return;
  if (!is_outlet[jn[v]]) {
    A[v] = area[v];
    tau[v] = 0.0;
    return;
  }
  double celerity = K[v] * rust_pow_half(A[v]);
  x[v] = 1.0 / celerity * dnext[v];
}

// Response time and elevation for the chains of one light depth, head to
// tail (generator.rs:162-203).  A head's receiver lies on a chain of a
// smaller light depth, finished in an earlier round; the basin's outlet
// (jn = root) was finished in round 0.
__global__ void k_tau_round(uint32_t c0, uint32_t c1, const uint32_t* __restrict__ chain_base,
                            const uint32_t* __restrict__ chain_len, const uint32_t* __restrict__ perm,
                            const uint32_t* __restrict__ next, const uint32_t* __restrict__ jn,
                            const double* __restrict__ x, const double* __restrict__ U,
                            const double* __restrict__ tan_ms, const double* __restrict__ dnext,
                            double* __restrict__ tau, double* __restrict__ h, int* __restrict__ flags) {
  uint32_t c = c0 + gid();
  bool changed = false;
  if (c < c1) {
    uint32_t base = chain_base[c], len = chain_len[c];
    uint32_t k = perm[base];
    uint32_t p = next[k];
    uint32_t o = jn[k];
    double tau_p, h_p, tau_o, h_o;
    if (p == k) {
      // the outlet: rt[o] += rt[o] + x_o with rt[o] = 0; h_o' = h_o + U_o*max(0,0)
      double t0 = 0.0 + (0.0 + x[k]);
      double ho = h[k];
      double nh = ho + U[k] * rust_max(t0 - t0, 0.0);
      double tm = tan_ms[k];
      double slope = (nh - ho) / dnext[k];  // receiver is itself, not yet updated
      if (slope > tm) nh = ho + tm * dnext[k];
      changed |= (nh != ho);
      tau[k] = t0;
      h[k] = nh;
      tau_p = t0;
      h_p = nh;
      tau_o = t0;
      h_o = nh;
    } else {
      tau_p = tau[p];
      h_p = h[p];
      tau_o = tau[o];
      h_o = h[o];
      double t = 0.0 + (tau_p + x[k]);
      double nh = h_o + U[k] * rust_max(t - tau_o, 0.0);
      double tm = tan_ms[k];
      double d = dnext[k];
      double slope = (nh - h_p) / d;
      if (slope > tm) nh = h_p + tm * d;
      double old = h[k];
      changed |= (nh != old);
      tau[k] = t;
      h[k] = nh;
      tau_p = t;
      h_p = nh;
    }
    for (uint32_t t = 1; t < len; ++t) {
      k = perm[base + t];
      double tk = 0.0 + (tau_p + x[k]);
      double nh = h_o + U[k] * rust_max(tk - tau_o, 0.0);
      double tm = tan_ms[k];
      double d = dnext[k];
      double slope = (nh - h_p) / d;
      if (slope > tm) nh = h_p + tm * d;
      double old = h[k];
      changed |= (nh != old);
      tau[k] = tk;
      h[k] = nh;
      tau_p = tk;
      h_p = nh;
    }
  }
  if (__any(changed) && (threadIdx.x & 63u) == 0) flags[FL_FLAG_CHANGED] = 1;
}

// --------------------------------------------------------- launchers -----

namespace {
inline dim3 grid_for(uint64_t n, uint32_t block) { return dim3((uint32_t)((n + block - 1) / block)); }
}  // namespace

thread_local fl_profiler* fl_current_profiler = nullptr;

hipEvent_t fl_profiler::get() {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

void fl_profiler::record(hipStream_t s, const char* name, hipEvent_t a, hipEvent_t b) {
  pending.push_back({name, a, b});
  if (pending.size() >= 8192) flush();
  (void)s;
}

void fl_profiler::flush() {
  for (const Rec& r : pending) {
    float ms = 0.f;
    (void)hipEventSynchronize(r.b);
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    auto& t = totals[r.name];
    t.first += 1;
    t.second += ms;
    pool.push_back(r.a);
    pool.push_back(r.b);
  }
  pending.clear();
}

void fl_profiler::reset() {
  flush();
  totals.clear();
}

fl_profiler::~fl_profiler() {
  flush();
  for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

#define FL_LAUNCH_GRID(kernel, grid, block, stream, ...)                           \
  do {                                                                            \
    fl_profiler* prof_ = fl_current_profiler;                                     \
    if (prof_) {                                                                  \
      hipEvent_t a_ = prof_->get(), b_ = prof_->get();                            \
      (void)hipEventRecord(a_, (stream));                                         \
      hipLaunchKernelGGL(kernel, (grid), (block), 0, (stream), __VA_ARGS__);     \
      (void)hipEventRecord(b_, (stream));                                         \
      prof_->record((stream), #kernel, a_, b_);                                   \
    } else {                                                                      \
      hipLaunchKernelGGL(kernel, (grid), (block), 0, (stream), __VA_ARGS__);     \
    }                                                                             \
  } while (0)

#define FL_LAUNCH(kernel, count, stream, ...)                                              \
  do {                                                                                     \
    if ((count) > 0) FL_LAUNCH_GRID(kernel, grid_for((count), 256), dim3(256), (stream), __VA_ARGS__); \
  } while (0)

void fl_launch_receivers(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w) {
  FL_LAUNCH(k_receivers, t.n, s, t.n, t.row_ptr, t.col, t.len, w.h, t.is_outlet, w.next, w.dnext, w.flags);
}

void fl_launch_sub_init(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w) {
  FL_LAUNCH(k_copy_u32, t.n, s, t.n, w.next, w.sub);
}

void fl_launch_sub_jump(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, bool flip) {
  const uint32_t* in = flip ? w.sub2 : w.sub;
  uint32_t* out = flip ? w.sub : w.sub2;
  FL_LAUNCH(k_sub_jump, t.n, s, t.n, in, out, w.flags + FL_FLAG_JUMP);
}

void fl_launch_lake_join(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w) {
  FL_LAUNCH(k_fill_u64, t.n, s, t.n, w.lake_key, ~0ull);
  FL_LAUNCH(k_lake_keys, t.n, s, t.n, t.row_ptr, t.col, t.rank, w.sub, t.is_outlet, (unsigned long long*)w.lake_key);
  FL_LAUNCH(k_lake_flip, t.n, s, t.n, t.row_ptr, t.col, t.len, t.order, (const unsigned long long*)w.lake_key,
            t.is_outlet, w.next, w.dnext);
}

void fl_launch_children(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w) {
  FL_LAUNCH(k_children, t.n, s, t.n, t.row_ptr, t.col, w.next, w.A, w.coff, w.ccnt, w.clist, w.hc,
            w.counters + FL_CNT_CHILDREN);
  FL_LAUNCH(k_chain_init, t.n, s, t.n, w.next, w.hc, w.jc, w.pc, w.jn, w.lw);
}

void fl_launch_chain_jump(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, bool flip) {
  if (!flip)
    FL_LAUNCH(k_chain_jump, t.n, s, t.n, w.jc, w.pc, w.jn, w.lw, w.jc2, w.pc2, w.jn2, w.lw2, w.flags + FL_FLAG_JUMP);
  else
    FL_LAUNCH(k_chain_jump, t.n, s, t.n, w.jc2, w.pc2, w.jn2, w.lw2, w.jc, w.pc, w.jn, w.lw, w.flags + FL_FLAG_JUMP);
}

void fl_launch_chain_count(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w) {
  FL_LAUNCH(k_chain_len, t.n, s, t.n, w.hc, w.jc, w.pc, w.jn, t.is_outlet, w.clen);
  FL_LAUNCH(k_bucket_count, t.n, s, t.n, w.jc, w.jn, w.lw, t.is_outlet, w.clen, w.bcnt, w.bnodes,
            w.counters + FL_CNT_MAXLD);
}

void fl_launch_chain_layout(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, uint32_t nb) {
  FL_LAUNCH_GRID(k_bucket_scan, dim3(1), dim3(1024), s, nb, w.bcnt, w.bnodes, w.boff, w.bnoff);
  FL_LAUNCH(k_bucket_place, t.n, s, t.n, w.jc, w.jn, w.lw, t.is_outlet, w.clen, w.boff, w.bnoff, w.bcur, w.bncur,
            w.chain_base, w.chain_len, w.head_base);
  FL_LAUNCH(k_chain_perm, t.n, s, t.n, w.jc, w.pc, w.jn, t.is_outlet, w.head_base, w.perm);
}

void fl_launch_da_round(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, uint32_t c0, uint32_t c1) {
  FL_LAUNCH(k_da_round, c1 - c0, s, c0, c1, w.chain_base, w.chain_len, w.perm, w.coff, w.ccnt, w.clist, w.hc,
            t.area, w.A);
}

void fl_launch_celerity(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w) {
  FL_LAUNCH(k_celerity, t.n, s, t.n, w.jn, t.is_outlet, t.erod, t.area, w.A, w.dnext, w.x, w.tau);
}

void fl_launch_tau_round(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, uint32_t c0, uint32_t c1) {
  FL_LAUNCH(k_tau_round, c1 - c0, s, c0, c1, w.chain_base, w.chain_len, w.perm, w.next, w.jn, w.x, t.uplift,
            t.tan_ms, w.dnext, w.tau, w.h, w.flags);
}

void fl_launch_fill_u32(hipStream_t s, uint32_t* p, uint64_t count, uint32_t v) {
  FL_LAUNCH(k_fill_u32, count, s, (uint32_t)count, p, v);
}

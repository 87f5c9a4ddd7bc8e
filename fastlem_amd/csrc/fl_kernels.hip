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
//     k_chain_jump      chain head, position in chain, light depth, root
//     k_bucket_*        chains bucketed by light depth, laid out contiguously
//     k_chain_perm      sites scattered to chain-major positions, with
//                                             each site's light children
//   sweeps, one round per light depth (deepest first for drainage area)
//     k_da_fold         generator.rs:154-159  drainage area along the chains,
//                       generator.rs:172      exact fold order; then each
//                                             site's 1/(K*A^0.5) * distance
//     k_tau_fold        generator.rs:162-203  response time (+ clamped
//                                             elevation), head to tail
//     k_da_tail, k_tau_tail                   the deep rounds (few chains
//                                             each) in one launch per sweep
//     k_tau_scatter     generator.rs:177-202  elevation, changed flag
//
// Per-site kernels take their site ranges XCD by XCD (xcd_block): each of
// the 8 XCDs works on one contiguous eighth of the sites.
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
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fl_kernels.h"

#define FL_NONE 0xffffffffu
#define FL_LIST 4                   // light children listed per chain position
#define FL_LIST_OVERFLOW 0x10000u   // pcnt flag: more light children than listed
#define FL_ROW 8                    // graph row entries loaded together

namespace {

// Workgroups are dealt to the GPU's 8 XCDs in turn (workgroup b to XCD
// b % 8), each XCD with an L2 of its own.  The per-site kernels hand each XCD
// one contiguous eighth of the index range instead, so that sites close in
// the breadth-first labels, whose neighbour gathers touch the same lines,
// share an L2.  A bijection of [0, gridDim.x) for any grid size.
__device__ __forceinline__ uint32_t xcd_block() {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// bs: the workgroup size the kernel was launched with, passed as an argument
// (blockDim.x is read from the dispatch's implicit arguments by a vector load
// that every wavefront would wait for before its first site load)
__device__ __forceinline__ uint32_t gid(uint32_t bs) { return xcd_block() * bs + threadIdx.x; }

// (pointer, count) pairs of the pointer jumping: pointer in the low word,
// count in the high word below FL_FIN, the mark of a pair whose pointer is
// known to be a fixed point (a chain head, a root).  Counts stay below 2^31
// (they count sites).
#define FL_FIN 0x80000000u
__device__ __forceinline__ unsigned long long pair(uint32_t ptr, uint32_t cnt) {
  return ((unsigned long long)cnt << 32) | ptr;
}
__device__ __forceinline__ uint32_t pair_ptr(unsigned long long p) { return (uint32_t)p; }
__device__ __forceinline__ uint32_t pair_cnt(unsigned long long p) { return (uint32_t)(p >> 32) & ~FL_FIN; }
__device__ __forceinline__ bool pair_fin(unsigned long long p) { return (p >> 63) != 0; }

// Is slot s of row k (starting at b) a child of k?
__device__ __forceinline__ bool is_child(unsigned long long m, uint32_t deg, uint32_t s, uint32_t b,
                                         const uint32_t* __restrict__ col, const uint32_t* __restrict__ next,
                                         uint32_t k) {
  return deg <= 64 ? ((m >> s) & 1ull) != 0 : next[col[b + s]] == k;
}

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
// A receiver depends on the site's and its neighbours' elevations only: with
// chg_prev (per-site flags, raised where the previous iteration, whose tree
// had no lake joins, changed the elevation of the site or of a neighbour), a
// site whose neighbourhood did not change keeps its previous receiver and
// distance.  chg_out (this iteration's flags) is zeroed here for
// k_tau_scatter to mark.
//
// With `dirty` (cleared by the previous iteration's k_tau_scatter), a site
// whose receiver or receiver distance differs from the previous tree's
// marks the previous basins of itself, its new and its old receiver as
// recomputed (the marks of k_dirty, made here without a launch of their
// own).
__global__ void k_receivers(uint32_t bs, uint32_t n, const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                            const double* __restrict__ len, const double* __restrict__ h,
                            const uint8_t* __restrict__ is_outlet, uint32_t* __restrict__ next,
                            double* __restrict__ dnext, int* __restrict__ flags,
                            const uint8_t* __restrict__ chg_prev, const uint32_t* __restrict__ next_prev,
                            const double* __restrict__ dnext_prev, const uint32_t* __restrict__ root_prev,
                            uint8_t* __restrict__ dirty, uint8_t* __restrict__ chg_out) {
  uint32_t i = gid(bs);
  if (i >= n) return;
  chg_out[i] = 0;  // this iteration's k_tau_scatter marks the changed neighbourhoods
  // the site's own words are loaded together, before the outlet test
  const bool prev = chg_prev || dirty;
  const uint8_t outlet = is_outlet[i];
  const uint32_t b = rp[i], e = rp[i + 1];
  const uint32_t pj = prev ? next_prev[i] : 0u;
  const double pd = prev ? dnext_prev[i] : 0.0;
  if (outlet) {
    next[i] = i;
    dnext[i] = 1.0;
    return;
  }
  if (chg_prev && chg_prev[i] == 0) {
    if (pj == i) flags[FL_FLAG_LAKE] = 1;
    next[i] = pj;
    dnext[i] = pd;
    return;
  }
  // a row is read FL_ROW entries at a time, every load of a group issued
  // before the first is used (indices past the row's end repeat its last
  // entry and are not evaluated)
  uint32_t c[FL_ROW];
  const double hi = h[i];
  double best = 0.0, d = 1.0;
  uint32_t bj = i;
  for (uint32_t k = b; k < e; k += FL_ROW) {
#pragma unroll
    for (int j = 0; j < FL_ROW; ++j) c[j] = col[min(k + j, e - 1)];
    double hj[FL_ROW], lk[FL_ROW];
#pragma unroll
    for (int j = 0; j < FL_ROW; ++j) {
      hj[j] = h[c[j]];
      lk[j] = len[min(k + j, e - 1)];
    }
    // slots in row order: the earliest of equal slopes is kept
#pragma unroll
    for (int j = 0; j < FL_ROW; ++j) {
      if (k + j < e && hi > hj[j]) {
        double s = (hi - hj[j]) / lk[j];
        if (s > best) {
          best = s;
          bj = c[j];
          d = lk[j];  // rows hold no duplicate edges: this slot is has_edge(i, bj)'s
        }
      }
    }
  }
  if (bj == i) flags[FL_FLAG_LAKE] = 1;
  next[i] = bj;
  dnext[i] = d;
  if (dirty) {
    if (bj != pj || __double_as_longlong(d) != __double_as_longlong(pd)) {
      dirty[root_prev[i]] = 1;
      dirty[root_prev[bj]] = 1;
      dirty[root_prev[pj]] = 1;
    }
  }
}

// ------------------------------------------------------------ lake join ---

__global__ void k_copy_u32(uint32_t bs, uint32_t n, const uint32_t* __restrict__ src, uint32_t* __restrict__ dst) {
  uint32_t i = gid(bs);
  if (i < n) dst[i] = src[i];
}

__global__ void k_fill_u64(uint32_t bs, uint32_t n, uint64_t* __restrict__ dst, uint64_t v) {
  uint32_t i = gid(bs);
  if (i < n) dst[i] = v;
}

__global__ void k_fill_u32(uint32_t bs, uint32_t n, uint32_t* __restrict__ dst, uint32_t v) {
  uint32_t i = gid(bs);
  if (i < n) dst[i] = v;
}

// One pointer-jumping round: out[i] = in[in[i]]; raises *flag if anything moved.
__global__ void k_sub_jump(uint32_t bs, uint32_t n, const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                           int* __restrict__ flag) {
  uint32_t i = gid(bs);
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
__global__ void k_lake_keys(uint32_t bs, uint32_t n, const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                            const uint32_t* __restrict__ rank, const uint32_t* __restrict__ sub,
                            const uint8_t* __restrict__ is_outlet, unsigned long long* __restrict__ key) {
  uint32_t i = gid(bs);
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
__global__ void k_lake_flip(uint32_t bs, uint32_t n, const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                            const double* __restrict__ len, const uint32_t* __restrict__ order,
                            const unsigned long long* __restrict__ key, const uint8_t* __restrict__ is_outlet,
                            uint32_t* __restrict__ next, double* __restrict__ dnext) {
  uint32_t s = gid(bs);
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
// cmask[u] bit s marks slot s as a child (rows of up to 64 entries; longer
// rows are rescanned by the consumers).  The heavy child (first with the
// largest previous drainage area) continues u's chain; the choice only
// shapes the work, never the results.
// Basins of the previous iteration's tree that are recomputed: those holding
// a site whose receiver or receiver distance changed, or that site's old or
// new receiver (k_receivers marks those of raw receivers; this kernel runs
// after a lake join changed receivers).  Any other basin keeps every parent pointer and distance of
// its tree, and no site of another basin attaches to it, so its drainage
// areas, response times and elevations come out bit for bit as before.
__global__ void k_dirty(uint32_t bs, uint32_t n, const uint32_t* __restrict__ next, const double* __restrict__ dnext,
                        const uint32_t* __restrict__ next_prev, const double* __restrict__ dnext_prev,
                        const uint32_t* __restrict__ root_prev, uint8_t* __restrict__ dirty) {
  uint32_t v = gid(bs);
  if (v >= n) return;
  const uint32_t a = next[v], b = next_prev[v];
  if (a != b || __double_as_longlong(dnext[v]) != __double_as_longlong(dnext_prev[v])) {
    dirty[root_prev[v]] = 1;
    dirty[root_prev[a]] = 1;
    dirty[root_prev[b]] = 1;
  }
}

// Child slots and heavy child of every active site; act[u] (read by the
// iteration's later kernels): u belongs to a dirty basin (every site when
// `dirty` is null) of a terrain still iterating (tid, t_active: per-site
// terrain and per-terrain mask of a union, or null).  Also the chain and
// root pointers for pointer jumping, written by each site for its children
// (every child of an active site is active, and every active site is a root
// or a child of one):
//   cp: (chain parent, 1), (v, 0) at a chain head
//   rt: (receiver, 1 at a head that is not a root: a light edge), (v, 0) at a root
// An inactive site keeps its basin: its root is carried to this iteration's.
__global__ void k_children(uint32_t bs, uint32_t n, const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                           const uint32_t* __restrict__ next, const double* __restrict__ prev_area,
                           unsigned long long* __restrict__ cmask, uint32_t* __restrict__ hc,
                           uint8_t* __restrict__ act, const uint8_t* __restrict__ dirty,
                           const uint32_t* __restrict__ root_prev, const uint32_t* __restrict__ tid,
                           const uint8_t* __restrict__ t_active, unsigned long long* __restrict__ cp,
                           unsigned long long* __restrict__ rt, uint32_t* __restrict__ root_out) {
  uint32_t u = gid(bs);
  if (u >= n) return;
  // the site's own words together (an inactive site reads its receiver and
  // row for nothing: one load latency less for the active ones)
  const uint32_t ru = root_prev[u], nu = next[u], b = rp[u], e = rp[u + 1];
  const uint32_t tu = tid ? tid[u] : 0u;
  bool a = !dirty || dirty[ru];
  if (tid) a = a && t_active[tu];
  act[u] = a ? 1 : 0;
  if (!a) {
    root_out[u] = ru;
    return;
  }
  if (nu == u) {
    cp[u] = pair(u, FL_FIN);
    rt[u] = pair(u, FL_FIN);
  }
  uint32_t heavy = FL_NONE;
  unsigned long long m = 0;
  double best = 0.0;
  for (uint32_t k = b; k < e; k += FL_ROW) {
    // FL_ROW neighbours at a time, every load issued before the first is
    // used (indices past the row's end repeat its last entry, not evaluated)
    uint32_t c[FL_ROW], nx[FL_ROW];
    double av[FL_ROW];
#pragma unroll
    for (int j = 0; j < FL_ROW; ++j) c[j] = col[min(k + j, e - 1)];
#pragma unroll
    for (int j = 0; j < FL_ROW; ++j) {
      nx[j] = next[c[j]];
      av[j] = prev_area[c[j]];
    }
#pragma unroll
    for (int j = 0; j < FL_ROW; ++j) {
      if (k + j < e && nx[j] == u) {
        const uint32_t v = c[j];
        if (heavy == FL_NONE || av[j] > best) {
          best = av[j];
          heavy = v;
        }
        if (k + j - b < 64) m |= 1ull << (k + j - b);
        // a light child heads its chain (the heavy one is set below)
        cp[v] = pair(v, FL_FIN);
        rt[v] = pair(u, 1u);
      }
    }
  }
  cmask[u] = m;
  hc[u] = heavy;
  if (heavy != FL_NONE) {
    cp[heavy] = pair(u, 1u);
    rt[heavy] = pair(u, 0u);
  }
}

// One jumping round on both pointer systems, in place:
// (a, d) <- (a''', d + d' + d'' + d''') with (a', d') the pair of a,
// (a'', d'') that of a' and (a''', d''') that of a'' (three hops: the reach
// quadruples per round).  A pair read while its
// owner rewrites it is either its old or its new value (aligned 64-bit
// accesses are single-copy atomic), and both are (ancestor, count) pairs of
// the same path, so rounds may mix them.  After convergence
// cp[v] = (chain head, position in chain), rt[v] = (root of v's tree, light
// edges between v and the root = light depth of v's chain).
//
// Fixed points (a head's cp, a root's rt: pairs (x, 0) of x itself) carry
// FL_FIN, and so does a pair that reaches one: a site finding its pointer at
// a fixed point marks its pair, and a hop inherits the mark of the pair it
// copies (the pair of a'' is marked when a''' is a fixed point).  A marked
// pair never changes and its site loads nothing more.  A round flags, in its
// own control word (zeroed at the start of the iteration), whether some
// pair is still unmarked after it: a round without the flag proves
// convergence in the round that completes it, and the rounds launched after
// it return at once.
__global__ void k_chain_jump(uint32_t bs, uint32_t n, unsigned long long* cp, unsigned long long* rt,
                             uint32_t* __restrict__ flag_set, const uint32_t* __restrict__ prev_flag,
                             const uint8_t* __restrict__ act) {
  uint32_t v = gid(bs);
  // the site's own pairs are read together with the previous round's flag
  // (an inactive site's pairs are stale and unused: no active site points
  // at one, as a dirty basin holds its sites' whole paths)
  const bool on = v < n && act[v] != 0;
  unsigned long long c = 0, r = 0;
  if (on) {
    c = __atomic_load_n(cp + v, __ATOMIC_RELAXED);
    r = __atomic_load_n(rt + v, __ATOMIC_RELAXED);
  }
  // a round after a round that left every pair marked has nothing to do
  if (prev_flag && *prev_flag == 0) return;
  bool open = false;
  if (on) {
    const bool oc = !pair_fin(c), orr = !pair_fin(r);
    const uint32_t a = pair_ptr(c), b = pair_ptr(r);
    const unsigned long long ca = oc ? __atomic_load_n(cp + a, __ATOMIC_RELAXED) : 0ull;
    const unsigned long long rb = orr ? __atomic_load_n(rt + b, __ATOMIC_RELAXED) : 0ull;
    // at a fixed point already (a is a head, b a root): mark; else hop
    const bool fc = oc && pair_ptr(ca) == a, fr = orr && pair_ptr(rb) == b;
    const bool mc = oc && !fc, mr = orr && !fr;
    const unsigned long long ca2 = mc ? __atomic_load_n(cp + pair_ptr(ca), __ATOMIC_RELAXED) : 0ull;
    const unsigned long long rb2 = mr ? __atomic_load_n(rt + pair_ptr(rb), __ATOMIC_RELAXED) : 0ull;
    // a third hop (a fixed point's pair is (x, 0) of x itself: hops past one
    // stay there and add nothing)
    const unsigned long long ca3 = mc ? __atomic_load_n(cp + pair_ptr(ca2), __ATOMIC_RELAXED) : 0ull;
    const unsigned long long rb3 = mr ? __atomic_load_n(rt + pair_ptr(rb2), __ATOMIC_RELAXED) : 0ull;
    if (fc) __atomic_store_n(cp + v, c | (1ull << 63), __ATOMIC_RELAXED);
    if (fr) __atomic_store_n(rt + v, r | (1ull << 63), __ATOMIC_RELAXED);
    if (mc) {
      const bool f = pair_fin(ca3);
      __atomic_store_n(
          cp + v,
          pair(pair_ptr(ca3), (pair_cnt(c) + pair_cnt(ca) + pair_cnt(ca2) + pair_cnt(ca3)) | (f ? FL_FIN : 0u)),
          __ATOMIC_RELAXED);
      open = open || !f;
    }
    if (mr) {
      const bool f = pair_fin(rb3);
      __atomic_store_n(
          rt + v,
          pair(pair_ptr(rb3), (pair_cnt(r) + pair_cnt(rb) + pair_cnt(rb2) + pair_cnt(rb3)) | (f ? FL_FIN : 0u)),
          __ATOMIC_RELAXED);
      open = open || !f;
    }
  }
  if (__any(open) && (threadIdx.x & 63u) == 0) *flag_set = 1;
}

// Buckets: 2 * light depth + (chain length >= long_min), so every round
// has a range of short chains (one lane each) followed by long ones (one
// wavefront each); long_min is a launch parameter (tuned; any value >= 1
// gives the same results).
#define FL_LDS_BUCKETS 128
#define FL_BUCKET_BLOCK 1024
#define FL_BUCKET_PER_THREAD 2

// The chain tails among a thread's FL_BUCKET_PER_THREAD sites of the
// workgroup's range starting at `base`, with their pairs: activity and heavy
// child of every site loaded first (clamped in range, all together), the
// pairs of the tails second, the outlet flag of their roots third.
struct fl_tail {
  unsigned long long r, c;
  bool t;
};
__device__ __forceinline__ void load_tails(uint32_t base, uint32_t n, const uint32_t* __restrict__ hc,
                                           const unsigned long long* __restrict__ cp,
                                           const unsigned long long* __restrict__ rt,
                                           const uint8_t* __restrict__ is_outlet, const uint8_t* __restrict__ act,
                                           fl_tail (&q)[FL_BUCKET_PER_THREAD]) {
  uint8_t a[FL_BUCKET_PER_THREAD];
  uint32_t h[FL_BUCKET_PER_THREAD];
#pragma unroll
  for (uint32_t j = 0; j < FL_BUCKET_PER_THREAD; ++j) {
    const uint32_t v = base + j * FL_BUCKET_BLOCK + threadIdx.x, vv = v < n ? v : n - 1;
    a[j] = act[vv];
    h[j] = hc[vv];
  }
#pragma unroll
  for (uint32_t j = 0; j < FL_BUCKET_PER_THREAD; ++j) {
    const uint32_t v = base + j * FL_BUCKET_BLOCK + threadIdx.x;
    q[j].t = v < n && a[j] && h[j] == FL_NONE;
    q[j].r = q[j].t ? rt[v] : 0ull;
    q[j].c = q[j].t ? cp[v] : 0ull;
  }
#pragma unroll
  for (uint32_t j = 0; j < FL_BUCKET_PER_THREAD; ++j)
    if (q[j].t) q[j].t = is_outlet[pair_ptr(q[j].r)] != 0;
}

// Light-depth histogram of the chains (count and total length), aggregated
// per workgroup in LDS before one global atomic per (workgroup, bucket).
// Each chain is counted by its tail (the member without a heavy child),
// which knows the chain's length (its position + 1) and light depth.
__global__ void __launch_bounds__(FL_BUCKET_BLOCK)
k_bucket_count(const uint32_t* __restrict__ guard, uint32_t n, const uint32_t* __restrict__ hc, const unsigned long long* __restrict__ cp,
               const unsigned long long* __restrict__ rt, const uint8_t* __restrict__ is_outlet,
               uint32_t* __restrict__ bcnt, uint32_t* __restrict__ bnodes, uint32_t* __restrict__ maxld,
               uint32_t long_min, const uint8_t* __restrict__ act) {
  const uint32_t base = xcd_block() * (FL_BUCKET_BLOCK * FL_BUCKET_PER_THREAD);
  fl_tail q[FL_BUCKET_PER_THREAD];
  load_tails(base, n, hc, cp, rt, is_outlet, act, q);  // in flight while the guard is read
  if (*guard) return;  // pointer jumping has not converged: the iteration is redone
  __shared__ uint32_t s_c[FL_LDS_BUCKETS], s_n[FL_LDS_BUCKETS];
  __shared__ uint32_t s_max;
  for (uint32_t t = threadIdx.x; t < FL_LDS_BUCKETS; t += FL_BUCKET_BLOCK) {
    s_c[t] = 0;
    s_n[t] = 0;
  }
  if (threadIdx.x == 0) s_max = 0;
  __syncthreads();
  for (uint32_t j = 0; j < FL_BUCKET_PER_THREAD; ++j) {
    if (q[j].t) {
      uint32_t lw = pair_cnt(q[j].r), len = pair_cnt(q[j].c) + 1, l = 2u * lw + (len >= long_min ? 1u : 0u);
      if (l < FL_LDS_BUCKETS) {
        atomicAdd(s_c + l, 1u);
        atomicAdd(s_n + l, len);
      } else {
        atomicAdd(bcnt + l, 1u);
        atomicAdd(bnodes + l, len);
      }
      atomicMax(&s_max, lw);
    }
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < FL_LDS_BUCKETS; t += FL_BUCKET_BLOCK) {
    if (s_c[t]) {
      atomicAdd(bcnt + t, s_c[t]);
      atomicAdd(bnodes + t, s_n[t]);
    }
  }
  if (threadIdx.x == 0) atomicMax(maxld, s_max);
}

// Exclusive scans of the bucket histograms over the light depths actually
// present (2 * (counters[FL_CNT_MAXLD] + 1) buckets), one workgroup.  The first
// FL_CTRL_HEAD + 1 offsets are mirrored into the control block; offsets up to
// `fill` are written.
#define FL_SCAN_BLOCK 1024  // k_bucket_scan: one workgroup of this size
__global__ void __launch_bounds__(FL_SCAN_BLOCK)
k_bucket_scan(uint32_t n, uint32_t fill, uint32_t* __restrict__ ctrl,
                              const uint32_t* __restrict__ bcnt, const uint32_t* __restrict__ bnodes,
                              uint32_t* __restrict__ boff, uint32_t* __restrict__ bnoff) {
  __shared__ uint32_t s_a[FL_SCAN_BLOCK], s_b[FL_SCAN_BLOCK];
  __shared__ uint32_t carry_a, carry_b;
  uint32_t* counters = ctrl + FL_NUM_FLAGS;
  uint32_t nb = 2 * (counters[FL_CNT_MAXLD] + 1);
  if (nb > 2 * n + 2) nb = 2 * n + 2;
  uint32_t* head_c = counters + FL_NUM_COUNTERS;
  uint32_t* head_n = head_c + FL_CTRL_HEAD + 1;
  if (threadIdx.x == 0) {
    carry_a = 0;
    carry_b = 0;
  }
  __syncthreads();
  for (uint32_t base = 0; base < nb; base += FL_SCAN_BLOCK) {
    uint32_t i = base + threadIdx.x;
    uint32_t a = i < nb ? bcnt[i] : 0, b = i < nb ? bnodes[i] : 0;
    s_a[threadIdx.x] = a;
    s_b[threadIdx.x] = b;
    __syncthreads();
    for (uint32_t off = 1; off < FL_SCAN_BLOCK; off <<= 1) {
      uint32_t ta = threadIdx.x >= off ? s_a[threadIdx.x - off] : 0;
      uint32_t tb = threadIdx.x >= off ? s_b[threadIdx.x - off] : 0;
      __syncthreads();
      s_a[threadIdx.x] += ta;
      s_b[threadIdx.x] += tb;
      __syncthreads();
    }
    if (i < nb) {
      uint32_t oa = carry_a + s_a[threadIdx.x] - a, ob = carry_b + s_b[threadIdx.x] - b;
      boff[i] = oa;
      bnoff[i] = ob;
      if (i <= FL_CTRL_HEAD) {
        head_c[i] = oa;
        head_n[i] = ob;
      }
    }
    __syncthreads();
    if (threadIdx.x == FL_SCAN_BLOCK - 1) {
      carry_a += s_a[threadIdx.x];
      carry_b += s_b[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    boff[nb] = carry_a;
    bnoff[nb] = carry_b;
    if (nb <= FL_CTRL_HEAD) {
      head_c[nb] = carry_a;
      head_n[nb] = carry_b;
    }
    counters[FL_CNT_CHAINS] = carry_a;
    counters[FL_CNT_POSITIONS] = carry_b;
  }
  // offsets nb+1..fill repeat the totals: rounds launched beyond the deepest
  // light depth find empty ranges
  for (uint32_t i = nb + 1 + threadIdx.x; i <= fill; i += FL_SCAN_BLOCK) {
    boff[i] = carry_a;
    bnoff[i] = carry_b;
  }
}

// Give every chain (through its tail) a slot in its light-depth bucket and a
// contiguous range of positions: LDS atomics rank the chains of a workgroup,
// one global atomic per (workgroup, bucket) reserves the ranges.  Order
// inside a bucket is immaterial to the results.  Chain q's record starts
// (base, len) (k_chain_perm adds the receiver's and the outlet's positions),
// its head knows q and its base, and the base's receiver slot starts empty.
__device__ __forceinline__ void place_chain(uint32_t q, uint32_t pos, uint32_t len, uint32_t hd,
                                            uint4* __restrict__ cmeta, uint32_t* __restrict__ head_chain,
                                            uint32_t* __restrict__ head_base, uint32_t* __restrict__ pdest) {
  reinterpret_cast<uint2*>(cmeta)[2ull * q] = make_uint2(pos, len);
  head_chain[hd] = q;
  head_base[hd] = pos;
  pdest[pos] = FL_NONE;
}

__global__ void __launch_bounds__(FL_BUCKET_BLOCK)
k_bucket_place(const uint32_t* __restrict__ guard, uint32_t n, const uint32_t* __restrict__ hc, const unsigned long long* __restrict__ cp,
               const unsigned long long* __restrict__ rt, const uint8_t* __restrict__ is_outlet,
               const uint32_t* __restrict__ boff,
               const uint32_t* __restrict__ bnoff, uint32_t* __restrict__ bcur, uint32_t* __restrict__ bncur,
               uint4* __restrict__ cmeta, uint32_t* __restrict__ head_chain, uint32_t* __restrict__ pdest,
               uint32_t* __restrict__ head_base, uint32_t long_min, const uint8_t* __restrict__ act) {
  const uint32_t base = xcd_block() * (FL_BUCKET_BLOCK * FL_BUCKET_PER_THREAD);
  fl_tail tl[FL_BUCKET_PER_THREAD];
  load_tails(base, n, hc, cp, rt, is_outlet, act, tl);  // in flight while the guard is read
  if (*guard) return;  // pointer jumping has not converged: the iteration is redone
  __shared__ uint32_t s_c[FL_LDS_BUCKETS], s_n[FL_LDS_BUCKETS], s_bc[FL_LDS_BUCKETS], s_bn[FL_LDS_BUCKETS];
  for (uint32_t t = threadIdx.x; t < FL_LDS_BUCKETS; t += FL_BUCKET_BLOCK) {
    s_c[t] = 0;
    s_n[t] = 0;
  }
  __syncthreads();
  uint32_t lc[FL_BUCKET_PER_THREAD], ln[FL_BUCKET_PER_THREAD];
  for (uint32_t j = 0; j < FL_BUCKET_PER_THREAD; ++j) {
    lc[j] = FL_NONE;
    ln[j] = 0;
    if (tl[j].t) {
      const unsigned long long c = tl[j].c;
      uint32_t len = pair_cnt(c) + 1, l = 2u * pair_cnt(tl[j].r) + (len >= long_min ? 1u : 0u);
      if (l < FL_LDS_BUCKETS) {
        lc[j] = atomicAdd(s_c + l, 1u);
        ln[j] = atomicAdd(s_n + l, len);
      } else {
        uint32_t q = boff[l] + atomicAdd(bcur + l, 1u);
        uint32_t pos = bnoff[l] + atomicAdd(bncur + l, len);
        place_chain(q, pos, len, pair_ptr(c), cmeta, head_chain, head_base, pdest);
      }
    }
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < FL_LDS_BUCKETS; t += FL_BUCKET_BLOCK) {
    if (s_c[t]) {
      s_bc[t] = atomicAdd(bcur + t, s_c[t]);
      s_bn[t] = atomicAdd(bncur + t, s_n[t]);
    }
  }
  __syncthreads();
  for (uint32_t j = 0; j < FL_BUCKET_PER_THREAD; ++j) {
    if (lc[j] == FL_NONE) continue;
    const unsigned long long c = tl[j].c;
    uint32_t len = pair_cnt(c) + 1;
    uint32_t l = 2u * pair_cnt(tl[j].r) + (len >= long_min ? 1u : 0u);
    uint32_t q = boff[l] + s_bc[l] + lc[j];
    uint32_t pos = bnoff[l] + s_bn[l] + ln[j];
    place_chain(q, pos, len, pair_ptr(c), cmeta, head_chain, head_base, pdest);
  }
}

// Scatter sites to their chain-major positions with the inputs the sweeps
// read there: the drainage-area fold's list counts, area, erodibility and
// receiver distance (with max slopes also uplift, tan(max slope), distance
// and old elevation), and for each listed light child the slot its chain's
// head fills with its drainage area (pdest at the child chain's base).  A
// chain's record gets the positions of the head's receiver (FL_NONE for an
// outlet) and of the basin's outlet.  Sites outside every basin keep their
// elevation, drainage area = own area and response time 0 (generator.rs:
// 144-145), as the reference leaves them.
struct fl_perm_out {
  uint32_t* perm;
  uint32_t* pcnt;
  uint32_t* pdest;
  double* parea;
  double* pK;
  double* pdn;
  double* up;
  double* tanp;
  double* dp;
  double* hop;
  uint4* cmeta;
};

__global__ void k_chain_perm(uint32_t bs, const uint32_t* __restrict__ guard, uint32_t n, const unsigned long long* __restrict__ cp,
                             const unsigned long long* __restrict__ rt, const uint32_t* __restrict__ next,
                             const uint8_t* __restrict__ is_outlet, const uint32_t* __restrict__ head_base,
                             const uint32_t* __restrict__ head_chain, fl_perm_out o,
                             const double* __restrict__ area, const double* __restrict__ K,
                             const double* __restrict__ dnext, const double* __restrict__ U,
                             const double* __restrict__ tan_ms, int clamp, double* __restrict__ A,
                             double* __restrict__ tau, const double* __restrict__ h, double* __restrict__ h2,
                             const uint32_t* __restrict__ maxld, uint32_t* __restrict__ bcnt,
                             uint32_t* __restrict__ bnodes, uint32_t* __restrict__ bcur,
                             uint32_t* __restrict__ bncur, const uint32_t* __restrict__ rp,
                             const uint32_t* __restrict__ col, const unsigned long long* __restrict__ cmask,
                             const uint32_t* __restrict__ hc, const uint8_t* __restrict__ act,
                             uint32_t* __restrict__ root) {
  uint32_t v = gid(bs);
  // the site's own words are read together with the guard (one load
  // latency; a site of a lake basin reads some it does not use)
  const uint32_t mld = *maxld;
  const bool on = v < n && act[v];
  unsigned long long rv = 0ull, cv = 0ull, m = 0ull;
  uint32_t hq = 0u, b = 0u, e = 0u, hk = 0u, p = 0u;
  double Kv = 0.0, dv = 0.0, av = 0.0;
  if (on) {
    rv = rt[v];
    cv = cp[v];
    hq = head_chain[v];  // meaningful at a chain head only
    b = rp[v];
    e = rp[v + 1];
    hk = hc[v];
    m = cmask[v];
    p = next[v];
    Kv = K[v];
    dv = dnext[v];
    av = area[v];
  }
  if (*guard) return;  // pointer jumping has not converged: the iteration is redone
  // the bucket histograms are used up (count, scan, place): clear the
  // entries this decomposition touched for the next one
  const uint32_t nb = 2u * (mld + 1u);
  for (uint32_t i = v; i <= nb && i < 2u * n + 4u; i += gridDim.x * bs) {
    bcnt[i] = 0;
    bnodes[i] = 0;
    bcur[i] = 0;
    bncur[i] = 0;
  }
  if (!on) return;
  const uint32_t r = pair_ptr(rv), hd = pair_ptr(cv), deg = e - b;
  root[v] = r;
  // the second level of loads, issued together: the root's outlet flag, the
  // chain's base, the row (up to FL_ROW entries) and, at a chain head, the
  // receiver's pair and the outlet's base
  const uint8_t outlet = is_outlet[r];
  const uint32_t base = head_base[hd];
  const bool head = hd == v;
  const unsigned long long cpp = head ? cp[p] : 0ull;
  const uint32_t hbr = head ? head_base[r] : 0u;
  uint32_t cc[FL_ROW];
  const bool short_row = deg <= FL_ROW;
  if (short_row && deg > 0) {
#pragma unroll
    for (int j = 0; j < FL_ROW; ++j) cc[j] = col[min(b + j, e - 1)];
  }
  if (!outlet) {
    A[v] = av;
    tau[v] = 0.0;
    h2[v] = h[v];
    return;
  }
  const uint32_t P = base + pair_cnt(cv);
  o.perm[P] = v;
  o.pK[P] = Kv;
  o.pdn[P] = dv;
  if (clamp) {
    o.up[P] = U[v];
    o.tanp[P] = tan_ms[v];
    o.dp[P] = dv;
    o.hop[P] = h[v];
  }
  // the light children in fold order (descending slot), those before the
  // heavy child first: the first FL_LIST get slots FL_LIST * P + i for
  // their chains' heads to fill; more set the overflow bit and the fold
  // reads the graph instead
  {
    uint32_t npre = 0, npost = 0;
    bool after = false;
    if (short_row) {
      // the children's bases loaded together (a row of at most FL_ROW
      // entries: the child mask covers it)
      uint32_t hb[FL_ROW];
#pragma unroll
      for (int j = 0; j < FL_ROW; ++j) {
        hb[j] = 0u;
        if (j < (int)deg && ((m >> j) & 1ull) && cc[j] != hk) hb[j] = head_base[cc[j]];
      }
#pragma unroll
      for (int j = FL_ROW - 1; j >= 0; --j) {
        if (j >= (int)deg || !((m >> j) & 1ull)) continue;
        if (cc[j] == hk) {
          after = true;
          continue;
        }
        uint32_t idx = npre + npost;
        if (idx < FL_LIST) o.pdest[hb[j]] = FL_LIST * P + idx;
        if (after) ++npost;
        else ++npre;
      }
    } else {
      for (uint32_t s = deg; s-- > 0;) {
        if (!is_child(m, deg, s, b, col, next, v)) continue;
        uint32_t ch = col[b + s];
        if (ch == hk) {
          after = true;
          continue;
        }
        uint32_t idx = npre + npost;
        if (idx < FL_LIST) o.pdest[head_base[ch]] = FL_LIST * P + idx;
        if (after) ++npost;
        else ++npre;
      }
    }
    bool over = npre + npost > FL_LIST;
    o.pcnt[P] = over ? FL_LIST_OVERFLOW : (npre | (npost << 8));
    o.parea[P] = av;
  }
  if (head) {
    const uint32_t pp = (p == v) ? FL_NONE : head_base[pair_ptr(cpp)] + pair_cnt(cpp);
    reinterpret_cast<uint2*>(o.cmeta)[2ull * hq + 1] = make_uint2(pp, hbr);
  }
}

// --------------------------------------------------------------- sweeps ---

#define FL_BATCH 8

// Drainage area of the site at chain position P, folded in the reference's
// order (generator.rs:154-159 accumulates children in reverse BFS order, so
// A_k = ((a_k + A_{c_m}) + A_{c_{m-1}}) + ... + A_{c_1} over the children in
// row order).  The heavy child's value is not known here: the fold is split
// into `pre` (the part before it) and the light children after it, posts
// p0..p3 in fold order (more than four are re-read by da_post_rest).  Light
// children head chains of a deeper round, already finished.
//
// Absent posts are -0.0, and the chain fold starts from r = -0.0 at the
// tail: x + (-0.0) == x exactly for every x (signed zeros included), so
//     A = ((pre + r) + p0) + p1
// is the reference's value at every position without a branch; positions
// with three or more posts (about 1% of sites) add p2, p3, ... after it.
#define FL_NEG0 (-0.0)
#define FL_POSTS 4

__device__ __forceinline__ void da_partial(uint32_t k, const uint32_t* __restrict__ rp,
                                           const uint32_t* __restrict__ col, const uint32_t* __restrict__ next,
                                           const unsigned long long* __restrict__ cmask,
                                           const uint32_t* __restrict__ hc, const double* __restrict__ area,
                                           const double* __restrict__ A, double& pre, double& p0, double& p1,
                                           double& p2, double& p3, uint32_t& np) {
  uint32_t b = rp[k], deg = rp[k + 1] - b, hk = hc[k];
  unsigned long long m = cmask[k];
  double acc = area[k];
  p0 = p1 = p2 = p3 = FL_NEG0;
  np = 0;
  bool after = false;
  for (uint32_t s = deg; s-- > 0;) {
    if (!is_child(m, deg, s, b, col, next, k)) continue;
    uint32_t c = col[b + s];
    if (c == hk) {
      after = true;
      continue;
    }
    double v = A[c];
    if (!after) {
      acc = acc + v;
    } else {
      // selects, not an indexed store: keeps the posts in registers
      p3 = (np == 3) ? v : p3;
      p2 = (np == 2) ? v : p2;
      p1 = (np == 1) ? v : p1;
      p0 = (np == 0) ? v : p0;
      ++np;
    }
  }
  pre = acc;
}

// The light children after the heavy one beyond the first FL_POSTS.
__device__ __forceinline__ double da_post_rest(double r, uint32_t k, const uint32_t* __restrict__ rp,
                                               const uint32_t* __restrict__ col, const uint32_t* __restrict__ next,
                                               const unsigned long long* __restrict__ cmask,
                                               const uint32_t* __restrict__ hc, const double* __restrict__ A) {
  uint32_t b = rp[k], deg = rp[k + 1] - b, hk = hc[k];
  unsigned long long m = cmask[k];
  bool after = false;
  uint32_t seen = 0;
  for (uint32_t s = deg; s-- > 0;) {
    if (!is_child(m, deg, s, b, col, next, k)) continue;
    uint32_t c = col[b + s];
    if (c == hk) {
      after = true;
      continue;
    }
    if (after) {
      if (seen >= FL_POSTS) r = r + A[c];
      ++seen;
    }
  }
  return r;
}

// The terms after ((pre + r) + p0) + p1 of a position with np >= 3 posts.
__device__ __forceinline__ double da_tail_terms(double r, uint32_t np, double c, double d, uint32_t k,
                                                const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                                                const uint32_t* __restrict__ next,
                                                const unsigned long long* __restrict__ cmask,
                                                const uint32_t* __restrict__ hc, const double* __restrict__ A) {
  r = r + c;
  if (np >= 4) r = r + d;
  if (np > FL_POSTS) r = da_post_rest(r, k, rp, col, next, cmask, hc, A);
  return r;
}

// The partial folds of a listed position from the values of its listed light
// children (v_i = the i-th listed child's A, or anything when the list has
// fewer entries: such slots are replaced by -0.0 before any arithmetic).
__device__ __forceinline__ void da_combine(uint32_t cn, double area_k, double v0, double v1, double v2, double v3,
                                           double& pre, double& p0, double& p1, double& p2, double& p3,
                                           uint32_t& np) {
  const uint32_t npre = cn & 0xffu, nl = npre + ((cn >> 8) & 0xffu);
  v0 = nl > 0 ? v0 : FL_NEG0;
  v1 = nl > 1 ? v1 : FL_NEG0;
  v2 = nl > 2 ? v2 : FL_NEG0;
  v3 = nl > 3 ? v3 : FL_NEG0;
  // pre: x + (-0.0) == x, so the terms past npre are added as -0.0
  pre = (((area_k + (npre > 0 ? v0 : FL_NEG0)) + (npre > 1 ? v1 : FL_NEG0)) + (npre > 2 ? v2 : FL_NEG0)) +
        (npre > 3 ? v3 : FL_NEG0);
  // posts: list entries npre, npre+1, ... (entries past nl read as -0.0)
  p0 = npre == 0 ? v0 : npre == 1 ? v1 : npre == 2 ? v2 : npre == 3 ? v3 : FL_NEG0;
  p1 = npre == 0 ? v1 : npre == 1 ? v2 : npre == 2 ? v3 : FL_NEG0;
  p2 = npre == 0 ? v2 : npre == 1 ? v3 : FL_NEG0;
  p3 = npre == 0 ? v3 : FL_NEG0;
  np = nl - npre;
}

#define FL_DA_BLOCK 4

// Everything the drainage-area fold reads at chain position P, in
// chain-major order (written by k_chain_perm, the children's values by the
// deeper rounds): one level of loads, no gathers.
//   k: the site; cn: npre | npost << 8, or the overflow flag;
//   v0..v3: the listed light children's A (cval[FL_LIST * P + i]);
//   area, K, d: the site's area, erodibility and receiver distance.
struct fl_da_pos {
  uint32_t k, cn;
  double v0, v1, v2, v3, area, K, d;
};

// The arrays of a drainage-area round.
struct fl_da_args {
  const uint32_t* perm;
  const uint32_t* pcnt;
  const double* cval;   // [FL_LIST n] light children's A by (position, list slot)
  const double* parea;
  const double* pK;
  const double* pdn;
  const uint32_t* pdest;  // at a chain's base position: its head's slot in cval, or FL_NONE
  double* cval_out;
  double* xp;
  // the graph, for positions with more light children than listed
  const uint32_t* rp;
  const uint32_t* col;
  const uint32_t* next;
  const unsigned long long* cmask;
  const uint32_t* hc;
  const double* area;
  double* A;
};

__device__ __forceinline__ fl_da_pos da_load(const fl_da_args& g, uint32_t P) {
  const double2* __restrict__ cv = reinterpret_cast<const double2*>(g.cval);
  const double2 a = cv[2ull * P], b = cv[2ull * P + 1];
  return fl_da_pos{g.perm[P], g.pcnt[P], a.x, a.y, b.x, b.y, g.parea[P], g.pK[P], g.pdn[P]};
}

// A site's drainage area is final: store it and the response-time input of
// its position, x = 1.0 / (K * a^0.5) * d (generator.rs:172-173, the
// reference's association).
__device__ __forceinline__ void da_store(const fl_da_args& g, uint32_t P, uint32_t k, double a, double K, double d) {
  g.A[k] = a;
  double celerity = K * rust_pow_half(a);
  g.xp[P] = 1.0 / celerity * d;
}

// Drainage area of one light depth: the sequential fold along each chain,
// tail to head, A = ((pre + A_heavy) + p0) + p1 (+ p2 + p3 ...), with the
// partial folds of each position made from its listed children's values.
// Long chains [c1, c2) take one 64-lane wavefront each: tiles of 64
// positions are prepared by the whole wavefront and staged in LDS, lane 0
// folds them eight at a time (reads first, then the dependent adds;
// positions with three or more posts marked by a ballot), the wavefront
// stores the results.  Short chains [c0, c1) take one lane each,
// FL_DA_BLOCK positions at a time.  A chain's head hands its A to its
// receiver's list slot (pdest).  Bounds from the bucket offsets of round l.
// One wavefront's LDS tile of a drainage-area round.
struct fl_da_lds {
  double p[64], a[64], b[64], c[64], d[64], r[64];
  uint32_t n[64], k[64];
};

// Synchronisation of one wavefront's lanes around its own LDS tile: a
// wavefront-scope fence (orders the LDS accesses; the wavefront's LDS
// operations execute in order) and a wave barrier.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wavefront `wave` of `nwaves` in drainage-area round l (short chains
// [c0, c1), long chains [c1, c2)); cmeta: (base, len, ...) per chain.  A
// long chain's tiles are loaded D - 1 ahead of the fold (D = 2 or 3: the
// registers a workgroup of 1024 lanes has allow two).
template <int D>
__device__ __forceinline__ void da_round(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t wave, uint32_t nwaves,
                                         uint32_t lane, fl_da_lds& L, const uint4* __restrict__ cmeta,
                                         const fl_da_args& g) {
  const uint2* __restrict__ cm2 = reinterpret_cast<const uint2*>(cmeta);
  // long chains: one per block (uniform per block).  Tiles of 64 positions
  // from the tail, software-pipelined D deep: a tile's loads are issued
  // D - 1 tiles before its fold, unconditionally and at positions clamped
  // into the chain, into registers of their own (the loop is unrolled D
  // times: moving a register a load is still writing would wait for it), so
  // waiting for one tile never waits for the tiles loaded after it.  No tile
  // input depends on this round's results (light children head chains of
  // deeper rounds).
  for (uint32_t c = c1 + wave; c < c2; c += nwaves) {
    const uint2 m = cm2[2ull * c];
    const uint32_t base = m.x, len = m.y;
    const uint32_t pd = g.pdest[base];
    // position of this lane in the tile ending at `e` (e > 0; lanes past
    // the tile read its first position)
    auto pos_of = [&](uint32_t e) {
      const uint32_t cs = e > 64 ? e - 64 : 0;
      return base + (lane < e - cs ? cs + lane : cs);
    };
    // the tile ending `ahead` tiles below `e`, or the tail tile past the head
    auto ahead_of = [&](uint32_t e, uint32_t ahead) { return e > 64 * ahead ? e - 64 * ahead : len; };
    uint32_t end = len;
    double r = FL_NEG0;  // lane 0's running value
    // fold the tile ending at `end` from cur, then load into cur the tile D
    // tiles further towards the head
    auto step = [&](fl_da_pos& cur) {
      const uint32_t cs = end > 64 ? end - 64 : 0, cnt = end - cs;
      const bool mine = lane < cnt;
      double p = 0.0, a = 0.0, b = 0.0, x = 0.0, y = 0.0;
      uint32_t np = 0;
      if (__any(mine && (cur.cn & FL_LIST_OVERFLOW))) {
        // a position with more light children than listed (rare) reads the
        // graph; kept apart so that the usual tiles issue no load here
        if (mine) {
          if (cur.cn & FL_LIST_OVERFLOW)
            da_partial(cur.k, g.rp, g.col, g.next, g.cmask, g.hc, g.area, g.A, p, a, b, x, y, np);
          else
            da_combine(cur.cn, cur.area, cur.v0, cur.v1, cur.v2, cur.v3, p, a, b, x, y, np);
        }
      } else {
        da_combine(cur.cn, cur.area, cur.v0, cur.v1, cur.v2, cur.v3, p, a, b, x, y, np);
        if (!mine) {
          p = a = b = x = y = 0.0;
          np = 0;
        }
      }
      L.p[lane] = p;
      L.a[lane] = a;
      L.b[lane] = b;
      L.c[lane] = x;
      L.d[lane] = y;
      L.n[lane] = np;
      L.k[lane] = mine ? cur.k : FL_NONE;
      // positions with a second post, and with three or more (those add p2,
      // p3 and, past FL_POSTS posts, read the graph: rare, and such "deep"
      // tiles take a fold of their own, so that no other fold holds a load)
      const unsigned long long post2 = __ballot(np >= 2), rare = __ballot(np >= 3), post4 = __ballot(np >= 4);
      const bool deep = __any(np > FL_POSTS);
      const uint32_t kk = cur.k;
      const double Kk = cur.K, dk = cur.d;
      wave_sync();
      if (lane == 0 && !deep) {
        // x + (-0.0) == x: positions add the -0.0 pads of the posts they do
        // not have, so a batch of eight folds without branches, with as
        // many dependent adds a position as the most posts in the batch
        // require (two for none or one post, up to five for four).  Every
        // operand of a batch is read from the tile before its adds: no wait
        // inside them, and no load from global memory in this loop (waiting
        // for one would wait for the tiles in flight too)
        int e = (int)cnt - 1;
        for (; e >= 7; e -= 8) {
          double tp[8], ta[8], tb[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            tp[u] = L.p[e - u];
            ta[u] = L.a[e - u];
            tb[u] = L.b[e - u];
          }
          const uint32_t sh = (uint32_t)(e - 7);
          const uint32_t m2 = (uint32_t)(post2 >> sh) & 0xffu, m3 = (uint32_t)(rare >> sh) & 0xffu;
          if (m3 == 0) {
            if (m2 == 0) {
#pragma unroll
              for (int u = 0; u < 8; ++u) {
                r = (tp[u] + r) + ta[u];
                L.r[e - u] = r;
              }
            } else {
#pragma unroll
              for (int u = 0; u < 8; ++u) {
                r = ((tp[u] + r) + ta[u]) + tb[u];
                L.r[e - u] = r;
              }
            }
          } else {
            // (the third and fourth posts in two halves of four positions:
            // registers)
            const bool four = ((uint32_t)(post4 >> sh) & 0xffu) != 0;
#pragma unroll
            for (int h = 0; h < 8; h += 4) {
              double tc[4], td[4];
#pragma unroll
              for (int q = 0; q < 4; ++q) {
                tc[q] = L.c[e - h - q];
                td[q] = L.d[e - h - q];
              }
              if (!four) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                  r = (((tp[h + q] + r) + ta[h + q]) + tb[h + q]) + tc[q];
                  L.r[e - h - q] = r;
                }
              } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                  r = ((((tp[h + q] + r) + ta[h + q]) + tb[h + q]) + tc[q]) + td[q];
                  L.r[e - h - q] = r;
                }
              }
            }
          }
        }
        for (; e >= 0; --e) {
          r = ((((L.p[e] + r) + L.a[e]) + L.b[e]) + L.c[e]) + L.d[e];
          L.r[e] = r;
        }
      }
      if (lane == 0 && deep) {
        // a tile with a position of more posts than staged: those read the
        // graph (da_tail_terms)
        for (int e = (int)cnt - 1; e >= 0; --e) {
          r = ((L.p[e] + r) + L.a[e]) + L.b[e];
          if ((rare >> e) & 1ull)
            r = da_tail_terms(r, L.n[e], L.c[e], L.d[e], L.k[e], g.rp, g.col, g.next, g.cmask, g.hc, g.A);
          L.r[e] = r;
        }
      }
      wave_sync();
      // the tile D tiles further, loaded after the fold: a fold that waits
      // for a load of its own (a deep tile) then waits for no tile's loads
      cur = da_load(g, pos_of(ahead_of(end, D)));
      if (mine) da_store(g, base + cs + lane, kk, L.r[lane], Kk, dk);
      wave_sync();  // the tile arrays are rewritten next
      end = cs;
    };
    fl_da_pos q0 = da_load(g, pos_of(end));
    fl_da_pos q1 = da_load(g, pos_of(ahead_of(end, 1)));
    if (D >= 3) {
      fl_da_pos q2 = da_load(g, pos_of(ahead_of(end, 2)));
      for (;;) {
        step(q0);
        if (end == 0) break;
        step(q1);
        if (end == 0) break;
        step(q2);
        if (end == 0) break;
      }
    } else {
      for (;;) {
        step(q0);
        if (end == 0) break;
        step(q1);
        if (end == 0) break;
      }
    }
    // lane 0's value is the head's A
    if (lane == 0 && pd != FL_NONE) g.cval_out[pd] = r;
  }
  // short chains: one lane each, the last blocks first (they hold no long
  // chain).  When one pass covers them, the 64-chain groups of the wavefronts
  // that share an XCD (workgroup index mod 8) are contiguous: chains placed
  // together hold nearby sites, whose site-ordered stores then share an L2.
  const uint32_t groups = (c1 - c0 + 63u) >> 6;
  uint32_t sw = nwaves - 1 - wave;
  if (groups <= nwaves && sw < groups) {
    const uint32_t q = groups >> 3, rq = groups & 7u, x = sw & 7u;
    sw = x * q + (x < rq ? x : rq) + (sw >> 3);
  }
  for (uint32_t c = c0 + sw * 64u + lane; c < c1; c += nwaves * 64u) {
    const uint2 m = cm2[2ull * c];
    const uint32_t base = m.x, len = m.y;
    const uint32_t pd = g.pdest[base];
    double r = FL_NEG0;
    uint32_t end = len;
    while (end > 0) {
      const uint32_t cs = end > FL_DA_BLOCK ? end - FL_DA_BLOCK : 0, cnt = end - cs;
      fl_da_pos q[FL_DA_BLOCK];
#pragma unroll
      for (uint32_t u = 0; u < FL_DA_BLOCK; ++u)
        if (u < cnt) q[u] = da_load(g, base + end - 1 - u);
#pragma unroll
      for (uint32_t u = 0; u < FL_DA_BLOCK; ++u) {
        if (u < cnt) {
          double pre, p0, p1, p2, p3;
          uint32_t np;
          if (q[u].cn & FL_LIST_OVERFLOW)
            da_partial(q[u].k, g.rp, g.col, g.next, g.cmask, g.hc, g.area, g.A, pre, p0, p1, p2, p3, np);
          else
            da_combine(q[u].cn, q[u].area, q[u].v0, q[u].v1, q[u].v2, q[u].v3, pre, p0, p1, p2, p3, np);
          r = ((pre + r) + p0) + p1;
          if (np >= 3) r = da_tail_terms(r, np, p2, p3, q[u].k, g.rp, g.col, g.next, g.cmask, g.hc, g.A);
          da_store(g, base + end - 1 - u, q[u].k, r, q[u].K, q[u].d);
        }
      }
      end = cs;
    }
    if (pd != FL_NONE) g.cval_out[pd] = r;
  }
}

__global__ void __launch_bounds__(64)
k_da_fold(const uint32_t* __restrict__ guard, uint32_t l, const uint32_t* __restrict__ boff,
          const uint4* __restrict__ cmeta, fl_da_args g) {
  // the round's bounds are read together with the guard (one load latency;
  // the bounds enter the test so that their loads stay above it: offsets are
  // ascending, c2 < c0 never holds once the guard is clear)
  const uint32_t c0 = boff[2 * l], c1 = boff[2 * l + 1], c2 = boff[2 * l + 2];
  if (*guard != 0 || c2 < c0) return;  // pointer jumping has not converged: the iteration is redone
  __shared__ fl_da_lds L;
  da_round<3>(c0, c1, c2, blockIdx.x, gridDim.x, threadIdx.x, L, cmeta, g);
}

// Drainage-area rounds maxld, maxld - 1, ..., lsplit (the small deep ones) in
// one workgroup of FL_TAIL_WAVES wavefronts, a workgroup barrier between rounds (a
// round's results are read by the next one, within the workgroup): one
// launch instead of one per round.  Rounds above `cap` are left out.
__global__ void __launch_bounds__(64 * FL_TAIL_WAVES)
k_da_tail(const uint32_t* __restrict__ guard, uint32_t lsplit, uint32_t cap, const uint32_t* __restrict__ maxld,
          const uint32_t* __restrict__ boff, const uint4* __restrict__ cmeta, fl_da_args g) {
  uint32_t top = *maxld;
  if (*guard != 0 || top > 0xfffffffeu) return;  // pointer jumping has not converged: the iteration is redone
  top = top < cap ? top : cap;
  __shared__ fl_da_lds L[FL_TAIL_WAVES];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u, nwaves = FL_TAIL_WAVES;
  for (uint32_t l = top + 1; l-- > lsplit;) {
    const uint32_t c0 = boff[2 * l], c1 = boff[2 * l + 1], c2 = boff[2 * l + 2];
    da_round<2>(c0, c1, c2, wave, nwaves, lane, L[wave], cmeta, g);
    __syncthreads();
  }
}

// Response time along the chains of one light depth, head to tail:
// rt_i = 0.0 + (rt_next + x_i) (generator.rs:173 with rt_i initially 0),
// the basin outlet first (its receiver is itself).  A head's receiver lies
// on a chain finished in an earlier round.  With clamp, also the elevation
// recurrence of generator.rs:177-202, whose max-slope limit depends on the
// receiver's new elevation.  Long chains [c1, c2) take one wavefront each
// (blocks [0, c2 - c1)): tiles of 64 positions are staged in LDS, lane 0
// runs the recurrence, the wavefront stores; short chains [c0, c1) one lane
// each (the remaining blocks).
// One wavefront's LDS tile of a response-time round.
struct fl_tau_lds {
  double x[64], u[64], t[64], d[64], tau[64], h[64];
};

// Wavefront `wave` of `nwaves` in response-time round l (short chains
// [c0, c1), long chains [c1, c2)).
//
// The reference's step rt = 0.0 + (rt_next + x) adds 0.0, which changes
// nothing unless y = rt_next + x is -0.0 (0.0 + -0.0 = +0.0; a NaN y is
// already quiet and passes unchanged).  A sum is -0.0 only when both terms
// are, and no running value is: a basin's outlet starts at 0.0 + (...),
// never -0.0, and every other chain starts from its receiver's value.  So
// the chains take the single add (one dependent operation per position
// instead of two, bit for bit the same).
template <bool clamp>
__device__ __forceinline__ void
tau_round(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t wave, uint32_t nwaves, uint32_t lane, fl_tau_lds& L,
          const uint4* __restrict__ cmeta, const double* __restrict__ xp,
          double* __restrict__ taup, const double* __restrict__ up, const double* __restrict__ tanp,
          const double* __restrict__ dp, const double* __restrict__ hop, double* __restrict__ hnp) {
  // long chains: one per block (uniform per block).  Tiles of 64 positions
  // from the head, loaded three tiles ahead of their fold, unconditionally
  // at positions clamped into the chain and into registers of their own
  // (the loop is unrolled three times), so waiting for one tile never waits
  // for the tiles loaded after it.
  for (uint32_t c = c1 + wave; c < c2; c += nwaves) {
    // base, length, position of the head's receiver (FL_NONE at the basin
    // outlet) and of the basin outlet
    const uint4 m = cmeta[c];
    const uint32_t base = m.x, len = m.y;
    // lane 0's recurrence state (at a basin outlet 0.0 + x_o, the
    // reference's 0.0 + (0.0 + x_o): the same value, see above)
    double tau = 0.0, tau_o = 0.0, h_p = 0.0, h_o = 0.0;
    const uint32_t pp = m.z;
    if (lane == 0 && pp != FL_NONE) {
      tau = taup[pp];
      if (clamp) {
        uint32_t ro = m.w;
        tau_o = taup[ro];
        h_o = hnp[ro];
        h_p = hnp[pp];
      }
    }
    struct TauIn {
      double x, u, t, d;
    };
    auto tile_in = [&](uint32_t t0) {
      const uint32_t P = base + (t0 + lane < len ? t0 + lane : len - 1);
      TauIn q{xp[P], 0.0, 0.0, 0.0};
      if (clamp) {
        q.u = up[P];
        q.t = tanp[P];
        q.d = dp[P];
      }
      return q;
    };
    uint32_t t = 0;
    // fold the tile starting at t from q, then load into q the tile three
    // tiles further
    auto step = [&](TauIn& q) {
      const uint32_t cnt = (len - t) < 64 ? (len - t) : 64;
      const bool in = lane < cnt;
      L.x[lane] = in ? q.x : 0.0;
      if (clamp) {
        L.u[lane] = in ? q.u : 0.0;
        L.t[lane] = in ? q.t : 0.0;
        L.d[lane] = in ? q.d : 0.0;
      }
      wave_sync();
      q = tile_in(t + 192);
      if (lane == 0) {
        if (!clamp) {
          // batches of eight, all within the tile (positions past the chain
          // fold zeros into values nothing reads)
          for (uint32_t j = 0; j < cnt; j += 8) {
            double tx[8];
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u) tx[u] = L.x[j + u];
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u) {
              tau = tau + tx[u];  // 0.0 + y == y here (see above)
              L.tau[j + u] = tau;
            }
          }
        } else {
          uint32_t j = 0;
          if (t == 0 && pp == FL_NONE) {
            // the outlet: rt_o = 0.0 + (0.0 + x_o); h_o' = h_o + U_o * max(0, 0)
            tau = 0.0 + (0.0 + L.x[0]);
            L.tau[0] = tau;
            double ho = hop[base];
            double nh = ho + L.u[0] * rust_max(tau - tau, 0.0);
            double slope = (nh - ho) / L.d[0];
            if (slope > L.t[0]) nh = ho + L.t[0] * L.d[0];
            L.h[0] = nh;
            h_p = nh;
            h_o = nh;
            tau_o = tau;
            j = 1;
          }
          while (j < cnt) {
            const uint32_t mm = (cnt - j) < 8 ? (cnt - j) : 8;
            double tx[8], tu[8], tt[8], td[8];
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u) {
              if (u < mm) {
                tx[u] = L.x[j + u];
                tu[u] = L.u[j + u];
                tt[u] = L.t[j + u];
                td[u] = L.d[j + u];
              }
            }
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u) {
              if (u < mm) {
                tau = tau + tx[u];  // 0.0 + y == y here (see above)
                double nh = h_o + tu[u] * rust_max(tau - tau_o, 0.0);
                double slope = (nh - h_p) / td[u];
                if (slope > tt[u]) nh = h_p + tt[u] * td[u];
                L.tau[j + u] = tau;
                L.h[j + u] = nh;
                h_p = nh;
              }
            }
            j += mm;
          }
        }
      }
      wave_sync();
      if (in) {
        taup[base + t + lane] = L.tau[lane];
        if (clamp) hnp[base + t + lane] = L.h[lane];
      }
      wave_sync();  // the tile arrays are rewritten next
      t += 64;
    };
    TauIn q0 = tile_in(0), q1 = tile_in(64), q2 = tile_in(128);
    for (;;) {
      step(q0);
      if (t >= len) break;
      step(q1);
      if (t >= len) break;
      step(q2);
      if (t >= len) break;
    }
  }
  // short chains: one lane each, the last blocks first (they hold no long
  // chain); the groups of the wavefronts that share an XCD are contiguous
  // when one pass covers them
  const uint32_t groups = (c1 - c0 + 63u) >> 6;
  uint32_t sw = nwaves - 1 - wave;
  if (groups <= nwaves && sw < groups) {
    const uint32_t q = groups >> 3, rq = groups & 7u, x = sw & 7u;
    sw = x * q + (x < rq ? x : rq) + (sw >> 3);
  }
  for (uint32_t c = c0 + sw * 64u + lane; c < c1; c += nwaves * 64u) {
    const uint4 m = cmeta[c];
    const uint32_t base = m.x, len = m.y, pp = m.z;
    double tau, tau_o = 0.0, h_p = 0.0, h_o = 0.0;
    uint32_t t = 0;
    if (pp == FL_NONE) {
      tau = 0.0 + (0.0 + xp[base]);
      taup[base] = tau;
      if (clamp) {
        double ho = hop[base];
        double nh = ho + up[base] * rust_max(tau - tau, 0.0);
        double tm = tanp[base], d = dp[base];
        double slope = (nh - ho) / d;  // the receiver is the outlet itself, not yet updated
        if (slope > tm) nh = ho + tm * d;
        hnp[base] = nh;
        h_p = nh;
        h_o = nh;
        tau_o = tau;
      }
      t = 1;
    } else {
      tau = taup[pp];
      if (clamp) {
        uint32_t ro = m.w;
        tau_o = taup[ro];
        h_o = hnp[ro];
        h_p = hnp[pp];
      }
    }
    if (!clamp) {
      while (t < len) {
        uint32_t cnt = (len - t) < FL_BATCH ? (len - t) : FL_BATCH;
        // every slot loads (slots past the chain read its last position): no
        // branch per load, so the batch's loads are in flight together
        double vx[FL_BATCH];
#pragma unroll
        for (uint32_t q = 0; q < FL_BATCH; ++q) vx[q] = xp[base + t + (q < cnt ? q : cnt - 1)];
#pragma unroll
        for (uint32_t q = 0; q < FL_BATCH; ++q) {
          if (q < cnt) {
            tau = tau + vx[q];  // 0.0 + y == y here (see tau_round)
            taup[base + t + q] = tau;
          }
        }
        t += cnt;
      }
    } else {
      while (t < len) {
        uint32_t cnt = (len - t) < FL_BATCH ? (len - t) : FL_BATCH;
        double vx[FL_BATCH], vu[FL_BATCH], vt[FL_BATCH], vd[FL_BATCH];
#pragma unroll
        for (uint32_t q = 0; q < FL_BATCH; ++q) {
          const uint32_t P = base + t + (q < cnt ? q : cnt - 1);  // every slot loads (see above)
          vx[q] = xp[P];
          vu[q] = up[P];
          vt[q] = tanp[P];
          vd[q] = dp[P];
        }
#pragma unroll
        for (uint32_t q = 0; q < FL_BATCH; ++q) {
          if (q < cnt) {
            tau = tau + vx[q];  // 0.0 + y == y here (see tau_round)
            double nh = h_o + vu[q] * rust_max(tau - tau_o, 0.0);
            double slope = (nh - h_p) / vd[q];
            if (slope > vt[q]) nh = h_p + vt[q] * vd[q];
            taup[base + t + q] = tau;
            hnp[base + t + q] = nh;
            h_p = nh;
          }
        }
        t += cnt;
      }
    }
  }
}

template <bool clamp>
__global__ void __launch_bounds__(64)
k_tau_fold(const uint32_t* __restrict__ guard, uint32_t l, const uint32_t* __restrict__ boff,
           const uint4* __restrict__ cmeta, const double* __restrict__ xp, double* __restrict__ taup,
           const double* __restrict__ up, const double* __restrict__ tanp, const double* __restrict__ dp,
           const double* __restrict__ hop, double* __restrict__ hnp) {
  // the round's bounds are read together with the guard (one load latency;
  // the bounds enter the test so that their loads stay above it: offsets are
  // ascending, c2 < c0 never holds once the guard is clear)
  const uint32_t c0 = boff[2 * l], c1 = boff[2 * l + 1], c2 = boff[2 * l + 2];
  if (*guard != 0 || c2 < c0) return;  // pointer jumping has not converged: the iteration is redone
  __shared__ fl_tau_lds L;
  tau_round<clamp>(c0, c1, c2, blockIdx.x, gridDim.x, threadIdx.x, L, cmeta, xp, taup, up, tanp, dp, hop, hnp);
}

// Response-time rounds lsplit, lsplit + 1, ..., maxld (the small deep ones)
// in one workgroup of FL_TAIL_WAVES wavefronts, a workgroup barrier between rounds (a
// round reads the response times and elevations of the previous ones): one
// launch instead of one per round.  Rounds above `cap` are left out.
template <bool clamp>
__global__ void __launch_bounds__(64 * FL_TAIL_WAVES)
k_tau_tail(const uint32_t* __restrict__ guard, uint32_t lsplit, uint32_t cap, const uint32_t* __restrict__ maxld,
           const uint32_t* __restrict__ boff, const uint4* __restrict__ cmeta, const double* __restrict__ xp,
          double* __restrict__ taup, const double* __restrict__ up, const double* __restrict__ tanp,
          const double* __restrict__ dp, const double* __restrict__ hop, double* __restrict__ hnp) {
  uint32_t top = *maxld;
  if (*guard != 0 || top > 0xfffffffeu) return;  // pointer jumping has not converged: the iteration is redone
  top = top < cap ? top : cap;
  __shared__ fl_tau_lds L[FL_TAIL_WAVES];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u, nwaves = FL_TAIL_WAVES;
  for (uint32_t l = lsplit; l <= top; ++l) {
    const uint32_t c0 = boff[2 * l], c1 = boff[2 * l + 1], c2 = boff[2 * l + 2];
    tau_round<clamp>(c0, c1, c2, wave, nwaves, lane, L[wave], cmeta, xp, taup, up, tanp, dp, hop, hnp);
    __syncthreads();
  }
}

// Back to site order: response time, new elevation (computed here when no
// site has a max slope: h_i' = h_o' + U_i * max(rt_i - rt_o, 0) with
// h_o' = h_o + U_o * max(0, 0), generator.rs:178-180) and the changed flag
// (generator.rs:201).  One thread per site, which finds its chain position
// from its converged chain pair (head, position): the site-ordered accesses
// are coalesced, the position-ordered ones are gathers.  Sites outside every
// basin were settled by k_chain_perm.  In a batch (tid != null) a terrain
// that stopped iterating keeps its elevations and each terrain has its own
// flag.
__global__ void k_tau_scatter(uint32_t bs, const uint32_t* __restrict__ guard, uint32_t n, const unsigned long long* __restrict__ cp,
                              const unsigned long long* __restrict__ rt, const uint8_t* __restrict__ is_outlet,
                              const uint32_t* __restrict__ head_base, const double* __restrict__ taup,
                              const double* __restrict__ hnp, const double* __restrict__ U,
                              const double* __restrict__ h, double* __restrict__ h2, double* __restrict__ tau,
                              int* __restrict__ flags, int clamp, const uint32_t* __restrict__ tid,
                              const uint8_t* __restrict__ act, uint32_t* __restrict__ tchanged,
                              uint8_t* __restrict__ chg, uint8_t* __restrict__ dirty,
                              const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col) {
  const uint32_t v = gid(bs);
  // the basin marks were read by this iteration's k_children: clear them for
  // the next iteration's receivers (also when the iteration is redone: a
  // redo marks again from its own receivers)
  if (v < n) dirty[v] = 0;
  // the site's own words are read together with the guard (one load
  // latency), then the second level (outlet flag, positions, the outlet's
  // elevation and uplift) together
  const bool in = v < n;
  const bool on = in && act[v];
  const double hv = in ? h[v] : 0.0;
  unsigned long long rv = 0ull, cv = 0ull;
  uint32_t rb = 0u, re = 0u, g = 0u;  // the row, for the change marks
  double uv = 0.0;
  if (on) {
    rv = rt[v];
    cv = cp[v];
    rb = rp[v];
    re = rp[v + 1];
    uv = U[v];
    g = tid ? tid[v] : 0u;
  }
  if (*guard != 0) return;  // pointer jumping has not converged: the iteration is redone
  bool changed = false;
  if (in && !on) {
    // a site of a basin that keeps its tree (or of a terrain that stopped)
    // keeps its elevation
    h2[v] = hv;
  } else if (on) {
    const uint32_t o = pair_ptr(rv);
    const uint8_t outlet = is_outlet[o];
    // (a site of a lake basin reads these for nothing; its values were
    // written by k_chain_perm)
    const uint32_t P = head_base[pair_ptr(cv)] + pair_cnt(cv);
    const uint32_t Po = clamp ? 0u : head_base[o];
    const double h_o = clamp ? 0.0 : h[o], u_o = clamp ? 0.0 : U[o];
    if (outlet) {
      const double tk = taup[P];
      double nh;
      if (clamp) {
        nh = hnp[P];
      } else {
        const double to = taup[Po];
        const double ho = h_o + u_o * rust_max(to - to, 0.0);
        nh = (v == o) ? ho : ho + uv * rust_max(tk - to, 0.0);
      }
      tau[v] = tk;
      h2[v] = nh;
      changed = nh != hv;
      if (changed && tchanged) tchanged[g] = 1;
    }
  }
  // neighbourhood change flags for the next iteration's receivers (zeroed by
  // this iteration's k_receivers; sites outside every basin kept their
  // elevation in k_chain_perm): a changed site marks itself and its
  // neighbours (rows are symmetric)
  if (changed) {
    chg[v] = 1;
    for (uint32_t k = rb; k < re; k += FL_ROW) {
      uint32_t c[FL_ROW];  // loaded together (past the row's end: its last entry again)
#pragma unroll
      for (int j = 0; j < FL_ROW; ++j) c[j] = col[min(k + j, re - 1)];
#pragma unroll
      for (int j = 0; j < FL_ROW; ++j) chg[c[j]] = 1;
    }
  }
  if (__any(changed) && (threadIdx.x & 63u) == 0) flags[FL_FLAG_CHANGED] = 1;
}

// ----------------------------------------------------------- compaction ---

namespace {
// Terrain of new site v: the last q with new_off[q] <= v.
__device__ __forceinline__ uint32_t terrain_of(const fl_compact_map& m, uint32_t v) {
  uint32_t lo = 0, hi = m.count;  // new_off[lo] <= v < new_off[hi]
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (m.new_off[mid] <= v) lo = mid;
    else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ uint32_t shift_site(uint32_t x, const fl_compact_map& m, uint32_t q) {
  return x == FL_NONE ? FL_NONE : x - m.old_off[q] + m.new_off[q];
}
}  // namespace

// New sites [v0, v1) of a union of n_new sites and nnz_new entries, taken
// from the terrains of one source union (m, terrain q of the map becomes
// terrain tid0 + q of the new union).
__global__ void k_compact_sites(uint32_t bs, uint32_t v0, uint32_t v1, uint32_t n_new, uint32_t nnz_new, uint32_t tid0,
                                fl_compact_map m, fl_dev_terrain ot, fl_dev_work ow, fl_dev_terrain nt,
                                fl_dev_work nw, uint32_t* __restrict__ new_tid) {
  uint32_t v = v0 + gid(bs);
  if (v >= v1) return;
  uint32_t q = terrain_of(m, v);
  uint32_t src = v - m.new_off[q] + m.old_off[q];
  nt.row_ptr[v] = ot.row_ptr[src] - m.old_eoff[q] + m.new_eoff[q];
  if (v == n_new - 1) nt.row_ptr[n_new] = nnz_new;
  nt.area[v] = ot.area[src];
  nt.erod[v] = ot.erod[src];
  nt.uplift[v] = ot.uplift[src];
  nt.tan_ms[v] = ot.tan_ms[src];
  nt.init_h[v] = ot.init_h[src];
  nt.is_outlet[v] = ot.is_outlet[src];
  nt.rank[v] = shift_site(ot.rank[src], m, q);
  nt.order[v] = shift_site(ot.order[src], m, q);  // rank positions share the terrain's site range
  nw.h[v] = ow.h[src];
  nw.A[v] = ow.A[src];
  new_tid[v] = tid0 + q;
}

__global__ void k_compact_entries(uint32_t bs, uint32_t v0, uint32_t v1, fl_compact_map m, fl_dev_terrain ot,
                                  fl_dev_terrain nt) {
  uint32_t v = v0 + gid(bs);
  if (v >= v1) return;
  uint32_t q = terrain_of(m, v);
  uint32_t src = v - m.new_off[q] + m.old_off[q];
  uint32_t b = ot.row_ptr[src], e = ot.row_ptr[src + 1];
  uint32_t d = m.new_eoff[q] - m.old_eoff[q];
  for (uint32_t k = b; k < e; ++k) {
    nt.col[k + d] = ot.col[k] - m.old_off[q] + m.new_off[q];
    nt.len[k + d] = ot.len[k];
  }
}

// Verification: the reference's traversal of one basin
// (DrainageBasin::construct, drainage_basin.rs:13-36): start at the outlet,
// then for each site in order append its children in row order.  One
// workgroup; children counts are prefix-summed per chunk of 1024 sites so
// the appended order equals the sequential one.
__global__ void __launch_bounds__(1024)
k_basin_order(uint32_t outlet, uint32_t cap, const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
              const uint32_t* __restrict__ next, const unsigned long long* __restrict__ cmask,
              uint32_t* __restrict__ order, uint32_t* __restrict__ len_out) {
  __shared__ uint32_t s_scan[1024];
  __shared__ uint32_t s_tail;
  uint32_t tid = threadIdx.x;
  if (tid == 0) {
    order[0] = outlet;
    s_tail = 1;
  }
  __syncthreads();
  uint32_t head = 0;
  while (head < s_tail) {
    uint32_t end = s_tail;
    for (uint32_t chunk = head; chunk < end; chunk += 1024) {
      uint32_t i = chunk + tid;
      uint32_t u = FL_NONE, cnt = 0, b = 0, deg = 0;
      unsigned long long m = 0;
      if (i < end) {
        u = order[i];
        b = rp[u];
        deg = rp[u + 1] - b;
        m = cmask[u];
        for (uint32_t s = 0; s < deg; ++s) cnt += is_child(m, deg, s, b, col, next, u) ? 1u : 0u;
      }
      s_scan[tid] = cnt;
      __syncthreads();
      for (uint32_t off = 1; off < 1024; off <<= 1) {
        uint32_t t = tid >= off ? s_scan[tid - off] : 0;
        __syncthreads();
        s_scan[tid] += t;
        __syncthreads();
      }
      uint32_t base = s_tail + s_scan[tid] - cnt;
      if (i < end && cnt) {
        for (uint32_t s = 0; s < deg; ++s)
          if (is_child(m, deg, s, b, col, next, u) && base < cap) order[base++] = col[b + s];
      }
      __syncthreads();
      if (tid == 1023) s_tail = min(cap, s_tail + s_scan[1023]);
      __syncthreads();
    }
    head = end;
  }
  if (tid == 0) *len_out = s_tail;
}

// Zero the control block and the first `nb` bucket entries.
__global__ void k_clear_iter(uint32_t bs, uint32_t nb, uint32_t* __restrict__ ctrl, uint32_t* __restrict__ bcnt,
                             uint32_t* __restrict__ bnodes, uint32_t* __restrict__ bcur,
                             uint32_t* __restrict__ bncur) {
  uint32_t i = gid(bs);
  if (i < FL_CTRL_WORDS) ctrl[i] = 0;
  if (i < nb) {
    bcnt[i] = 0;
    bnodes[i] = 0;
    bcur[i] = 0;
    bncur[i] = 0;
  }
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

void fl_profiler::record(hipStream_t s, const char* name, hipEvent_t a, hipEvent_t b, uint64_t units) {
  pending.push_back({name, a, b, units});  // flushed on query: the host may fill in units later
  (void)s;
}

void fl_profiler::flush() {
  for (const Rec& r : pending) {
    float ms = 0.f;
    (void)hipEventSynchronize(r.b);
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    auto& t = totals[r.name];
    t.launches += 1;
    t.ms += ms;
    t.units += r.units;
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

// `units`: the launch's work items for the profile (sites or positions).  A
// profiled launch passes its two events to the dispatch itself
// (hipExtLaunchKernelGGL), which stamps them with the kernel's own start and
// end: the kernel's execution time, without the queue's dispatch gap.
#define FL_LAUNCH_GRID(kernel, grid, block, units, stream, ...)                    \
  do {                                                                            \
    fl_profiler* prof_ = fl_current_profiler;                                     \
    if (prof_) {                                                                  \
      hipEvent_t a_ = prof_->get(), b_ = prof_->get();                            \
      hipExtLaunchKernelGGL(kernel, (grid), (block), 0, (stream), a_, b_, 0,      \
                            __VA_ARGS__);                                         \
      prof_->record((stream), #kernel, a_, b_, (units));                          \
    } else {                                                                      \
      hipLaunchKernelGGL(kernel, (grid), (block), 0, (stream), __VA_ARGS__);     \
    }                                                                             \
  } while (0)

// Per-site kernels: one thread per item in workgroups of fl_site_block
// threads (the tuned workgroup size of the call in progress).
thread_local uint32_t fl_site_block = 256;
// The workgroup size is also the kernel's first argument, taken into a local
// first: a launch recorded for later replay holds the value of its time.
#define FL_LAUNCH_U(kernel, count, units, stream, ...)                                                    \
  do {                                                                                                      \
    const uint32_t bs_ = fl_site_block;                                                                     \
    if ((count) > 0)                                                                                        \
      FL_LAUNCH_GRID(kernel, grid_for((count), bs_), dim3(bs_), (units), (stream), bs_, __VA_ARGS__);      \
  } while (0)
#define FL_LAUNCH(kernel, count, stream, ...) FL_LAUNCH_U(kernel, count, count, stream, __VA_ARGS__)

void fl_launch_receivers(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, bool keep_unchanged,
                         bool mark) {
  FL_LAUNCH(k_receivers, t.n, s, t.n, t.row_ptr, t.col, t.len, w.h, t.is_outlet, w.next, w.dnext, w.flags,
            keep_unchanged ? w.chg_prev : (const uint8_t*)nullptr, w.next_prev, w.dnext_prev, w.root_prev,
            mark ? w.dirty : (uint8_t*)nullptr, w.chg);
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

void fl_launch_dirty(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w) {
  FL_LAUNCH(k_dirty, t.n, s, t.n, w.next, w.dnext, w.next_prev, w.dnext_prev, w.root_prev, w.dirty);
}

void fl_launch_children(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, bool full, const uint32_t* tid,
                        const uint8_t* t_active) {
  FL_LAUNCH(k_children, t.n, s, t.n, t.row_ptr, t.col, w.next, w.A, (unsigned long long*)w.cmask, w.hc, w.act,
            full ? (const uint8_t*)nullptr : (const uint8_t*)w.dirty, w.root_prev, tid, t_active,
            (unsigned long long*)w.cp, (unsigned long long*)w.rt, w.root);
}

void fl_launch_chain_jump(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, uint32_t round) {
  uint32_t r = round < FL_MAX_JUMP_ROUNDS ? round : FL_MAX_JUMP_ROUNDS - 1;
  uint32_t* fset = w.ctrl + FL_CTRL_JUMPS + r;
  const uint32_t* prev = r > 0 ? w.ctrl + FL_CTRL_JUMPS + r - 1 : nullptr;
  FL_LAUNCH(k_chain_jump, t.n, s, t.n, (unsigned long long*)w.cp, (unsigned long long*)w.rt, fset, prev, w.act);
}

void fl_launch_chain_count(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w) {
  uint32_t per = FL_BUCKET_BLOCK * FL_BUCKET_PER_THREAD;
  FL_LAUNCH_GRID(k_bucket_count, dim3((t.n + per - 1) / per), dim3(FL_BUCKET_BLOCK), t.n, s, w.guard, t.n, w.hc,
                 (const unsigned long long*)w.cp, (const unsigned long long*)w.rt, t.is_outlet, w.bcnt, w.bnodes, w.counters + FL_CNT_MAXLD,
                 w.long_min, w.act);
}

void fl_launch_chain_layout(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, uint32_t rounds) {
  uint64_t fill = std::min<uint64_t>(2ull * rounds, 2ull * t.n + 2);
  FL_LAUNCH_GRID(k_bucket_scan, dim3(1), dim3(FL_SCAN_BLOCK), 0, s, t.n, (uint32_t)fill, w.ctrl, w.bcnt, w.bnodes, w.boff,
                 w.bnoff);
  uint32_t per = FL_BUCKET_BLOCK * FL_BUCKET_PER_THREAD;
  FL_LAUNCH_GRID(k_bucket_place, dim3((t.n + per - 1) / per), dim3(FL_BUCKET_BLOCK), t.n, s, w.guard, t.n, w.hc,
                 (const unsigned long long*)w.cp, (const unsigned long long*)w.rt, t.is_outlet, w.boff, w.bnoff,
                 w.bcur, w.bncur, w.cmeta, w.head_chain, w.pdest, w.head_base, w.long_min, w.act);
  fl_perm_out o{w.perm, w.pcnt, w.pdest, w.parea, w.pK, w.pdn, w.up, w.tanp, w.dp, w.hop, w.cmeta};
  FL_LAUNCH(k_chain_perm, t.n, s, w.guard, t.n, (const unsigned long long*)w.cp, (const unsigned long long*)w.rt,
            w.next, t.is_outlet, w.head_base, w.head_chain, o, t.area, t.erod, w.dnext, t.uplift, t.tan_ms,
            t.any_max_slope ? 1 : 0, w.A, w.tau, w.h, w.h2, w.counters + FL_CNT_MAXLD, w.bcnt, w.bnodes, w.bcur,
            w.bncur, t.row_ptr, t.col, (const unsigned long long*)w.cmask, w.hc, w.act, w.root);
}

namespace {
fl_da_args da_args(const fl_dev_terrain& t, const fl_dev_work& w) {
  return fl_da_args{w.perm, w.pcnt,   w.cval, w.parea, w.pK, w.pdn, w.pdest, w.cval, w.xp,
                    t.row_ptr, t.col, w.next, (const unsigned long long*)w.cmask, w.hc, t.area, w.A};
}
}  // namespace

void fl_launch_da_round(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, uint32_t l,
                        const fl_round_grid& g, bool clamp) {
  (void)clamp;  // the max-slope inputs are laid out by k_chain_perm
  if (g.fold_blocks)
    FL_LAUNCH_GRID(k_da_fold, dim3(g.fold_blocks), dim3(64), g.positions, s, w.guard, l, w.boff, w.cmeta,
                   da_args(t, w));
}

void fl_launch_da_tail(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, uint32_t lsplit, uint32_t cap,
                       uint64_t units, bool clamp) {
  (void)clamp;
  FL_LAUNCH_GRID(k_da_tail, dim3(1), dim3(64 * FL_TAIL_WAVES), units, s, w.guard, lsplit, cap,
                 w.counters + FL_CNT_MAXLD, w.boff, w.cmeta, da_args(t, w));
}

void fl_launch_tau_tail(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, uint32_t lsplit, uint32_t cap,
                        uint64_t units, bool clamp) {
  (void)t;
  if (clamp)
    FL_LAUNCH_GRID(k_tau_tail<true>, dim3(1), dim3(64 * FL_TAIL_WAVES), units, s, w.guard, lsplit, cap,
                   w.counters + FL_CNT_MAXLD, w.boff, w.cmeta, w.xp, w.taup, w.up, w.tanp, w.dp, w.hop, w.hnp);
  else
    FL_LAUNCH_GRID(k_tau_tail<false>, dim3(1), dim3(64 * FL_TAIL_WAVES), units, s, w.guard, lsplit, cap,
                   w.counters + FL_CNT_MAXLD, w.boff, w.cmeta, w.xp, w.taup, w.up, w.tanp, w.dp, w.hop, w.hnp);
}

void fl_launch_tau_round(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, uint32_t l,
                         const fl_round_grid& g, bool clamp) {
  (void)t;
  if (!g.fold_blocks) return;
  if (clamp)
    FL_LAUNCH_GRID(k_tau_fold<true>, dim3(g.fold_blocks), dim3(64), g.positions, s, w.guard, l, w.boff, w.cmeta,
                   w.xp, w.taup, w.up, w.tanp, w.dp, w.hop, w.hnp);
  else
    FL_LAUNCH_GRID(k_tau_fold<false>, dim3(g.fold_blocks), dim3(64), g.positions, s, w.guard, l, w.boff, w.cmeta,
                   w.xp, w.taup, w.up, w.tanp, w.dp, w.hop, w.hnp);
}

void fl_launch_tau_scatter(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, bool clamp,
                           const uint32_t* tid, const uint8_t* active, uint32_t* tchanged) {
  FL_LAUNCH(k_tau_scatter, t.n, s, w.guard, t.n, (const unsigned long long*)w.cp, (const unsigned long long*)w.rt,
            t.is_outlet, w.head_base, w.taup, w.hnp, t.uplift, w.h, w.h2, w.tau, w.flags, clamp ? 1 : 0, tid, w.act,
            tchanged, w.chg, w.dirty, t.row_ptr, t.col);
}

void fl_launch_clear_iter(hipStream_t s, const fl_dev_work& w, uint32_t nb) {
  uint32_t cnt = nb > FL_CTRL_WORDS ? nb : FL_CTRL_WORDS;
  FL_LAUNCH(k_clear_iter, cnt, s, nb, w.ctrl, w.bcnt, w.bnodes, w.bcur, w.bncur);
}

void fl_launch_fill_u32(hipStream_t s, uint32_t* p, uint64_t count, uint32_t v) {
  FL_LAUNCH(k_fill_u32, count, s, (uint32_t)count, p, v);
}

void fl_launch_basin_order(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, uint32_t outlet,
                           uint32_t* order, uint32_t* len_out) {
  FL_LAUNCH_GRID(k_basin_order, dim3(1), dim3(1024), t.n, s, outlet, t.n, t.row_ptr, t.col, w.next,
                 (const unsigned long long*)w.cmask, order, len_out);
}

void fl_launch_compact(hipStream_t s, const fl_compact_map& m, const fl_dev_terrain& ot, const fl_dev_work& ow,
                       const fl_dev_terrain& nt, const fl_dev_work& nw, uint32_t* new_tid, uint32_t v0, uint32_t v1,
                       uint32_t tid0) {
  FL_LAUNCH(k_compact_sites, v1 - v0, s, v0, v1, nt.n, nt.nnz, tid0, m, ot, ow, nt, nw, new_tid);
  FL_LAUNCH(k_compact_entries, v1 - v0, s, v0, v1, m, ot, nt);
}

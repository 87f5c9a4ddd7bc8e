// fl_order.hip -- device part of the lake-traversal visit order.
//
// The traversal of stream_tree.rs:184-240 pops the outlets first (all keys
// 0.0) and then follows Prim's algorithm from the outlet set.  When the
// edge lengths of a terrain are positive and pairwise distinct, its Prim
// phase is determined by the minimum spanning forest of the graph with the
// terrain's outlets contracted into one vertex S.  This file
//   * checks the lengths per terrain (a hash set of their bit patterns);
//   * builds the forest (Boruvka, each terrain's outlets pre-merged into one
//     component) and sorts its edges by length;
//   * turns the forest into Prim ranks without a sequential pass.
//
// Prim ranks from the forest.  Root the forest at S and give every other
// site v the weight w(v) of the edge to its parent (the edge's position in
// length order).  For the path S = u_0, u_1, ..., u_k = v let key(v) be the
// weights w(u_i) that exceed every weight after them on the path (the path's
// suffix maxima, a decreasing sequence ending in w(v)).  Prim from S visits
// the sites in lexicographic order of their keys, a key before every key it
// is a prefix of: the heaviest edge e of a subtree separates the part that
// holds the subtree's entry point, which Prim completes first, from the part
// below e, which it enters at e's lower end; every key below e starts with
// w(e), every other key with a smaller weight, and both parts recurse.
// key(v) is key(q(v)) followed by w(v), where q(v) is the nearest proper
// ancestor of v with a heavier parent edge (S, of infinite weight, if none).
// So the order is the preorder of the tree of q with each site's children
// taken in increasing weight.  Steps:
//   1. rooting: an Euler tour of the forest (arcs sorted by source, the
//      successor of arc u->v the arc after v->u in v's list), cut where it
//      returns to S, ranked by pointer jumping; of the two arcs of an edge
//      the earlier one points away from S;
//   2. q by pointer jumping: while w(q(v)) < w(v), q(v) <- q(q(v)) (every
//      site skipped is lighter than q(v), hence lighter than v);
//   3. the children of each q-tree site in increasing weight: the forest's
//      edges are in length order already, a stable radix sort by q groups
//      them;
//   4. preorder: an Euler tour of the q-tree ranked by pointer jumping,
//      counting the down steps from the tour's start.
// The outlet phase (host, fl_outlet_pop_order) ranks the outlets; terrains
// whose lengths repeat or are not positive are ranked by the exact heap
// simulation on the host (fl_lake_visit_order).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include <chrono>
#include <string>
#include <vector>

#include "fl_internal.h"
#include "fl_kernels.h"

#define FL_NONE 0xffffffffu

namespace {

// Workgroup b runs on XCD b % 8; each XCD takes one contiguous eighth of the
// index range (a bijection of [0, gridDim.x)), so neighbouring sites share an L2.
__device__ __forceinline__ uint32_t gid() {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u;
  return (x * q + (x < r ? x : r) + (b >> 3)) * blockDim.x + threadIdx.x;
}

__device__ __forceinline__ unsigned long long len_bits(double w) {
  return (unsigned long long)__double_as_longlong(w);
}

__device__ __forceinline__ uint32_t hash64(unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return (uint32_t)k;
}

// (pointer, count) pairs of the list ranking: pointer in the low word.
__device__ __forceinline__ unsigned long long pr(uint32_t ptr, uint32_t cnt) {
  return ((unsigned long long)cnt << 32) | ptr;
}
__device__ __forceinline__ uint32_t pr_ptr(unsigned long long p) { return (uint32_t)p; }
__device__ __forceinline__ uint32_t pr_cnt(unsigned long long p) { return (uint32_t)(p >> 32); }

__global__ void k_fill32(uint32_t n, uint32_t* __restrict__ p, uint32_t v) {
  uint32_t i = gid();
  if (i < n) p[i] = v;
}

__global__ void k_fill64(uint32_t n, unsigned long long* __restrict__ p, unsigned long long v) {
  uint32_t i = gid();
  if (i < n) p[i] = v;
}

// ------------------------------------------------------- length check ----

// Every undirected edge (from its smaller endpoint): a terrain's lengths
// must be positive (not NaN) and pairwise distinct for the fast form; a hash
// set of their bit patterns per terrain (terrain g's table at toff[g], mask
// tmask[g]) finds repeats.  Equal lengths in different terrains do not
// interact.
__global__ void k_len_check(uint32_t n, const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                            const double* __restrict__ len, const uint32_t* __restrict__ tid,
                            const uint32_t* __restrict__ toff, const uint32_t* __restrict__ tmask,
                            unsigned long long* __restrict__ table, uint32_t* __restrict__ bad) {
  uint32_t i = gid();
  if (i >= n) return;
  const uint32_t g = tid ? tid[i] : 0u;
  unsigned long long* tab = table + toff[g];
  const uint32_t mask = tmask[g];
  for (uint32_t e = rp[i]; e < rp[i + 1]; ++e) {
    if (col[e] <= i) continue;
    double w = len[e];
    if (!(w > 0.0)) {
      bad[g] = 1;
      continue;
    }
    unsigned long long key = len_bits(w);
    uint32_t h = hash64(key) & mask;
    for (uint32_t probe = 0; probe <= mask; ++probe) {
      unsigned long long prev = atomicCAS(tab + h, 0ull, key);
      if (prev == 0ull) break;
      if (prev == key) {
        bad[g] = 1;
        break;
      }
      h = (h + 1) & mask;
    }
  }
}

// ------------------------------------------------------------- Boruvka ----

__global__ void k_bv_init(uint32_t n, const uint32_t* __restrict__ init_comp, uint32_t* __restrict__ comp,
                          unsigned long long* __restrict__ cbest) {
  uint32_t v = gid();
  if (v >= n) return;
  comp[v] = init_comp[v];
  cbest[v] = ~0ull;
}

// Lightest edge leaving each site's component; per component the minimum
// over its sites (positive lengths order as their bit patterns).
__global__ void k_bv_best(uint32_t n, const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                          const double* __restrict__ len, const uint32_t* __restrict__ comp,
                          unsigned long long* __restrict__ cbest, uint32_t* __restrict__ vbest) {
  uint32_t v = gid();
  if (v >= n) return;
  const uint32_t cv = comp[v], b0 = rp[v], e1 = rp[v + 1];
  uint32_t be = FL_NONE;
  unsigned long long best = ~0ull;
  // eight entries at a time, every load of a group issued before the first
  // is used (past the row's end: its last entry again, not evaluated)
  for (uint32_t e = b0; e < e1; e += 8) {
    uint32_t c[8], k[8];
    double w[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint32_t ej = e + j < e1 ? e + j : e1 - 1;
      c[j] = col[ej];
      w[j] = len[ej];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) k[j] = comp[c[j]];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (e + j >= e1 || k[j] == cv) continue;
      const unsigned long long b = len_bits(w[j]);
      if (b < best) {
        best = b;
        be = e + j;
      }
    }
  }
  vbest[v] = be;
  // the sites of a large component all aim at one word: an atomic only where
  // the word (possibly an earlier value: it only decreases) is not already
  // as small
  if (be != FL_NONE && best < __atomic_load_n(cbest + cv, __ATOMIC_RELAXED)) atomicMin(cbest + cv, best);
}

// The site owning its component's lightest edge records it (unique:
// lengths are distinct and the edge leaves the component from one side).
__global__ void k_bv_pick(uint32_t n, const double* __restrict__ len, const uint32_t* __restrict__ comp,
                          const unsigned long long* __restrict__ cbest, const uint32_t* __restrict__ vbest,
                          uint32_t* __restrict__ cedge, uint32_t* __restrict__ csrc) {
  uint32_t v = gid();
  if (v >= n) return;
  uint32_t be = vbest[v];
  if (be == FL_NONE) return;
  uint32_t c = comp[v];
  if (len_bits(len[be]) == cbest[c]) {
    cedge[c] = be;
    csrc[c] = v;
  }
}

// Hook each component along its lightest edge.  Two components choosing the
// same edge form the only possible cycle: the larger label hooks.  Each hook
// adds its edge to the forest.
__global__ void k_bv_hook(uint32_t n, const uint32_t* __restrict__ col, const uint32_t* __restrict__ comp,
                          const unsigned long long* __restrict__ cbest, const uint32_t* __restrict__ cedge,
                          const uint32_t* __restrict__ csrc, uint32_t* __restrict__ parent,
                          uint32_t* __restrict__ msf_a, uint32_t* __restrict__ msf_b,
                          unsigned long long* __restrict__ msf_w, uint32_t* __restrict__ count) {
  uint32_t c = gid();
  bool hook = false;
  uint32_t e = 0, t = 0;
  unsigned long long w = 0;
  if (c < n) {
    parent[c] = c;
    w = cbest[c];
    if (comp[c] == c && w != ~0ull) {
      e = cedge[c];
      t = comp[col[e]];
      hook = !(cbest[t] == w && c < t);
    }
  }
  if (hook) parent[c] = t;
  // forest slots: one counter atomic per wavefront (the order of the slots
  // is immaterial: the forest's edges are sorted by length afterwards)
  const unsigned long long m = __ballot(hook);
  if (m == 0) return;
  const uint32_t lane = threadIdx.x & 63u, leader = (uint32_t)__builtin_ctzll(m);
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(count, (uint32_t)__builtin_popcountll(m));
  base = __builtin_amdgcn_readlane(base, leader);
  if (hook) {
    const uint32_t slot = base + (uint32_t)__builtin_popcountll(m & ((1ull << lane) - 1ull));
    msf_a[slot] = csrc[c];
    msf_b[slot] = col[e];
    msf_w[slot] = w;
  }
}

// One in-place doubling round on the hook pointers; a round after one that
// moved nothing returns at once.  A pointer read while its owner rewrites it
// is its old or its new value (aligned 32-bit accesses are single-copy
// atomic), both ancestors on the same hook path.
__global__ void k_bv_jump(uint32_t n, uint32_t* parent, uint32_t* __restrict__ flag_set,
                          const uint32_t* __restrict__ prev_flag) {
  if (prev_flag && *prev_flag == 0) return;
  uint32_t c = gid();
  bool moved = false;
  if (c < n) {
    uint32_t p = __atomic_load_n(parent + c, __ATOMIC_RELAXED);
    uint32_t pp = __atomic_load_n(parent + p, __ATOMIC_RELAXED);
    if (pp != p) {
      __atomic_store_n(parent + c, pp, __ATOMIC_RELAXED);
      moved = true;
    }
  }
  if (__any(moved) && (threadIdx.x & 63u) == 0) *flag_set = 1;
}

__global__ void k_bv_relabel(uint32_t n, const uint32_t* __restrict__ parent, uint32_t* __restrict__ comp,
                             unsigned long long* __restrict__ cbest) {
  uint32_t v = gid();
  if (v >= n) return;
  comp[v] = parent[comp[v]];
  cbest[v] = ~0ull;
}

__global__ void k_iota(uint32_t m, uint32_t* __restrict__ out) {
  uint32_t i = gid();
  if (i < m) out[i] = i;
}

__global__ void k_gather_pairs(uint32_t m, const uint32_t* __restrict__ idx, const uint32_t* __restrict__ a,
                               const uint32_t* __restrict__ b, uint32_t* __restrict__ a2, uint32_t* __restrict__ b2) {
  uint32_t i = gid();
  if (i >= m) return;
  a2[i] = a[idx[i]];
  b2[i] = b[idx[i]];
}

// ------------------------------------------------- ranks from the forest ----

// Components holding an outlet are reached by the traversal.
__global__ void k_mark_reached(uint32_t n, const uint8_t* __restrict__ is_outlet, const uint32_t* __restrict__ comp,
                               uint8_t* __restrict__ reached) {
  uint32_t v = gid();
  if (v < n && is_outlet[v]) reached[comp[v]] = 1;
}

// The two arcs of forest edge j (in length order) keyed by their source's
// label (an outlet's label is its terrain's S); edges of unreached
// components get the key FL_NONE and sort last.
__global__ void k_arcs(uint32_t m, const uint32_t* __restrict__ a2, const uint32_t* __restrict__ b2,
                       const uint32_t* __restrict__ icomp, const uint32_t* __restrict__ comp,
                       const uint8_t* __restrict__ reached, uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
  uint32_t j = gid();
  if (j >= m) return;
  const uint32_t a = a2[j], b = b2[j];
  const bool ok = reached[comp[a]] != 0;
  key[2 * j] = ok ? icomp[a] : FL_NONE;
  key[2 * j + 1] = ok ? icomp[b] : FL_NONE;
  val[2 * j] = 2 * j;
  val[2 * j + 1] = 2 * j + 1;
}

// Position of every arc in source order, and the first position of every
// source's list.
__global__ void k_arc_pos(uint32_t M, const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sarc,
                          uint32_t* __restrict__ pos, uint32_t* __restrict__ start) {
  uint32_t p = gid();
  if (p >= M) return;
  pos[sarc[p]] = p;
  const uint32_t u = skey[p];
  if (u != FL_NONE && (p == 0 || skey[p - 1] != u)) start[u] = p;
}

// Euler-tour successor of the arc at position p (u->v): the arc after v->u
// in v's list, cyclically; the tour is cut where it would start S's list
// again.  Element p = (successor position or FL_NONE, 1).
__global__ void k_arc_tour(uint32_t M, const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sarc,
                           const uint32_t* __restrict__ pos, const uint32_t* __restrict__ start,
                           const uint8_t* __restrict__ is_outlet, unsigned long long* __restrict__ tour) {
  uint32_t p = gid();
  if (p >= M) return;
  if (skey[p] == FL_NONE) {
    tour[p] = pr(FL_NONE, 1u);
    return;
  }
  const uint32_t t = pos[sarc[p] ^ 1u];
  const uint32_t v = skey[t];
  const bool last = !(t + 1 < M && skey[t + 1] == v);
  uint32_t nx = last ? start[v] : t + 1;
  if (last && is_outlet[v]) nx = FL_NONE;  // back at S after its last arc
  tour[p] = pr(nx, 1u);
}

// One in-place doubling round of a list ranking over elements [0, M) (or the
// elements of the children list, see k_qtour_jump): (p, c) <- (p', c + c') with
// (p', c') the pair of p; FL_NONE ends a list.  Pairs read while their owner
// rewrites them are old or new values of the same list, both valid (aligned
// 64-bit accesses are single-copy atomic).
__global__ void k_list_jump(uint32_t M, unsigned long long* list, uint32_t* __restrict__ flag_set,
                            const uint32_t* __restrict__ prev_flag) {
  if (prev_flag && *prev_flag == 0) return;
  uint32_t p = gid();
  bool moved = false;
  if (p < M) {
    unsigned long long x = __atomic_load_n(list + p, __ATOMIC_RELAXED);
    uint32_t nx = pr_ptr(x);
    if (nx != FL_NONE) {
      // two hops: (p'', c + c' + c''), the reach triples per round
      unsigned long long y = __atomic_load_n(list + nx, __ATOMIC_RELAXED);
      const uint32_t ny = pr_ptr(y);
      if (ny != FL_NONE) {
        unsigned long long z = __atomic_load_n(list + ny, __ATOMIC_RELAXED);
        __atomic_store_n(list + p, pr(pr_ptr(z), pr_cnt(x) + pr_cnt(y) + pr_cnt(z)), __ATOMIC_RELAXED);
      } else {
        __atomic_store_n(list + p, pr(FL_NONE, pr_cnt(x) + pr_cnt(y)), __ATOMIC_RELAXED);
      }
      moved = true;
    }
  }
  if (__any(moved) && (threadIdx.x & 63u) == 0) *flag_set = 1;
}

// ------------------------------------------------ subtree sizes ----
//
// Sizes of the subtrees of a forest given by its parent pointers (next; a
// root points at itself), from an Euler tour ranked by list ranking: arc 2v
// enters non-root site v from its parent, arc 2v + 1 leaves it, and the
// 2 (size(v) - 1) arcs between them are those of v's subtree.  The sites of
// a terrain that is not active (tid, t_active) count as roots: their
// pointers may be stale.
__device__ __forceinline__ bool st_root(uint32_t v, uint32_t p, const uint32_t* tid, const uint8_t* t_active) {
  return p == v || (tid && !t_active[tid[v]]);
}

// Children lists built with exchanges: a site's children come in no
// particular order, which sizes do not depend on.
__global__ void k_st_link(uint32_t n, const uint32_t* __restrict__ next, const uint32_t* __restrict__ tid,
                          const uint8_t* __restrict__ t_active, uint32_t* __restrict__ head,
                          uint32_t* __restrict__ sib) {
  uint32_t v = gid();
  if (v >= n) return;
  const uint32_t p = next[v];
  sib[v] = st_root(v, p, tid, t_active) ? FL_NONE : atomicExch(head + p, v);
}

// (successor, 1) per arc: into v, then v's first child or back out of v;
// out of v, then v's next sibling or out of v's parent (the tour's end when
// the parent is a root: (FL_NONE, 1)).  A root's two slots are (FL_NONE, 0).
__global__ void k_st_tour(uint32_t n, const uint32_t* __restrict__ next, const uint32_t* __restrict__ tid,
                          const uint8_t* __restrict__ t_active, const uint32_t* __restrict__ head,
                          const uint32_t* __restrict__ sib, unsigned long long* __restrict__ tour) {
  uint32_t v = gid();
  if (v >= n) return;
  const uint32_t p = next[v];
  if (st_root(v, p, tid, t_active)) {
    tour[2ull * v] = pr(FL_NONE, 0);
    tour[2ull * v + 1] = pr(FL_NONE, 0);
    return;
  }
  const uint32_t c = head[v], sb = sib[v];
  tour[2ull * v] = pr(c != FL_NONE ? 2u * c : 2u * v + 1u, 1);
  tour[2ull * v + 1] = pr(sb != FL_NONE ? 2u * sb : (st_root(p, next[p], tid, t_active) ? FL_NONE : 2u * p + 1u), 1);
}

// size(v) = (arcs to the end from the arc into v - those from the arc out of
// v + 1) / 2; a root gets n.
__global__ void k_st_size(uint32_t n, const uint32_t* __restrict__ next, const uint32_t* __restrict__ tid,
                          const uint8_t* __restrict__ t_active, const unsigned long long* __restrict__ tour,
                          double* __restrict__ size) {
  uint32_t v = gid();
  if (v >= n) return;
  if (st_root(v, next[v], tid, t_active)) {
    size[v] = (double)n;
    return;
  }
  const uint32_t a = pr_cnt(tour[2ull * v]), b = pr_cnt(tour[2ull * v + 1]);
  size[v] = (double)((a - b + 1u) / 2u);
}

// Parent and weight of every forest site: of an edge's two arcs the one
// earlier in the tour (more arcs after it) points away from S.  The weight
// is the edge's position in length order; S weighs FL_NONE (infinite).
__global__ void k_edge_dir(uint32_t m, const uint32_t* __restrict__ a2, const uint32_t* __restrict__ b2,
                           const uint32_t* __restrict__ icomp, const uint32_t* __restrict__ key,
                           const uint32_t* __restrict__ pos, const unsigned long long* __restrict__ tour,
                           uint32_t* __restrict__ child, uint32_t* __restrict__ par, uint32_t* __restrict__ wt) {
  uint32_t j = gid();
  if (j >= m) return;
  if (key[2 * j] == FL_NONE) {
    child[j] = FL_NONE;
    return;
  }
  const uint32_t da = pr_cnt(tour[pos[2 * j]]), db = pr_cnt(tour[pos[2 * j + 1]]);
  const uint32_t c = da > db ? b2[j] : a2[j];
  const uint32_t p = da > db ? icomp[a2[j]] : icomp[b2[j]];
  child[j] = c;
  par[c] = p;
  wt[c] = j;
}

// q(c) <- q(q(c)) while q(c) is lighter than c, in place (a value read while
// being rewritten is an older or newer pointer of the same site, both valid).
__global__ void k_q_jump(uint32_t m, const uint32_t* __restrict__ child, const uint32_t* __restrict__ wt,
                         uint32_t* q, uint32_t* __restrict__ flag_set, const uint32_t* __restrict__ prev_flag) {
  if (prev_flag && *prev_flag == 0) return;
  uint32_t j = gid();
  bool moved = false;
  if (j < m) {
    const uint32_t c = child[j];
    if (c != FL_NONE) {
      const uint32_t qc = __atomic_load_n(q + c, __ATOMIC_RELAXED);
      if (wt[qc] < j) {
        // up to two steps a round: past qc, and past its q when that is
        // lighter too
        const uint32_t q2 = __atomic_load_n(q + qc, __ATOMIC_RELAXED);
        const uint32_t q3 = wt[q2] < j ? __atomic_load_n(q + q2, __ATOMIC_RELAXED) : q2;
        __atomic_store_n(q + c, q3, __ATOMIC_RELAXED);
        moved = true;
      }
    }
  }
  if (__any(moved) && (threadIdx.x & 63u) == 0) *flag_set = 1;
}

// The q-tree's children, keyed by their q, in increasing weight (edge order).
__global__ void k_q_keys(uint32_t m, const uint32_t* __restrict__ child, const uint32_t* __restrict__ q,
                         uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
  uint32_t j = gid();
  if (j >= m) return;
  const uint32_t c = child[j];
  key[j] = c == FL_NONE ? FL_NONE : q[c];
  val[j] = c;
}

// First child of every q-tree site and next sibling of every child.
__global__ void k_q_links(uint32_t m, const uint32_t* __restrict__ qs, const uint32_t* __restrict__ cs,
                          uint32_t* __restrict__ fch, uint32_t* __restrict__ nsib) {
  uint32_t p = gid();
  if (p >= m) return;
  const uint32_t u = qs[p];
  if (u == FL_NONE) return;
  const uint32_t c = cs[p];
  if (p == 0 || qs[p - 1] != u) fch[u] = c;
  nsib[c] = (p + 1 < m && qs[p + 1] == u) ? cs[p + 1] : FL_NONE;
}

// Euler tour of the q-tree, elements 2c (into c) and 2c + 1 (back out of c)
// for every child c, stored backwards: element e = (predecessor or FL_NONE,
// 1 for an element into a site, else 0).  Each element is written by its
// predecessor's site (exactly one), the first element of each tour by its
// S (an outlet label with children).
__global__ void k_q_tour(uint32_t m, uint32_t n, const uint32_t* __restrict__ qs, const uint32_t* __restrict__ cs,
                         const uint32_t* __restrict__ q, const uint32_t* __restrict__ fch,
                         const uint32_t* __restrict__ nsib, const uint8_t* __restrict__ is_outlet,
                         unsigned long long* __restrict__ qt) {
  uint32_t p = gid();
  if (p < m && qs[p] != FL_NONE) {
    const uint32_t c = cs[p];
    const uint32_t f = fch[c];
    const uint32_t s1 = f != FL_NONE ? 2 * f : 2 * c + 1;  // after entering c
    qt[s1] = pr(2 * c, (s1 & 1u) ? 0u : 1u);
    const uint32_t sb = nsib[c], qc = q[c];
    const uint32_t s2 = sb != FL_NONE ? 2 * sb : (is_outlet[qc] ? FL_NONE : 2 * qc + 1);  // after leaving c
    if (s2 != FL_NONE) qt[s2] = pr(2 * c + 1, (s2 & 1u) ? 0u : 1u);
  }
  // tour starts: the first child of each S
  if (p < n && is_outlet[p]) {
    const uint32_t f = fch[p];
    if (f != FL_NONE) qt[2 * f] = pr(FL_NONE, 1u);
  }
}

// As k_list_jump over the tour elements of the children (element 2c + b of
// child cs[p / 2], b = p & 1).
__global__ void k_qtour_jump(uint32_t M, const uint32_t* __restrict__ qs, const uint32_t* __restrict__ cs,
                             unsigned long long* qt, uint32_t* __restrict__ flag_set,
                             const uint32_t* __restrict__ prev_flag) {
  if (prev_flag && *prev_flag == 0) return;
  uint32_t p = gid();
  bool moved = false;
  if (p < M && qs[p >> 1] != FL_NONE) {
    const uint32_t e = 2 * cs[p >> 1] + (p & 1u);
    unsigned long long x = __atomic_load_n(qt + e, __ATOMIC_RELAXED);
    uint32_t nx = pr_ptr(x);
    if (nx != FL_NONE) {
      // two hops, as k_list_jump
      unsigned long long y = __atomic_load_n(qt + nx, __ATOMIC_RELAXED);
      const uint32_t ny = pr_ptr(y);
      if (ny != FL_NONE) {
        unsigned long long z = __atomic_load_n(qt + ny, __ATOMIC_RELAXED);
        __atomic_store_n(qt + e, pr(pr_ptr(z), pr_cnt(x) + pr_cnt(y) + pr_cnt(z)), __ATOMIC_RELAXED);
      } else {
        __atomic_store_n(qt + e, pr(FL_NONE, pr_cnt(x) + pr_cnt(y)), __ATOMIC_RELAXED);
      }
      moved = true;
    }
  }
  if (__any(moved) && (threadIdx.x & 63u) == 0) *flag_set = 1;
}

// rank(c) = its terrain's first non-outlet rank + its preorder position
// (down steps from the tour's start up to c's, minus one).
__global__ void k_q_rank(uint32_t m, const uint32_t* __restrict__ qs, const uint32_t* __restrict__ cs,
                         const unsigned long long* __restrict__ qt, const uint32_t* __restrict__ tid,
                         const uint32_t* __restrict__ base, uint32_t* __restrict__ rank) {
  uint32_t p = gid();
  if (p >= m || qs[p] == FL_NONE) return;
  const uint32_t c = cs[p];
  rank[c] = base[tid ? tid[c] : 0u] + pr_cnt(qt[2 * c]) - 1u;
}

__global__ void k_scatter32(uint32_t count, const uint32_t* __restrict__ idx, const uint32_t* __restrict__ val,
                            uint32_t* __restrict__ dst) {
  uint32_t q = gid();
  if (q < count) dst[idx[q]] = val[q];
}

// The visited sites by rank: order[rank[v]] = v (ranks are distinct).
__global__ void k_order_of_rank(uint32_t n, const uint32_t* __restrict__ rank, uint32_t* __restrict__ order) {
  uint32_t v = gid();
  if (v < n && rank[v] != FL_NONE) order[rank[v]] = v;
}

inline dim3 grid_for(uint64_t n) { return dim3((uint32_t)((n + 255) / 256)); }

// Scratch of one visit-order build from an arena (the device pool:
// stream-ordered, no device-wide synchronisation); released when the build
// returns (pool buffers freed, region buffers rewound).
// Scratch of one call (fl_scratch) with byte-sized allocations.
struct DevBuf : fl_scratch {
  using fl_scratch::fl_scratch;
  hipError_t alloc(void** p, uint64_t bytes) { return fl_scratch::alloc((uint8_t**)p, bytes); }
};

#define HIPTRY(call)                                               \
  do {                                                             \
    hipError_t e_ = (call);                                        \
    if (e_ != hipSuccess) {                                        \
      *err = std::string(#call) + ": " + hipGetErrorString(e_);    \
      return e_ == hipErrorOutOfMemory ? FL_ERR_OUT_OF_MEMORY : FL_ERR_HIP; \
    }                                                              \
  } while (0)

// Hash tables of the length check: terrain g's at toff[g], 2^b >= its
// undirected edge count entries (entry_off: G + 1 union entry offsets).
uint64_t len_tables(uint32_t G, const uint32_t* entry_off, std::vector<uint32_t>& toff, std::vector<uint32_t>& tmask) {
  toff.assign(G, 0);
  tmask.assign(G, 0);
  uint64_t total = 0;
  for (uint32_t g = 0; g < G; g++) {
    uint64_t edges = (entry_off[g + 1] - entry_off[g]) / 2 + 1, cap = 1;
    while (cap < 2 * edges) cap <<= 1;
    toff[g] = (uint32_t)total;
    tmask[g] = (uint32_t)(cap - 1);
    total += cap;
  }
  return total;
}

// Rounds of pointer jumping per host check.
constexpr uint32_t kJumpBatch = 8;
constexpr uint32_t kMaxFlags = 256;

// The forest and rank buffers of fl_device_ranks (n sites): allocated from
// `buf`, also in a COUNT arena for sizing.
struct RankBufs {
  uint32_t *comp, *parent, *vbest, *cedge, *csrc, *ma, *mb, *icomp, *idx, *idx2, *a2, *b2, *flags;
  unsigned long long *cbest, *mw, *mw2;
  uint8_t* reached;
  uint32_t *key, *val, *skey, *sarc, *pos, *start;
  unsigned long long* tour;
  uint32_t *child, *par, *wt, *q, *qk, *qv, *qs, *cs, *fch, *nsib;
  unsigned long long* qt;
  void *tmp1, *tmp2;
  size_t tmp1_bytes = 0, tmp2_bytes = 0;
};

hipError_t rank_bufs(DevBuf& buf, uint32_t n, RankBufs& r) {
  hipError_t e;
  const uint64_t N4 = (uint64_t)n * 4, N8 = (uint64_t)n * 8;
  for (uint32_t** p : {&r.comp, &r.parent, &r.vbest, &r.cedge, &r.csrc, &r.ma, &r.mb, &r.icomp, &r.idx, &r.idx2,
                       &r.a2, &r.b2, &r.start, &r.child, &r.par, &r.wt, &r.q, &r.qk, &r.qv, &r.qs, &r.cs, &r.fch,
                       &r.nsib})
    if ((e = buf.alloc((void**)p, N4)) != hipSuccess) return e;
  for (uint32_t** p : {&r.key, &r.val, &r.skey, &r.sarc, &r.pos})
    if ((e = buf.alloc((void**)p, 2 * N4)) != hipSuccess) return e;
  for (unsigned long long** p : {&r.cbest, &r.mw, &r.mw2})
    if ((e = buf.alloc((void**)p, N8)) != hipSuccess) return e;
  for (unsigned long long** p : {&r.tour, &r.qt})
    if ((e = buf.alloc((void**)p, 2 * N8)) != hipSuccess) return e;
  if ((e = buf.alloc((void**)&r.reached, (uint64_t)n + 16)) != hipSuccess) return e;
  if ((e = buf.alloc((void**)&r.flags, kMaxFlags * 4)) != hipSuccess) return e;
  // radix sort scratch: edges by length (64-bit keys), arcs by source and
  // children by q (32-bit keys, up to 2n items)
  const int items = n > 0 ? (int)n : 1;
  if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, r.tmp1_bytes, r.mw, r.mw2, r.idx, r.idx2, items, 0, 64,
                                              buf.s)) != hipSuccess)
    return e;
  if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, r.tmp2_bytes, r.key, r.skey, r.val, r.sarc, 2 * items, 0,
                                              32, buf.s)) != hipSuccess)
    return e;
  if ((e = buf.alloc(&r.tmp1, r.tmp1_bytes)) != hipSuccess) return e;
  if ((e = buf.alloc(&r.tmp2, r.tmp2_bytes)) != hipSuccess) return e;
  return hipSuccess;
}

// Bytes of rank_bufs for n sites plus the terrain bases (G entries).
uint64_t rank_scratch_bytes(uint32_t n, uint64_t G) {
  fl_arena a;
  a.kind = fl_arena::COUNT;
  {
    DevBuf buf(nullptr, &a);
    RankBufs r;
    rank_bufs(buf, n, r);
    void* p;
    buf.alloc(&p, 4 * G);
  }
  return a.peak;
}

// Pointer-jumping rounds launched `batch` at a time, each flagging movement
// in its own word of r.flags (cleared beforehand); returns after the first
// batch whose last round moved nothing, or an error past kMaxFlags rounds.
template <class Launch>
int jump_to_fixpoint(hipStream_t s, uint32_t* d_flags, uint32_t nflags, uint32_t first_batch, Launch launch,
                     std::string* err, const char* what) {
  uint32_t h[kMaxFlags];
  uint32_t round = 0, target = first_batch;
  for (;;) {
    if (target > nflags) {
      *err = std::string(what) + " did not converge";
      return FL_ERR_HIP;
    }
    for (; round < target; round++) launch(d_flags + round, round ? d_flags + round - 1 : nullptr);
    HIPTRY(hipMemcpyAsync(h, d_flags, round * 4, hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    if (h[round - 1] == 0) return FL_OK;
    target = round + kJumpBatch;
  }
}

}  // namespace

uint64_t fl_msf_scratch_bytes(uint32_t n, uint32_t nnz) {
  // the length check's tables (terrain g's holds fewer than 2 * entries_g + 4
  // slots, and there are at most n terrains) and flags, then the rank buffers
  fl_arena a;
  a.kind = fl_arena::COUNT;
  uint64_t check;
  {
    DevBuf buf(nullptr, &a);
    void* p;
    buf.alloc(&p, 8ull * (2ull * nnz + 4ull * n + 4));
    buf.alloc(&p, 3ull * 4 * (n + 1));
    check = a.peak;
  }
  a.used = a.peak = 0;
  {
    DevBuf buf(nullptr, &a);
    RankBufs r;
    rank_bufs(buf, n, r);
    void* p;
    buf.alloc(&p, 4ull * (n + 1));  // terrain bases
  }
  uint64_t ranks = a.peak, scatter = fl_round256(8ull * (n + 1));
  uint64_t peak = check > ranks ? check : ranks;
  return peak > scatter ? peak : scatter;
}

int fl_device_len_check(hipStream_t s, uint32_t n, const uint32_t* rp, const uint32_t* col, const double* len,
                        const uint32_t* tid, uint32_t G, const uint32_t* entry_off, std::vector<uint8_t>& bad,
                        std::string* err, fl_arena* arena) {
  bad.assign(G, 0);
  if (n == 0) return FL_OK;
  DevBuf buf(s, arena);
  std::vector<uint32_t> toff, tmask;
  uint64_t entries = len_tables(G, entry_off, toff, tmask);
  unsigned long long* table = nullptr;
  uint32_t *d_meta = nullptr, *d_bad = nullptr;
  HIPTRY(buf.reserve(fl_round256(entries * 8) + fl_round256(3ull * 4 * G)));
  HIPTRY(buf.alloc((void**)&table, entries * 8));
  HIPTRY(buf.alloc((void**)&d_meta, 3ull * 4 * G));
  uint32_t *d_toff = d_meta, *d_mask = d_meta + G;
  d_bad = d_meta + 2 * G;
  std::vector<uint32_t> meta(3 * (size_t)G, 0);
  std::copy(toff.begin(), toff.end(), meta.begin());
  std::copy(tmask.begin(), tmask.end(), meta.begin() + G);
  HIPTRY(hipMemcpyAsync(d_meta, meta.data(), meta.size() * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_fill64, grid_for(entries), dim3(256), 0, s, (uint32_t)entries, table, 0ull);
  hipLaunchKernelGGL(k_len_check, grid_for(n), dim3(256), 0, s, n, rp, col, len, tid, d_toff, d_mask, table, d_bad);
  HIPTRY(hipMemcpyAsync(meta.data() + 2 * G, d_bad, (size_t)G * 4, hipMemcpyDeviceToHost, s));
  HIPTRY(hipStreamSynchronize(s));
  for (uint32_t g = 0; g < G; g++) bad[g] = meta[2 * G + g] != 0;
  return FL_OK;
}

int fl_device_ranks(hipStream_t s, uint32_t n, const uint32_t* rp, const uint32_t* col, const double* len,
                    const uint8_t* is_outlet, const uint32_t* tid, const std::vector<uint32_t>& init_comp,
                    const std::vector<uint32_t>& base, uint32_t* rank, fl_visit_stats* st, std::string* err,
                    fl_arena* arena) {
  if (n == 0) return FL_OK;
  auto t0 = std::chrono::steady_clock::now();
  DevBuf buf(s, arena);
  HIPTRY(buf.reserve(rank_scratch_bytes(n, base.size())));
  RankBufs r;
  HIPTRY(rank_bufs(buf, n, r));
  uint32_t* d_base = nullptr;
  HIPTRY(buf.alloc((void**)&d_base, base.size() * 4));
  const dim3 B(256), Gn = grid_for(n);
  HIPTRY(hipMemcpyAsync(d_base, base.data(), base.size() * 4, hipMemcpyHostToDevice, s));
  HIPTRY(hipMemcpyAsync(r.icomp, init_comp.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_fill32, dim3(1), B, 0, s, kMaxFlags, r.flags, 0u);
  hipLaunchKernelGGL(k_bv_init, Gn, B, 0, s, n, r.icomp, r.comp, r.cbest);
  // Boruvka: one host check per round
  uint32_t m = 0, h_count = 0;
  for (uint32_t round = 0;; ++round) {
    if (round > 64) {
      *err = "minimum spanning forest did not converge";
      return FL_ERR_HIP;
    }
    hipLaunchKernelGGL(k_fill32, dim3(1), B, 0, s, kMaxFlags - 1, r.flags + 1, 0u);
    hipLaunchKernelGGL(k_bv_best, Gn, B, 0, s, n, rp, col, len, r.comp, r.cbest, r.vbest);
    hipLaunchKernelGGL(k_bv_pick, Gn, B, 0, s, n, len, r.comp, r.cbest, r.vbest, r.cedge, r.csrc);
    hipLaunchKernelGGL(k_bv_hook, Gn, B, 0, s, n, col, r.comp, r.cbest, r.cedge, r.csrc, r.parent, r.ma, r.mb, r.mw,
                       r.flags);
    int rc = jump_to_fixpoint(
        s, r.flags + 1, kMaxFlags - 1, 4,
        [&](uint32_t* fset, const uint32_t* prev) {
          hipLaunchKernelGGL(k_bv_jump, Gn, B, 0, s, n, r.parent, fset, prev);
        },
        err, "forest hooking");
    if (rc != FL_OK) return rc;
    hipLaunchKernelGGL(k_bv_relabel, Gn, B, 0, s, n, r.parent, r.comp, r.cbest);
    HIPTRY(hipMemcpyAsync(&h_count, r.flags, 4, hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    if (h_count == m) break;  // no component hooked this round
    m = h_count;
    if (m >= n) {
      *err = "forest edge count out of range";
      return FL_ERR_HIP;
    }
  }
  auto t1 = std::chrono::steady_clock::now();
  if (st) {
    st->forest_edges = m;
    st->forest_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  }
  if (m == 0) return FL_OK;
  const dim3 Gm = grid_for(m), G2m = grid_for(2ull * m);
  // edges by length
  hipLaunchKernelGGL(k_iota, Gm, B, 0, s, m, r.idx);
  HIPTRY(hipcub::DeviceRadixSort::SortPairs(r.tmp1, r.tmp1_bytes, r.mw, r.mw2, r.idx, r.idx2, (int)m, 0, 64, s));
  hipLaunchKernelGGL(k_gather_pairs, Gm, B, 0, s, m, r.idx2, r.ma, r.mb, r.a2, r.b2);
  // 1. rooting at S
  hipLaunchKernelGGL(k_fill32, grid_for((n + 3) / 4), B, 0, s, (n + 3) / 4, (uint32_t*)r.reached, 0u);
  hipLaunchKernelGGL(k_mark_reached, Gn, B, 0, s, n, is_outlet, r.comp, r.reached);
  hipLaunchKernelGGL(k_arcs, Gm, B, 0, s, m, r.a2, r.b2, r.icomp, r.comp, r.reached, r.key, r.val);
  HIPTRY(hipcub::DeviceRadixSort::SortPairs(r.tmp2, r.tmp2_bytes, r.key, r.skey, r.val, r.sarc, (int)(2 * m), 0, 32,
                                            s));
  hipLaunchKernelGGL(k_arc_pos, G2m, B, 0, s, 2 * m, r.skey, r.sarc, r.pos, r.start);
  hipLaunchKernelGGL(k_arc_tour, G2m, B, 0, s, 2 * m, r.skey, r.sarc, r.pos, r.start, is_outlet, r.tour);
  // a list of L elements is ranked after ceil(log3 L) rounds (each round at
  // least triples every pointer's reach); one more proves it
  uint32_t lg = 1;
  for (uint64_t reach = 3; reach < 2ull * m; reach *= 3) lg++;
  hipLaunchKernelGGL(k_fill32, dim3(1), B, 0, s, kMaxFlags, r.flags, 0u);
  int rc = jump_to_fixpoint(
      s, r.flags, kMaxFlags, lg + 1,
      [&](uint32_t* fset, const uint32_t* prev) {
        hipLaunchKernelGGL(k_list_jump, G2m, B, 0, s, 2 * m, r.tour, fset, prev);
      },
      err, "forest tour ranking");
  if (rc != FL_OK) return rc;
  // S (every outlet label) is heavier than any edge
  hipLaunchKernelGGL(k_fill32, Gn, B, 0, s, n, r.wt, FL_NONE);
  hipLaunchKernelGGL(k_edge_dir, Gm, B, 0, s, m, r.a2, r.b2, r.icomp, r.key, r.pos, r.tour, r.child, r.par, r.wt);
  HIPTRY(hipMemcpyAsync(r.q, r.par, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
  // 2. q: the nearest heavier ancestor
  hipLaunchKernelGGL(k_fill32, dim3(1), B, 0, s, kMaxFlags, r.flags, 0u);
  rc = jump_to_fixpoint(
      s, r.flags, kMaxFlags, kJumpBatch,
      [&](uint32_t* fset, const uint32_t* prev) {
        hipLaunchKernelGGL(k_q_jump, Gm, B, 0, s, m, r.child, r.wt, r.q, fset, prev);
      },
      err, "heavier-ancestor jumping");
  if (rc != FL_OK) return rc;
  // 3. children in increasing weight under each q
  hipLaunchKernelGGL(k_q_keys, Gm, B, 0, s, m, r.child, r.q, r.qk, r.qv);
  HIPTRY(hipcub::DeviceRadixSort::SortPairs(r.tmp2, r.tmp2_bytes, r.qk, r.qs, r.qv, r.cs, (int)m, 0, 32, s));
  hipLaunchKernelGGL(k_fill32, Gn, B, 0, s, n, r.fch, FL_NONE);
  hipLaunchKernelGGL(k_q_links, Gm, B, 0, s, m, r.qs, r.cs, r.fch, r.nsib);
  // 4. preorder: the q-tree's tour ranked from its start (qt is indexed by
  // 2 * site + {0, 1}: 2n entries)
  hipLaunchKernelGGL(k_q_tour, grid_for(m > n ? m : n), B, 0, s, m, n, r.qs, r.cs, r.q, r.fch, r.nsib, is_outlet,
                     r.qt);
  hipLaunchKernelGGL(k_fill32, dim3(1), B, 0, s, kMaxFlags, r.flags, 0u);
  rc = jump_to_fixpoint(
      s, r.flags, kMaxFlags, lg + 1,
      [&](uint32_t* fset, const uint32_t* prev) {
        hipLaunchKernelGGL(k_qtour_jump, G2m, B, 0, s, 2 * m, r.qs, r.cs, r.qt, fset, prev);
      },
      err, "preorder ranking");
  if (rc != FL_OK) return rc;
  hipLaunchKernelGGL(k_q_rank, Gm, B, 0, s, m, r.qs, r.cs, r.qt, tid, d_base, rank);
  // the scratch is released (or rewound) on return: nothing may still read it
  HIPTRY(hipStreamSynchronize(s));
  if (st) st->ranks_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
  return FL_OK;
}

int fl_device_scatter_u32(hipStream_t s, uint32_t count, const uint32_t* idx, const uint32_t* val, uint32_t* dst,
                          std::string* err, fl_arena* arena) {
  if (count == 0) return FL_OK;
  DevBuf buf(s, arena);
  uint32_t* d = nullptr;
  HIPTRY(buf.alloc((void**)&d, 8ull * count));
  HIPTRY(hipMemcpyAsync(d, idx, (size_t)count * 4, hipMemcpyHostToDevice, s));
  HIPTRY(hipMemcpyAsync(d + count, val, (size_t)count * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_scatter32, grid_for(count), dim3(256), 0, s, count, d, d + count, dst);
  HIPTRY(hipStreamSynchronize(s));
  return FL_OK;
}

void fl_device_order(hipStream_t s, uint32_t n, const uint32_t* rank, uint32_t* order) {
  if (n) hipLaunchKernelGGL(k_order_of_rank, grid_for(n), dim3(256), 0, s, n, rank, order);
}

uint64_t fl_subtree_scratch_bytes(uint32_t n) {
  fl_arena a;
  a.kind = fl_arena::COUNT;
  {
    DevBuf buf(nullptr, &a);
    void* p;
    buf.alloc(&p, 4ull * n);
    buf.alloc(&p, 4ull * n);
    buf.alloc(&p, 16ull * n);
    buf.alloc(&p, 4ull * kMaxFlags);
  }
  return a.peak;
}

int fl_device_subtree_sizes(hipStream_t s, uint32_t n, const uint32_t* next, const uint32_t* tid,
                            const uint8_t* t_active, double* size, std::string* err, fl_arena* arena) {
  if (n == 0 || n >= (1u << 31)) return FL_OK;  // arc indices must fit 32 bits
  DevBuf buf(s, arena);
  HIPTRY(buf.reserve(fl_subtree_scratch_bytes(n)));
  uint32_t *head = nullptr, *sib = nullptr, *flags = nullptr;
  unsigned long long* tour = nullptr;
  HIPTRY(buf.alloc((void**)&head, 4ull * n));
  HIPTRY(buf.alloc((void**)&sib, 4ull * n));
  HIPTRY(buf.alloc((void**)&tour, 16ull * n));
  HIPTRY(buf.alloc((void**)&flags, 4ull * kMaxFlags));
  const dim3 B(256), Gn = grid_for(n), G2 = grid_for(2ull * n);
  hipLaunchKernelGGL(k_fill32, Gn, B, 0, s, n, head, FL_NONE);
  hipLaunchKernelGGL(k_fill32, dim3(1), B, 0, s, kMaxFlags, flags, 0u);
  hipLaunchKernelGGL(k_st_link, Gn, B, 0, s, n, next, tid, t_active, head, sib);
  hipLaunchKernelGGL(k_st_tour, Gn, B, 0, s, n, next, tid, t_active, head, sib, tour);
  // a list of L arcs is ranked after ceil(log3 L) rounds (each round at
  // least triples every pointer's reach), rounds after a quiet one return
  // at once; nothing is read back: the sizes are a hint, the stream goes on
  uint32_t lg = 1;
  for (uint64_t reach = 3; reach < 2ull * n; reach *= 3) lg++;
  for (uint32_t r = 0; r < lg && r < kMaxFlags; r++)
    hipLaunchKernelGGL(k_list_jump, G2, B, 0, s, 2u * n, tour, flags + r, r ? flags + r - 1 : nullptr);
  hipLaunchKernelGGL(k_st_size, Gn, B, 0, s, n, next, tid, t_active, tour, size);
  HIPTRY(hipGetLastError());
  return FL_OK;
}

// fl_create.hip -- device part of a terrain's upload: the caller's graph is
// checked, put in a locality order and relabelled on the GPU.
//
// The caller's arrays arrive as given ("raw").  In a union of G terrains,
// terrain g holds union sites [site_off[g], site_off[g+1]) and union entries
// [entry_off[g], entry_off[g+1]); its n_g + 1 row offsets sit at
// rp[site_off[g] + g ...] (terrain-local entry indices) and its neighbour
// indices are terrain-local.
//
// Checks (rows of at most FL_DEVICE_MAX_ROW entries; the host checks terrains
// with longer rows): per terrain, the first entry that is out of range, a
// self-loop or a repeat of an earlier entry of its row, and the first entry
// i->j whose row j holds no i -- each as the smallest entry index, which is
// the one a sequential scan of the rows meets first -- and whether the two
// entries of some edge hold different length bits.
//
// Locality order.  Neighbouring sites should get nearby labels, so that the
// gathers of the solve's kernels touch few cache lines (results do not depend
// on the labels).  A breadth-first order of a million sites is a sequential
// pass of random accesses (tens of milliseconds on a host core), so it runs
// over clusters of sites instead:
//   1. every site points at the neighbour (or itself) with the smallest hash
//      of its label; pointer jumping takes each site to the local minimum at
//      the end of its chain, which names its cluster (about deg + 1 sites);
//   2. clusters are numbered in label order of those minima, and their sites
//      grouped by a stable sort (label order inside a cluster);
//   3. each cluster's entries into other clusters (repeats mostly dropped)
//      go to the host, which orders the clusters breadth-first, starting each
//      component at its lowest-numbered cluster not yet reached;
//   4. sites take their cluster's position, in label order inside it.
// A terrain's clusters hold its own sites only and are numbered after those
// of the terrains before it, so every terrain keeps its site range.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "fl_internal.h"

#define FL_NONE 0xffffffffu

namespace {

// Workgroup b runs on XCD b % 8; each XCD takes one contiguous eighth of the
// index range (a bijection of [0, gridDim.x)), so neighbouring sites share an L2.
__device__ __forceinline__ uint32_t gid() {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u;
  return (x * q + (x < r ? x : r) + (b >> 3)) * blockDim.x + threadIdx.x;
}

dim3 grid_for(uint64_t n) { return dim3((uint32_t)std::max<uint64_t>((n + 255) / 256, 1)); }

// A bijection of 64-bit values (the finaliser of MurmurHash3): distinct
// labels get distinct hashes.
__device__ __forceinline__ unsigned long long mix64(unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

// Site i of the union: its terrain, the terrain's first site and entry, and
// the bounds of its raw row (terrain-local entry indices).
struct SiteRow {
  uint32_t g, s0, e0, ng, kb, ke;
};
__device__ __forceinline__ SiteRow site_row(const fl_raw_union& R, uint32_t i) {
  SiteRow r;
  r.g = R.tid ? R.tid[i] : 0u;
  r.s0 = R.site_off[r.g];
  r.e0 = R.entry_off[r.g];
  r.ng = R.site_off[r.g + 1] - r.s0;
  r.kb = R.rp[i + r.g];
  r.ke = R.rp[i + r.g + 1];
  return r;
}

__global__ void k_fill32(uint64_t n, uint32_t* __restrict__ p, uint32_t v) {
  uint64_t i = gid();
  if (i < n) p[i] = v;
}

// Terrain of every site: the last g with site_off[g] <= i (empty terrains
// share their offset with the next one).
__global__ void k_tid(uint32_t n, uint32_t G, const uint32_t* __restrict__ site_off, uint32_t* __restrict__ tid) {
  uint32_t i = gid();
  if (i >= n) return;
  uint32_t lo = 0, hi = G;  // site_off[lo] <= i < site_off[hi]
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (site_off[mid] <= i)
      lo = mid;
    else
      hi = mid;
  }
  tid[i] = lo;
}

// ----------------------------------------------------------------- checks --

// Out-of-range indices, self-loops and repeats within a row.
__global__ void k_check_rows(fl_raw_union R, uint32_t* __restrict__ first_bad) {
  uint32_t i = gid();
  if (i >= R.n) return;
  const SiteRow r = site_row(R, i);
  const uint32_t li = i - r.s0;
  const uint32_t* c = R.col + r.e0;
  for (uint32_t k = r.kb; k < r.ke; k++) {
    const uint32_t j = c[k];
    bool bad = j >= r.ng || j == li;
    for (uint32_t q = r.kb; q < k && !bad; q++) bad = c[q] == j;
    if (bad) {
      atomicMin(first_bad + r.g, k);
      return;
    }
  }
}

// Entries i->j whose row j holds no i; whether both entries of an edge carry
// the same length bits.  Entries that fail the row checks are skipped (they
// are reported first).
__global__ void k_check_sym(fl_raw_union R, const double* __restrict__ len, uint32_t* __restrict__ first_bad,
                            uint32_t* __restrict__ differs) {
  uint32_t i = gid();
  if (i >= R.n) return;
  const SiteRow r = site_row(R, i);
  const uint32_t li = i - r.s0;
  const uint32_t* c = R.col + r.e0;
  const double* w = len + r.e0;
  bool diff = false;
  for (uint32_t k = r.kb; k < r.ke; k++) {
    const uint32_t j = c[k];
    if (j >= r.ng || j == li) continue;
    const uint32_t jb = R.rp[r.s0 + j + r.g], je = R.rp[r.s0 + j + r.g + 1];
    uint32_t q = jb;
    while (q < je && c[q] != li) q++;
    if (q == je) {
      atomicMin(first_bad + r.g, k);
      return;
    }
    diff |= __double_as_longlong(w[k]) != __double_as_longlong(w[q]);
  }
  if (diff) atomicOr(differs + r.g, 1u);
}

// ----------------------------------------------------------- clustering ---

// p[i]: the neighbour of i, or i, with the smallest hash.
__global__ void k_hook(fl_raw_union R, uint32_t* __restrict__ p) {
  uint32_t i = gid();
  if (i >= R.n) return;
  const SiteRow r = site_row(R, i);
  const uint32_t* c = R.col + r.e0;
  uint32_t best = i;
  unsigned long long hb = mix64(i);
  for (uint32_t k = r.kb; k < r.ke; k++) {
    const uint32_t j = r.s0 + c[k];
    const unsigned long long h = mix64(j);
    if (h < hb) {
      hb = h;
      best = j;
    }
  }
  p[i] = best;
}

// One round of pointer jumping in place (every pointer leads to a smaller
// hash, so the pointers form a forest and each read is an ancestor, old or
// new).  Skipped once the previous round moved nothing.
__global__ void k_hook_jump(uint32_t n, uint32_t* __restrict__ p, uint32_t* __restrict__ moved,
                            const uint32_t* __restrict__ prev) {
  if (prev && *prev == 0) return;
  uint32_t i = gid();
  if (i >= n) return;
  const uint32_t a = p[i], b = p[a];
  if (a != b) {
    p[i] = b;
    *moved = 1;
  }
}

__global__ void k_root_flags(uint32_t n, const uint32_t* __restrict__ p, uint32_t* __restrict__ f) {
  uint32_t i = gid();
  if (i < n) f[i] = p[i] == i ? 1u : 0u;
}

// cid[i] = number of the cluster root p[i] (roots numbered in label order),
// from the inclusive root count y; val[i] = i for the grouping sort.
__global__ void k_cid(uint32_t n, const uint32_t* __restrict__ p, const uint32_t* __restrict__ y,
                      uint32_t* __restrict__ cid, uint32_t* __restrict__ val) {
  uint32_t i = gid();
  if (i >= n) return;
  cid[i] = y[p[i]] - 1u;
  val[i] = i;
}

__global__ void k_cluster_start(uint32_t n, uint32_t C, const uint32_t* __restrict__ skey,
                                uint32_t* __restrict__ cstart) {
  uint32_t x = gid();
  if (x >= n) return;
  if (x == 0 || skey[x] != skey[x - 1]) cstart[skey[x]] = x;
  if (x == 0) cstart[C] = n;
}

// Cluster c's entries into other clusters, members in sorted order and rows
// in entry order, a value equal to one of the last four kept skipped.  The
// counting pass writes cnt[c], the filling pass the same values at ptr[c].
template <bool kFill>
__global__ void k_cluster_nbrs(fl_raw_union R, uint32_t C, const uint32_t* __restrict__ cstart,
                               const uint32_t* __restrict__ ssite, const uint32_t* __restrict__ cid,
                               uint32_t* __restrict__ cnt_or_ptr, uint32_t* __restrict__ out) {
  uint32_t cl = gid();
  if (cl >= C) return;
  uint32_t l0 = FL_NONE, l1 = FL_NONE, l2 = FL_NONE, l3 = FL_NONE;
  uint32_t count = 0;
  uint32_t* dst = kFill ? out + cnt_or_ptr[cl] : nullptr;
  for (uint32_t x = cstart[cl]; x < cstart[cl + 1]; x++) {
    const uint32_t v = ssite[x];
    const SiteRow r = site_row(R, v);
    const uint32_t* c = R.col + r.e0;
    for (uint32_t k = r.kb; k < r.ke; k++) {
      const uint32_t d = cid[r.s0 + c[k]];
      if (d == cl || d == l0 || d == l1 || d == l2 || d == l3) continue;
      l3 = l2;
      l2 = l1;
      l1 = l0;
      l0 = d;
      if (kFill) dst[count] = d;
      count++;
    }
  }
  if (!kFill) cnt_or_ptr[cl] = count;
}

// Sites of the cluster at position j of the order take the labels
// noff[j], noff[j] + 1, ... in sorted (label) order.
__global__ void k_place(uint32_t C, const uint32_t* __restrict__ corder, const uint32_t* __restrict__ noff,
                        const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ ssite,
                        uint32_t* __restrict__ old_of_new, uint32_t* __restrict__ new_of_old) {
  uint32_t j = gid();
  if (j >= C) return;
  const uint32_t cl = corder[j], a = cstart[cl], b = cstart[cl + 1];
  uint32_t u = noff[j];
  for (uint32_t x = a; x < b; x++, u++) {
    const uint32_t v = ssite[x];
    old_of_new[u] = v;
    new_of_old[v] = u;
  }
}

// ------------------------------------------------------------- relabel ----

__global__ void k_new_degree(fl_raw_union R, const uint32_t* __restrict__ old_of_new, uint32_t* __restrict__ deg) {
  uint32_t u = gid();
  if (u >= R.n) return;
  const uint32_t o = old_of_new[u], g = R.tid ? R.tid[o] : 0u;
  deg[u] = R.rp[o + g + 1] - R.rp[o + g];
}

// Row of new site u: the entries of its old row in their order, neighbours
// relabelled (union labels).
__global__ void k_relabel_rows(fl_raw_union R, const double* __restrict__ len_raw,
                               const uint32_t* __restrict__ old_of_new, const uint32_t* __restrict__ new_of_old,
                               const uint32_t* __restrict__ row_ptr, uint32_t* __restrict__ col,
                               double* __restrict__ len) {
  uint32_t u = gid();
  if (u >= R.n) return;
  const SiteRow r = site_row(R, old_of_new[u]);
  const uint32_t* c = R.col + r.e0;
  const double* w = len_raw + r.e0;
  uint32_t e = row_ptr[u];
  for (uint32_t k = r.kb; k < r.ke; k++, e++) {
    col[e] = new_of_old[r.s0 + c[k]];
    len[e] = w[k];
  }
}

struct SiteArrays {
  const double* src[5];
  double* dst[5];
};

// dst[a][u] = src[a][old_of_new[u]]; a null source gives NaN.
__global__ void k_gather_sites(uint32_t n, const uint32_t* __restrict__ old_of_new, SiteArrays A) {
  uint32_t u = gid();
  if (u >= n) return;
  const uint32_t o = old_of_new[u];
#pragma unroll
  for (int a = 0; a < 5; a++) A.dst[a][u] = A.src[a] ? A.src[a][o] : __builtin_nan("");
}

__global__ void k_mark_u8(uint32_t count, const uint32_t* __restrict__ idx, uint8_t* __restrict__ dst) {
  uint32_t q = gid();
  if (q < count) dst[idx[q]] = 1;
}

#define HIPTRY(call)                                                       \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess) {                                                \
      *err = std::string(#call) + ": " + hipGetErrorString(e_);            \
      return e_ == hipErrorOutOfMemory ? FL_ERR_OUT_OF_MEMORY : FL_ERR_HIP; \
    }                                                                      \
  } while (0)

using Scratch = fl_scratch;

constexpr uint32_t kMaxFlags = 64;

// Buffers of the locality order for n sites and nnz entries (C <= n clusters).
struct LocBufs {
  uint32_t *p, *y, *cid, *val, *skey, *ssite, *cstart, *cnt, *cptr, *corder, *noff, *flags, *nbrs;
  void *tmp_scan, *tmp_sort, *tmp_scan2;
  size_t scan_bytes = 0, sort_bytes = 0, scan2_bytes = 0;
};

hipError_t loc_bufs(Scratch& S, uint32_t n, uint32_t nnz, LocBufs& b) {
  hipError_t e;
  for (uint32_t** q : {&b.p, &b.y, &b.cid, &b.val, &b.skey, &b.ssite, &b.corder, &b.noff})
    if ((e = S.alloc(q, n)) != hipSuccess) return e;
  if ((e = S.alloc(&b.cstart, (uint64_t)n + 1)) != hipSuccess) return e;
  if ((e = S.alloc(&b.cnt, (uint64_t)n + 1)) != hipSuccess) return e;
  if ((e = S.alloc(&b.cptr, (uint64_t)n + 1)) != hipSuccess) return e;
  if ((e = S.alloc(&b.flags, kMaxFlags)) != hipSuccess) return e;
  if ((e = S.alloc(&b.nbrs, nnz)) != hipSuccess) return e;
  const int items = n > 0 ? (int)n : 1;
  if ((e = hipcub::DeviceScan::InclusiveSum(nullptr, b.scan_bytes, b.val, b.y, items, S.s)) != hipSuccess) return e;
  if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, b.sort_bytes, b.cid, b.skey, b.val, b.ssite, items, 0, 32,
                                              S.s)) != hipSuccess)
    return e;
  if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, b.scan2_bytes, b.cnt, b.cptr, items + 1, S.s)) != hipSuccess)
    return e;
  if ((e = S.alloc((uint8_t**)&b.tmp_scan, b.scan_bytes)) != hipSuccess) return e;
  if ((e = S.alloc((uint8_t**)&b.tmp_sort, b.sort_bytes)) != hipSuccess) return e;
  if ((e = S.alloc((uint8_t**)&b.tmp_scan2, b.scan2_bytes)) != hipSuccess) return e;
  return hipSuccess;
}

struct RelBufs {
  uint32_t* deg;
  void* tmp;
  size_t tmp_bytes = 0;
};

hipError_t rel_bufs(Scratch& S, uint32_t n, uint32_t* row_ptr, RelBufs& b) {
  hipError_t e;
  if ((e = S.alloc(&b.deg, n)) != hipSuccess) return e;
  const int items = n > 0 ? (int)n : 1;
  if ((e = hipcub::DeviceScan::InclusiveSum(nullptr, b.tmp_bytes, b.deg, row_ptr, items, S.s)) != hipSuccess)
    return e;
  return S.alloc((uint8_t**)&b.tmp, b.tmp_bytes);
}

uint64_t loc_bytes(uint32_t n, uint32_t nnz) {
  fl_arena a;
  a.kind = fl_arena::COUNT;
  {
    Scratch S(nullptr, &a);
    LocBufs b;
    loc_bufs(S, n, nnz, b);
  }
  return a.peak;
}

uint64_t rel_bytes(uint32_t n) {
  fl_arena a;
  a.kind = fl_arena::COUNT;
  {
    Scratch S(nullptr, &a);
    RelBufs b;
    rel_bufs(S, n, nullptr, b);
  }
  return a.peak;
}

// Breadth-first order of C clusters (neighbour lists nb[ptr[c] .. ptr[c+1]),
// repeats allowed), each component from its lowest-numbered cluster; noff:
// the first new label of each position (cluster c holds cstart[c+1] -
// cstart[c] sites).
void cluster_bfs(uint32_t C, const uint32_t* ptr, const uint32_t* nb, const uint32_t* cstart, uint32_t* corder,
                 uint32_t* noff) {
  std::unique_ptr<uint64_t[]> vis(new uint64_t[(C + 63) / 64]());
  uint32_t head = 0, tail = 0, label = 0;
  auto visit = [&](uint32_t c) {
    vis[c >> 6] |= 1ull << (c & 63);
    corder[tail] = c;
    noff[tail++] = label;
    label += cstart[c + 1] - cstart[c];
  };
  for (uint32_t s = 0; s < C; s++) {
    if (vis[s >> 6] >> (s & 63) & 1) continue;
    visit(s);
    while (head < tail) {
      // the rows of the clusters a few places ahead are fetched early
      if (head + 8 < tail) __builtin_prefetch(ptr + corder[head + 8]);
      if (head + 4 < tail) __builtin_prefetch(nb + ptr[corder[head + 4]]);
      const uint32_t c = corder[head++];
      for (uint32_t e = ptr[c]; e < ptr[c + 1]; e++) {
        const uint32_t d = nb[e];
        if (!(vis[d >> 6] >> (d & 63) & 1)) visit(d);
      }
    }
  }
}

}  // namespace

void fl_device_fill_tid(hipStream_t s, uint32_t n, uint32_t G, const uint32_t* site_off, uint32_t* tid) {
  if (n) hipLaunchKernelGGL(k_tid, grid_for(n), dim3(256), 0, s, n, G, site_off, tid);
}

void fl_device_mark_u8(hipStream_t s, uint32_t n, uint32_t count, const uint32_t* idx, uint8_t* dst) {
  // whole words: every device buffer is a multiple of 256 bytes
  const uint64_t words = ((uint64_t)n + 3) / 4;
  hipLaunchKernelGGL(k_fill32, grid_for(words), dim3(256), 0, s, words, (uint32_t*)dst, 0u);
  if (count) hipLaunchKernelGGL(k_mark_u8, grid_for(count), dim3(256), 0, s, count, idx, dst);
}

int fl_device_check_graph(hipStream_t s, const fl_raw_union& R, const double* len, std::vector<uint32_t>& bad_entry,
                          std::vector<uint32_t>& asym_entry, std::vector<uint8_t>& len_differs, std::string* err,
                          fl_arena* arena) {
  const uint32_t G = R.G;
  bad_entry.assign(G, FL_NONE);
  asym_entry.assign(G, FL_NONE);
  len_differs.assign(G, 0);
  if (R.n == 0) return FL_OK;
  Scratch S(s, arena);
  uint32_t* d = nullptr;
  HIPTRY(S.alloc(&d, 3ull * G));
  hipLaunchKernelGGL(k_fill32, grid_for(2ull * G), dim3(256), 0, s, 2ull * G, d, FL_NONE);
  hipLaunchKernelGGL(k_fill32, grid_for(G), dim3(256), 0, s, (uint64_t)G, d + 2 * G, 0u);
  hipLaunchKernelGGL(k_check_rows, grid_for(R.n), dim3(256), 0, s, R, d);
  hipLaunchKernelGGL(k_check_sym, grid_for(R.n), dim3(256), 0, s, R, len, d + G, d + 2 * G);
  std::vector<uint32_t> h(3ull * G);
  HIPTRY(hipMemcpyAsync(h.data(), d, h.size() * 4, hipMemcpyDeviceToHost, s));
  HIPTRY(hipStreamSynchronize(s));
  for (uint32_t g = 0; g < G; g++) {
    bad_entry[g] = h[g];
    asym_entry[g] = h[G + g];
    len_differs[g] = h[2 * G + g] != 0;
  }
  return FL_OK;
}

int fl_device_locality(hipStream_t s, const fl_raw_union& R, uint32_t* old_of_new, uint32_t* new_of_old,
                       uint32_t* clusters, std::string* err, fl_arena* arena) {
  const uint32_t n = R.n;
  if (clusters) *clusters = 0;
  if (n == 0) return FL_OK;
  Scratch S(s, arena);
  HIPTRY(S.reserve(loc_bytes(n, R.nnz)));
  LocBufs b;
  HIPTRY(loc_bufs(S, n, R.nnz, b));
  const dim3 B(256), Gn = grid_for(n);
  // 1. hooks and their roots
  hipLaunchKernelGGL(k_hook, Gn, B, 0, s, R, b.p);
  hipLaunchKernelGGL(k_fill32, dim3(1), B, 0, s, (uint64_t)kMaxFlags, b.flags, 0u);
  uint32_t h_flags[kMaxFlags];
  uint32_t rounds = 0;
  for (uint32_t target = 4;; target += 4) {
    if (target > kMaxFlags) {
      *err = "locality clustering did not converge";
      return FL_ERR_HIP;
    }
    for (; rounds < target; rounds++)
      hipLaunchKernelGGL(k_hook_jump, Gn, B, 0, s, n, b.p, b.flags + rounds, rounds ? b.flags + rounds - 1 : nullptr);
    HIPTRY(hipMemcpyAsync(h_flags, b.flags, rounds * 4, hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    if (h_flags[rounds - 1] == 0) break;
  }
  // 2. clusters numbered and grouped
  hipLaunchKernelGGL(k_root_flags, Gn, B, 0, s, n, b.p, b.val);
  HIPTRY(hipcub::DeviceScan::InclusiveSum(b.tmp_scan, b.scan_bytes, b.val, b.y, (int)n, s));
  hipLaunchKernelGGL(k_cid, Gn, B, 0, s, n, b.p, b.y, b.cid, b.val);
  uint32_t C = 0;
  HIPTRY(hipMemcpyAsync(&C, b.y + (n - 1), 4, hipMemcpyDeviceToHost, s));
  HIPTRY(hipStreamSynchronize(s));
  if (C == 0 || C > n) {
    *err = "locality clustering produced no clusters";
    return FL_ERR_HIP;
  }
  int bits = 1;
  while (bits < 32 && (1ull << bits) < C) bits++;
  HIPTRY(hipcub::DeviceRadixSort::SortPairs(b.tmp_sort, b.sort_bytes, b.cid, b.skey, b.val, b.ssite, (int)n, 0, bits,
                                            s));
  hipLaunchKernelGGL(k_cluster_start, Gn, B, 0, s, n, C, b.skey, b.cstart);
  // 3. neighbour lists to the host, breadth-first order of the clusters
  const dim3 Gc = grid_for(C);
  hipLaunchKernelGGL(k_cluster_nbrs<false>, Gc, B, 0, s, R, C, b.cstart, b.ssite, b.cid, b.cnt, nullptr);
  hipLaunchKernelGGL(k_fill32, dim3(1), B, 0, s, (uint64_t)1, b.cnt + C, 0u);
  HIPTRY(hipcub::DeviceScan::ExclusiveSum(b.tmp_scan2, b.scan2_bytes, b.cnt, b.cptr, (int)C + 1, s));
  hipLaunchKernelGGL(k_cluster_nbrs<true>, Gc, B, 0, s, R, C, b.cstart, b.ssite, b.cid, b.cptr, b.nbrs);
  std::unique_ptr<uint32_t[]> h_ptr(new uint32_t[C + 1]), h_start(new uint32_t[C + 1]);
  HIPTRY(hipMemcpyAsync(h_ptr.get(), b.cptr, (C + 1ull) * 4, hipMemcpyDeviceToHost, s));
  HIPTRY(hipMemcpyAsync(h_start.get(), b.cstart, (C + 1ull) * 4, hipMemcpyDeviceToHost, s));
  HIPTRY(hipStreamSynchronize(s));
  const uint32_t total = h_ptr[C];
  if (total > R.nnz) {
    *err = "locality clustering: neighbour lists out of range";
    return FL_ERR_HIP;
  }
  std::unique_ptr<uint32_t[]> h_nb(new uint32_t[total > 0 ? total : 1]);
  if (total) HIPTRY(hipMemcpyAsync(h_nb.get(), b.nbrs, total * 4ull, hipMemcpyDeviceToHost, s));
  HIPTRY(hipStreamSynchronize(s));
  std::unique_ptr<uint32_t[]> h_order(new uint32_t[2ull * C]);
  cluster_bfs(C, h_ptr.get(), h_nb.get(), h_start.get(), h_order.get(), h_order.get() + C);
  // 4. labels
  HIPTRY(hipMemcpyAsync(b.corder, h_order.get(), C * 4ull, hipMemcpyHostToDevice, s));
  HIPTRY(hipMemcpyAsync(b.noff, h_order.get() + C, C * 4ull, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_place, Gc, B, 0, s, C, b.corder, b.noff, b.cstart, b.ssite, old_of_new, new_of_old);
  // the host arrays above must outlive the copy
  HIPTRY(hipStreamSynchronize(s));
  if (clusters) *clusters = C;
  return FL_OK;
}

int fl_device_relabel(hipStream_t s, const fl_raw_union& R, const double* len_raw, const uint32_t* old_of_new,
                      const uint32_t* new_of_old, uint32_t* row_ptr, uint32_t* col, double* len,
                      const double* const src[5], double* const dst[5], std::string* err, fl_arena* arena) {
  const uint32_t n = R.n;
  if (n == 0) return FL_OK;
  Scratch S(s, arena);
  HIPTRY(S.reserve(rel_bytes(n)));
  RelBufs b;
  HIPTRY(rel_bufs(S, n, row_ptr, b));
  const dim3 B(256), Gn = grid_for(n);
  hipLaunchKernelGGL(k_new_degree, Gn, B, 0, s, R, old_of_new, b.deg);
  hipLaunchKernelGGL(k_fill32, dim3(1), B, 0, s, (uint64_t)1, row_ptr, 0u);
  HIPTRY(hipcub::DeviceScan::InclusiveSum(b.tmp, b.tmp_bytes, b.deg, row_ptr + 1, (int)n, s));
  hipLaunchKernelGGL(k_relabel_rows, Gn, B, 0, s, R, len_raw, old_of_new, new_of_old, row_ptr, col, len);
  SiteArrays A;
  for (int a = 0; a < 5; a++) {
    A.src[a] = src[a];
    A.dst[a] = dst[a];
  }
  hipLaunchKernelGGL(k_gather_sites, Gn, B, 0, s, n, old_of_new, A);
  HIPTRY(hipGetLastError());
  return FL_OK;
}

uint64_t fl_create_scratch_bytes(uint32_t n, uint32_t nnz, uint32_t G) {
  return std::max({fl_round256(3ull * 4 * G), loc_bytes(n, nnz), rel_bytes(n)});
}

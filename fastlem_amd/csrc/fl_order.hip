// fl_order.hip -- device part of the lake-traversal visit order.
//
// The traversal of stream_tree.rs:184-240 pops the outlets first (all keys
// 0.0) and then follows Prim's algorithm from the outlet set.  When all edge
// lengths are positive and pairwise distinct the Prim phase is determined by
// the minimum spanning forest of the graph with the outlets contracted; this
// file computes that forest on the GPU (Boruvka with the outlets pre-merged
// into one component) and returns its edges sorted by length.  The host turns
// them into ranks (fl_ranks_from_msf) after the outlet phase
// (fl_outlet_pop_order).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

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

// Every undirected edge (from its smaller endpoint): lengths must be
// positive (not NaN) and pairwise distinct for the fast form; a hash set of
// the length bits finds repeats.
__global__ void k_len_check(uint32_t n, const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                            const double* __restrict__ len, unsigned long long* __restrict__ table, uint32_t mask,
                            uint32_t* __restrict__ bad) {
  uint32_t i = gid();
  if (i >= n) return;
  for (uint32_t e = rp[i]; e < rp[i + 1]; ++e) {
    if (col[e] <= i) continue;
    double w = len[e];
    if (!(w > 0.0)) {
      *bad = 1;
      continue;
    }
    unsigned long long key = len_bits(w);
    uint32_t h = hash64(key) & mask;
    for (uint32_t probe = 0; probe <= mask; ++probe) {
      unsigned long long prev = atomicCAS(table + h, 0ull, key);
      if (prev == 0ull) break;
      if (prev == key) {
        *bad = 1;
        break;
      }
      h = (h + 1) & mask;
    }
  }
}

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
  uint32_t cv = comp[v], be = FL_NONE;
  unsigned long long best = ~0ull;
  for (uint32_t e = rp[v]; e < rp[v + 1]; ++e) {
    if (comp[col[e]] == cv) continue;
    unsigned long long b = len_bits(len[e]);
    if (b < best) {
      best = b;
      be = e;
    }
  }
  vbest[v] = be;
  if (be != FL_NONE) atomicMin(cbest + cv, best);
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
  if (c >= n) return;
  parent[c] = c;
  if (comp[c] != c || cbest[c] == ~0ull) return;
  uint32_t e = cedge[c];
  uint32_t t = comp[col[e]];
  if (cbest[t] == cbest[c] && c < t) return;
  parent[c] = t;
  uint32_t slot = atomicAdd(count, 1u);
  msf_a[slot] = csrc[c];
  msf_b[slot] = col[e];
  msf_w[slot] = cbest[c];
}

__global__ void k_bv_jump(uint32_t n, const uint32_t* __restrict__ parent, uint32_t* __restrict__ parent2,
                          uint32_t* __restrict__ moved) {
  uint32_t c = gid();
  if (c >= n) return;
  uint32_t p = parent[c], pp = parent[p];
  parent2[c] = pp;
  if (pp != p) *moved = 1;
}

__global__ void k_bv_relabel(uint32_t n, const uint32_t* __restrict__ parent, uint32_t* __restrict__ comp,
                             unsigned long long* __restrict__ cbest) {
  uint32_t v = gid();
  if (v >= n) return;
  comp[v] = parent[comp[v]];
  cbest[v] = ~0ull;
}

__global__ void k_gather_pairs(uint32_t m, const uint32_t* __restrict__ idx, const uint32_t* __restrict__ a,
                               const uint32_t* __restrict__ b, uint32_t* __restrict__ a2, uint32_t* __restrict__ b2) {
  uint32_t i = gid();
  if (i >= m) return;
  a2[i] = a[idx[i]];
  b2[i] = b[idx[i]];
}

__global__ void k_iota(uint32_t m, uint32_t* __restrict__ out) {
  uint32_t i = gid();
  if (i < m) out[i] = i;
}

inline dim3 grid_for(uint64_t n) { return dim3((uint32_t)((n + 255) / 256)); }

// Scratch of one forest build, stream-ordered on the build's stream: no
// device-wide synchronisation, and the memory returns to the device's pool
// (which the contexts keep) for the next build.
struct DevBuf {
  hipStream_t s;
  std::vector<void*> ptrs;
  hipError_t alloc(void** p, size_t bytes) {
    hipError_t e = hipMallocAsync(p, bytes < 8 ? 8 : bytes, s);
    if (e == hipSuccess) ptrs.push_back(*p);
    return e;
  }
  ~DevBuf() {
    for (void* p : ptrs) (void)hipFreeAsync(p, s);
  }
};

#define HIPTRY(call)                                               \
  do {                                                             \
    hipError_t e_ = (call);                                        \
    if (e_ != hipSuccess) {                                        \
      *err = std::string(#call) + ": " + hipGetErrorString(e_);    \
      return e_ == hipErrorOutOfMemory ? FL_ERR_OUT_OF_MEMORY : FL_ERR_HIP; \
    }                                                              \
  } while (0)

}  // namespace

// Minimum spanning forest of the graph whose sites start in the components
// init_comp (host array: each terrain's outlets share one label, every other
// site is its own), edges sorted by increasing length, into
// (a_sorted, b_sorted).
// *applies is false (and nothing else is produced) when some edge length is
// not positive or two lengths are equal: then the traversal order depends on
// the heap's tie handling beyond the outlet phase, and the caller simulates
// it exactly instead.
int fl_device_msf(hipStream_t s, uint32_t n, uint32_t nnz, const uint32_t* rp, const uint32_t* col, const double* len,
                  const std::vector<uint32_t>& init_comp, std::vector<uint32_t>& a_sorted,
                  std::vector<uint32_t>& b_sorted, bool* applies, std::string* err) {
  *applies = false;
  DevBuf buf{s, {}};
  uint32_t* d_flag = nullptr;
  HIPTRY(buf.alloc((void**)&d_flag, 4 * sizeof(uint32_t)));
  HIPTRY(hipMemsetAsync(d_flag, 0, 4 * sizeof(uint32_t), s));
  uint32_t cap = 1;
  while (cap < nnz) cap <<= 1;
  unsigned long long* table = nullptr;
  HIPTRY(buf.alloc((void**)&table, (size_t)cap * 8));
  HIPTRY(hipMemsetAsync(table, 0, (size_t)cap * 8, s));
  hipLaunchKernelGGL(k_len_check, grid_for(n), dim3(256), 0, s, n, rp, col, len, table, cap - 1, d_flag);
  uint32_t h_flag[4] = {0, 0, 0, 0};
  HIPTRY(hipMemcpyAsync(h_flag, d_flag, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIPTRY(hipStreamSynchronize(s));
  if (h_flag[0]) return FL_OK;

  uint32_t *comp, *parent, *parent2, *vbest, *cedge, *csrc, *ma, *mb, *idx, *idx2, *a2, *b2;
  unsigned long long *cbest, *mw, *mw2;
  for (uint32_t** p : {&comp, &parent, &parent2, &vbest, &cedge, &csrc, &ma, &mb, &idx, &idx2, &a2, &b2})
    HIPTRY(buf.alloc((void**)p, (size_t)n * 4));
  for (unsigned long long** p : {&cbest, &mw, &mw2}) HIPTRY(buf.alloc((void**)p, (size_t)n * 8));
  HIPTRY(hipMemcpyAsync(parent, init_comp.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_bv_init, grid_for(n), dim3(256), 0, s, n, parent, comp, cbest);
  uint32_t m = 0;
  for (uint32_t round = 0;; ++round) {
    if (round > 64) {
      *err = "minimum spanning forest did not converge";
      return FL_ERR_HIP;
    }
    hipLaunchKernelGGL(k_bv_best, grid_for(n), dim3(256), 0, s, n, rp, col, len, comp, cbest, vbest);
    hipLaunchKernelGGL(k_bv_pick, grid_for(n), dim3(256), 0, s, n, len, comp, cbest, vbest, cedge, csrc);
    hipLaunchKernelGGL(k_bv_hook, grid_for(n), dim3(256), 0, s, n, col, comp, cbest, cedge, csrc, parent, ma, mb, mw,
                       d_flag + 1);
    for (uint32_t j = 0;; ++j) {
      HIPTRY(hipMemsetAsync(d_flag + 2, 0, sizeof(uint32_t), s));
      hipLaunchKernelGGL(k_bv_jump, grid_for(n), dim3(256), 0, s, n, parent, parent2, d_flag + 2);
      std::swap(parent, parent2);
      HIPTRY(hipMemcpyAsync(h_flag, d_flag, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      HIPTRY(hipStreamSynchronize(s));
      if (!h_flag[2]) break;
      if (j > 64) {
        *err = "forest hooking did not converge";
        return FL_ERR_HIP;
      }
    }
    hipLaunchKernelGGL(k_bv_relabel, grid_for(n), dim3(256), 0, s, n, parent, comp, cbest);
    if (h_flag[1] == m) break;  // no component hooked this round
    m = h_flag[1];
    if (m > n) {
      *err = "forest edge count out of range";
      return FL_ERR_HIP;
    }
  }
  a_sorted.assign(m, 0);
  b_sorted.assign(m, 0);
  if (m) {
    hipLaunchKernelGGL(k_iota, grid_for(m), dim3(256), 0, s, m, idx);
    size_t tmp_bytes = 0;
    HIPTRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, mw, mw2, idx, idx2, (int)m, 0, 64, s));
    void* tmp = nullptr;
    HIPTRY(buf.alloc(&tmp, tmp_bytes));
    HIPTRY(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, mw, mw2, idx, idx2, (int)m, 0, 64, s));
    hipLaunchKernelGGL(k_gather_pairs, grid_for(m), dim3(256), 0, s, m, idx2, ma, mb, a2, b2);
    HIPTRY(hipMemcpyAsync(a_sorted.data(), a2, (size_t)m * 4, hipMemcpyDeviceToHost, s));
    HIPTRY(hipMemcpyAsync(b_sorted.data(), b2, (size_t)m * 4, hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
  }
  *applies = true;
  return FL_OK;
}

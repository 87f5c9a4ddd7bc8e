// fl_internal.h -- shared declarations of the library's translation units.
#pragma once
#include "../../include/fastlem_hip.h"

struct fl_rng;

#ifdef __cplusplus
#include <hip/hip_runtime.h>

// Set the calling thread's fl_last_error() message.
void fl_internal_set_error(const char* msg);

// The context's stream and device (for the library's other parts).
int32_t fl_ctx_stream(fl_ctx* c, hipStream_t* stream, int* device);

#include <mutex>
#include <string>
#include <vector>

// fl_noalloc.hip: the static device scratch of fl_generate_f64_noalloc.
// fl_noalloc_scratch_init takes its address on `device` (which must be
// current), once; fl_noalloc_scratch returns it, its size and the mutex that
// serialises the calls using it.
int fl_noalloc_scratch_init(int device);
int fl_noalloc_scratch(int device, uint8_t** ptr, uint64_t* bytes, std::mutex** call_mu);

// Where device buffers come from.
//   POOL:   the device's memory pool, stream-ordered (hipMallocAsync /
//           hipFreeAsync);
//   REGION: a fixed device region handed out front to back, no allocation
//           call; buffers are released by rewinding `used`;
//   COUNT:  nowhere: the bytes are counted and the pointers stay null (sizing).
// Every buffer takes a multiple of 256 bytes, so the three agree on sizes.
struct fl_arena {
  enum Kind { POOL = 0, REGION = 1, COUNT = 2 };
  int kind = POOL;
  uint8_t* base = nullptr;
  uint64_t size = 0;  // REGION: bytes available
  uint64_t used = 0;  // REGION / COUNT: bytes handed out
  uint64_t peak = 0;  // REGION / COUNT: largest `used` seen
  hipMemPool_t pool = nullptr;  // POOL: the library's pool of the device (null: the default pool)
};
inline uint64_t fl_round256(uint64_t bytes) { return ((bytes < 8 ? 8 : bytes) + 255) & ~255ull; }
// A buffer of `bytes` from `a` (null = POOL) on stream s.  A full REGION
// returns hipErrorOutOfMemory and sets *p to null.
hipError_t fl_arena_alloc(fl_arena* a, hipStream_t s, void** p, uint64_t bytes);

// fl_order.hip, the device part of the lake-traversal visit order.  Scratch
// comes from `arena` (null: the device pool) and is released before return.
//
// Per terrain g of a union (terrain of each site: tid, device, null for one
// terrain; entry_off: G + 1 union entry offsets, host): bad[g] = 1 when its
// edge lengths are not all positive and pairwise distinct.
int fl_device_len_check(hipStream_t s, uint32_t n, const uint32_t* rp, const uint32_t* col, const double* len,
                        const uint32_t* tid, uint32_t G, const uint32_t* entry_off, std::vector<uint8_t>& bad,
                        std::string* err, fl_arena* arena);
struct fl_visit_stats {
  uint32_t forest_edges = 0;
  double forest_ms = 0.0;  // length order of the forest included
  double ranks_ms = 0.0;
};
// Prim ranks of the traversal after its outlet phase, from the minimum
// spanning forest of the graph with the sites of each init_comp component
// (host, n labels: each terrain's outlets share one outlet's label) merged:
// every site reached from an outlet through a forest edge gets
// rank[v] = d_base[tid[v]] + its position in Prim order.  Other entries of
// `rank` are not written.  Lengths must be positive and pairwise distinct
// within every component that holds an outlet.
int fl_device_ranks(hipStream_t s, uint32_t n, const uint32_t* rp, const uint32_t* col, const double* len,
                    const uint8_t* is_outlet, const uint32_t* tid, const std::vector<uint32_t>& init_comp,
                    const std::vector<uint32_t>& base, uint32_t* rank, fl_visit_stats* st, std::string* err,
                    fl_arena* arena);
// dst[idx[q]] = val[q] for the `count` host pairs.
int fl_device_scatter_u32(hipStream_t s, uint32_t count, const uint32_t* idx, const uint32_t* val, uint32_t* dst,
                          std::string* err, fl_arena* arena);
// order[rank[v]] = v for every site v with a rank (rank != UINT32_MAX).
void fl_device_order(hipStream_t s, uint32_t n, const uint32_t* rank, uint32_t* order);
// Peak scratch bytes of fl_device_len_check, fl_device_ranks and
// fl_device_scatter_u32 (up to n pairs) for n sites and nnz entries.
uint64_t fl_msf_scratch_bytes(uint32_t n, uint32_t nnz);

// fl_capi.cpp: lanes (continuous batching, used by the serving loop).  A
// member is one request's terrain in a lane: its caller tag, max_iteration,
// the relabelling of its sites and, once it stopped, its iteration count and
// elevations in the caller's order.
struct fl_lane_member {
  uint64_t tag = 0;
  uint32_t cap = 0xffffffffu;
  uint32_t iterations = 0;
  std::vector<uint32_t> new_of_old;
  std::vector<double> result;
};
struct fl_lane;
struct fl_terrain;
fl_lane* fl_lane_create(fl_ctx* c);
void fl_lane_destroy(fl_lane* lane);
// A terrain ready to join a lane: validated, relabelled, uploaded, with its
// visit order (on context c, synchronised).
int32_t fl_lane_prepare(fl_ctx* c, const fl_terrain_desc* d, fl_terrain** out, std::vector<uint32_t>* new_of_old);
// Between iterations only (fl_lane_quiet): the prepared terrains join the lane.
int32_t fl_lane_join(fl_lane* lane, std::vector<std::pair<fl_terrain*, fl_lane_member>>& join);
// One confirmed iteration; members that stopped in it are appended to `stopped`.
int32_t fl_lane_step(fl_lane* lane, std::vector<fl_lane_member>& stopped);
bool fl_lane_idle(const fl_lane* lane);
bool fl_lane_quiet(const fl_lane* lane);
// Stop enqueueing ahead, so that the lane becomes quiet (to join terrains).
void fl_lane_set_drain(fl_lane* lane);
uint64_t fl_lane_active_sites(const fl_lane* lane);
uint32_t fl_lane_terrains(const fl_lane* lane);
// HBM the union arrays of one terrain take (fl_request_device_bytes without
// the visit order's scratch).
uint64_t fl_terrain_arrays_bytes(const fl_terrain_desc* d);

// fl_setup.cpp: per-terrain forms on a union graph whose terrain occupies
// sites [s0, s0 + nl) (its rows reference only those sites).
int fl_outlet_pop_order_range(uint32_t s0, uint32_t nl, const uint32_t* row_ptr, const uint32_t* col,
                              const uint32_t* outlets, size_t k, uint32_t* order_out);
int fl_lake_visit_order_range(uint32_t s0, uint32_t nl, const uint32_t* row_ptr, const uint32_t* col,
                              const double* len, const uint32_t* outlets, size_t k, uint32_t rank_base,
                              uint32_t* rank_out);
// Non-outlet visit ranks from the spanning forest (outlets in any order).
// With site_off (groups + 1 entries: group g owns sites [site_off[g],
// site_off[g+1]) and no edge joins two groups) the groups are processed
// independently on the threads; without it, as one forest.
int fl_krt_ranks(size_t n, size_t groups, const uint32_t* outlet_off, const uint32_t* outlets,
                 const uint32_t* rank_off, const uint32_t* site_off, const uint32_t* msf_a, const uint32_t* msf_b,
                 size_t m, uint32_t* rank_out, unsigned threads);
// The one-forest form in two steps: the working arrays set up from the
// sites and outlets alone (while the forest is still being built), then the
// ranks from the forest.  fl_krt_prepare returns null and sets *rc on invalid
// outlets; fl_krt_ranks_prepared consumes the arrays (release them after).
struct fl_krt_work;
fl_krt_work* fl_krt_prepare(size_t n, size_t groups, const uint32_t* outlet_off, const uint32_t* outlets,
                            unsigned threads, int* rc);
int fl_krt_ranks_prepared(fl_krt_work* work, size_t groups, const uint32_t* outlet_off, const uint32_t* outlets,
                          const uint32_t* rank_off, const uint32_t* msf_a, const uint32_t* msf_b, size_t m,
                          uint32_t* rank_out, unsigned threads);
void fl_krt_release(fl_krt_work* work);
// Host threads for the one-time setup work (at most 16).
unsigned fl_host_threads();
extern "C" int fl_ranks_from_msf_groups(size_t n, size_t groups, const uint32_t* outlet_off,
                                        const uint32_t* outlet_order, const uint32_t* rank_off,
                                        const uint32_t* msf_a, const uint32_t* msf_b, size_t m,
                                        uint32_t* rank_out);
#endif

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
// Scratch taken from an arena and given back when the scope ends: pool
// buffers are freed in stream order, region arenas rewind to where the scope
// began.  reserve(bytes) on a pool takes one block of that size, from which
// the scope's later buffers are cut (one pool allocation and one free for
// the whole scope: the pool's cost grows with the number of its blocks).
struct fl_scratch {
  hipStream_t s;
  fl_arena* arena;
  uint64_t mark;
  std::vector<void*> ptrs;
  fl_arena block;  // the reserved block, as a region
  fl_scratch(hipStream_t s_, fl_arena* a) : s(s_), arena(a), mark(a ? a->used : 0) {}
  fl_scratch(const fl_scratch&) = delete;
  fl_scratch& operator=(const fl_scratch&) = delete;
  hipError_t reserve(uint64_t bytes) {
    if ((arena && arena->kind != fl_arena::POOL) || !ptrs.empty() || block.base) return hipSuccess;
    void* p = nullptr;
    hipError_t e = fl_arena_alloc(arena, s, &p, bytes);
    if (e != hipSuccess) return e;
    ptrs.push_back(p);
    block.kind = fl_arena::REGION;
    block.base = (uint8_t*)p;
    block.size = fl_round256(bytes);
    return hipSuccess;
  }
  template <class T>
  hipError_t alloc(T** p, uint64_t count) {
    if (block.base) return fl_arena_alloc(&block, s, (void**)p, count * sizeof(T));
    hipError_t e = fl_arena_alloc(arena, s, (void**)p, count * sizeof(T));
    if (e == hipSuccess && (!arena || arena->kind == fl_arena::POOL)) ptrs.push_back((void*)*p);
    return e;
  }
  ~fl_scratch() {
    for (void* p : ptrs) (void)hipFreeAsync(p, s);
    if (arena && arena->kind != fl_arena::POOL) arena->used = mark;
  }
};

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
// size[v] = the number of sites in v's subtree of the forest `next` (a root
// points at itself; roots get n): the heavy-child hint of an iteration with
// no drainage area of a similar tree (a solve's first, the one after lake
// joins).  Sites of terrains that are not active (tid and t_active, device;
// null for one terrain) count as roots.  Stream-ordered: nothing is read
// back (convergence is not checked: the sizes are only a hint).
int fl_device_subtree_sizes(hipStream_t s, uint32_t n, const uint32_t* next, const uint32_t* tid,
                            const uint8_t* t_active, double* size, std::string* err, fl_arena* arena);
uint64_t fl_subtree_scratch_bytes(uint32_t n);
// Peak scratch bytes of fl_device_len_check, fl_device_ranks and
// fl_device_scatter_u32 (up to n pairs) for n sites and nnz entries.
uint64_t fl_msf_scratch_bytes(uint32_t n, uint32_t nnz);

// fl_create.hip, the device part of an upload.  The caller's terrains as
// given ("raw"), packed as a union: terrain g holds union sites
// [site_off[g], site_off[g+1]) and union entries [entry_off[g],
// entry_off[g+1]); its n_g + 1 row offsets (terrain-local entry indices) sit
// at rp[site_off[g] + g ...] and its neighbour indices are terrain-local.
// Every pointer is a device pointer; tid (terrain of each site) is null for
// one terrain.  Row offsets must be checked (spanning each terrain's entries,
// non-decreasing) and rows at most FL_DEVICE_MAX_ROW long before any of
// these run; the locality order and the relabelling need graphs that passed
// fl_device_check_graph.
enum { FL_DEVICE_MAX_ROW = 64 };
struct fl_raw_union {
  uint32_t n = 0, nnz = 0, G = 1;
  const uint32_t* site_off = nullptr;   // [G + 1]
  const uint32_t* entry_off = nullptr;  // [G + 1]
  const uint32_t* tid = nullptr;        // [n] or null
  const uint32_t* rp = nullptr;         // [n + G]
  const uint32_t* col = nullptr;        // [nnz]
};
// tid[i] = the terrain holding union site i.
void fl_device_fill_tid(hipStream_t s, uint32_t n, uint32_t G, const uint32_t* site_off, uint32_t* tid);
// dst[0..n) = 0, then dst[idx[q]] = 1 for the `count` device indices.
void fl_device_mark_u8(hipStream_t s, uint32_t n, uint32_t count, const uint32_t* idx, uint8_t* dst);
// Per terrain: the smallest terrain-local entry index that is out of range,
// a self-loop or a repeat within its row (UINT32_MAX: none); the smallest
// entry i->j whose row j holds no i (among entries that pass the first
// check); whether the two entries of some edge hold different length bits.
int fl_device_check_graph(hipStream_t s, const fl_raw_union& R, const double* len, std::vector<uint32_t>& bad_entry,
                          std::vector<uint32_t>& asym_entry, std::vector<uint8_t>& len_differs, std::string* err,
                          fl_arena* arena);
// A locality order of the union's sites that keeps every terrain in its own
// site range (old_of_new, new_of_old: device, union labels).
int fl_device_locality(hipStream_t s, const fl_raw_union& R, uint32_t* old_of_new, uint32_t* new_of_old,
                       uint32_t* clusters, std::string* err, fl_arena* arena);
// The relabelled graph (row_ptr [n+1], col and len [nnz], union labels) and
// five per-site arrays dst[a][u] = src[a][old_of_new[u]] (a null source
// gives NaN).
int fl_device_relabel(hipStream_t s, const fl_raw_union& R, const double* len_raw, const uint32_t* old_of_new,
                      const uint32_t* new_of_old, uint32_t* row_ptr, uint32_t* col, double* len,
                      const double* const src[5], double* const dst[5], std::string* err, fl_arena* arena);
// Peak scratch of the three calls above for n sites, nnz entries, G terrains.
uint64_t fl_create_scratch_bytes(uint32_t n, uint32_t nnz, uint32_t G);

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
// The outlet phase from the outlets' own rows: outlet q (of k, in listed
// order) has deg[q] entries, of which nb[nb_ptr[q] .. nb_ptr[q+1]) lead to
// other outlets (their indices q'); order_out receives outlet indices.
int fl_outlet_pop_order_compact(size_t k, const uint32_t* deg, const uint32_t* nb_ptr, const uint32_t* nb,
                                uint32_t* order_out);
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

// fl_kernels.h -- device buffers of one terrain and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

// Per-kernel timing with HIP events recorded on the launch stream.
struct fl_profiler {
  struct Rec {
    const char* name;
    hipEvent_t a, b;
    uint64_t units;  // work items of the launch (sites or chain positions)
  };
  struct Total {
    uint64_t launches = 0;
    double ms = 0.0;
    uint64_t units = 0;
  };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, Total> totals;
  hipEvent_t get();
  void record(hipStream_t s, const char* name, hipEvent_t a, hipEvent_t b, uint64_t units);
  void flush();
  void reset();
  ~fl_profiler();
};

// Set by the C ABI for the duration of a call on a profiled context.
extern thread_local fl_profiler* fl_current_profiler;

// Control block read back once per phase: [flags | counters | boff head | bnoff head].
enum { FL_FLAG_LAKE = 0, FL_FLAG_JUMP = 1 /* and 2: alternating */, FL_FLAG_CHANGED = 3, FL_NUM_FLAGS = 4 };
enum { FL_CNT_MAXLD = 0, FL_CNT_CHAINS = 1, FL_CNT_POSITIONS = 2, FL_NUM_COUNTERS = 4 };
enum { FL_CTRL_HEAD = 512 };  // bucket offsets mirrored into the control block
enum { FL_MAX_JUMP_ROUNDS = 96 };  // one "pointers moved" word per doubling round
enum { FL_CTRL_JUMPS = FL_NUM_FLAGS + FL_NUM_COUNTERS + 2 * (FL_CTRL_HEAD + 1) };
enum { FL_CTRL_WORDS = FL_CTRL_JUMPS + FL_MAX_JUMP_ROUNDS };

// Most sites of a terrain or a union of terrains: the pointer jumping keeps
// its counts (chain positions, light depths) below 2^31, the top bit of the
// count word marking a pair at its fixed point.
#define FL_MAX_SITES 0x7fffffffull

// Read-only inputs of one terrain, resident in HBM.
struct fl_dev_terrain {
  uint32_t n = 0;
  uint32_t nnz = 0;
  uint32_t* row_ptr = nullptr;   // [n+1]
  uint32_t* col = nullptr;       // [nnz]
  double* len = nullptr;         // [nnz]
  double* area = nullptr;        // [n]
  double* erod = nullptr;        // [n]
  double* uplift = nullptr;      // [n]
  double* tan_ms = nullptr;      // [n] tan(max_slope), NaN when None
  double* init_h = nullptr;      // [n]
  bool any_max_slope = false;
  uint8_t* is_outlet = nullptr;  // [n]
  uint32_t* rank = nullptr;      // [n] lake-traversal visit rank (UINT32_MAX: never)
  uint32_t* order = nullptr;     // [n] visited sites by rank
};

// Per-iteration state and scratch of one terrain.
struct fl_dev_work {
  double* h = nullptr;      // [n] elevations (current)
  double* h2 = nullptr;     // [n] elevations being computed
  double* xp = nullptr;     // [n] 1/celerity*distance, chain-major
  double* taup = nullptr;   // [n] response time, chain-major
  double* hnp = nullptr;    // [n] new elevation, chain-major (max-slope form)
  double* up = nullptr;     // [n] uplift, chain-major (max-slope form)
  double* tanp = nullptr;   // [n] tan(max slope), chain-major (max-slope form)
  double* dp = nullptr;     // [n] receiver distance, chain-major (max-slope form)
  double* hop = nullptr;    // [n] old elevation, chain-major (max-slope form)
  double* A = nullptr;      // [n] drainage area (kept: heavy-child hint)
  double* tau = nullptr;    // [n] response time
  double* dnext = nullptr;  // [n] distance to the receiver
  uint32_t* next = nullptr; // [n]
  // lakes
  uint32_t* sub = nullptr;
  uint32_t* sub2 = nullptr;
  uint64_t* lake_key = nullptr;
  // decomposition
  uint64_t* cmask = nullptr;  // [n] child slots of each site
  uint32_t* hc = nullptr;     // [n] heavy child
  // drainage-area inputs in chain-major order (written by k_chain_perm):
  // list counts, own area, erodibility, receiver distance; the listed light
  // children's drainage areas (FL_LIST slots per position, fold order,
  // filled by the children's rounds: pdest at a chain's base position names
  // the slot of its head)
  uint32_t* pcnt = nullptr;   // [n] npre | npost << 8, or the overflow flag
  double* parea = nullptr;    // [n]
  double* pK = nullptr;       // [n]
  double* pdn = nullptr;      // [n]
  double* cval = nullptr;     // [FL_LIST n]
  uint32_t* pdest = nullptr;  // [n]

  // pointer jumping, two (pointer, count) pairs per site, pointer in the low
  // word: cp = (chain ancestor, distance to it), rt = (tree ancestor, light
  // edges to it); converged: (chain head, position), (root, light depth)
  uint64_t* cp = nullptr;
  uint64_t* rt = nullptr;
  uint32_t *bcnt = nullptr, *bnodes = nullptr, *boff = nullptr, *bnoff = nullptr, *bcur = nullptr,
           *bncur = nullptr;                                  // [n+2] light-depth buckets
  // chains by bucket: (base position, length, position of the head's
  // receiver or FL_NONE, position of the basin outlet)
  uint4* cmeta = nullptr;                                    // [n]
  uint32_t* head_chain = nullptr;                            // [n] at a chain head: its chain
  uint32_t* head_base = nullptr;                             // [n] at a chain head: its base position
  uint32_t* perm = nullptr;                                  // [n] chain-major site order
  uint32_t* ctrl = nullptr;                                  // [FL_CTRL_WORDS]
  // movement flag of the last pointer-jumping round launched: the kernels
  // after the decomposition do nothing unless it is zero (converged)
  const uint32_t* guard = nullptr;
  int* flags = nullptr;                                      // = ctrl
  uint32_t* counters = nullptr;                              // = ctrl + FL_NUM_FLAGS
  // chains of at least long_min sites fold in a wavefront, shorter ones in
  // one lane (a tuning: results do not depend on it)
  uint32_t long_min = 6;
  // Basins recomputed in an iteration.  A new elevation is a function of the
  // drainage tree alone (and the static parameters), so a basin of the
  // previous iteration's tree in which no site's receiver changed -- nor any
  // receiver changed to or from one of its sites -- keeps its tree and its
  // elevations bit for bit; only the other basins' sites are active.
  const uint32_t* next_prev = nullptr;  // [n] the previous iteration's tree
  const double* dnext_prev = nullptr;   // [n]
  const uint32_t* root_prev = nullptr;  // [n] its basin (root) of every site
  uint32_t* root = nullptr;             // [n] this iteration's, for the next one
  uint8_t* dirty = nullptr;             // [n] by root: basin to recompute
  uint8_t* act = nullptr;               // [n] site recomputed in this iteration
  // [n] the previous iteration changed the elevation of the site or of a
  // neighbour (zeroed by k_receivers, marked by k_tau_scatter)
  const uint8_t* chg_prev = nullptr;
  uint8_t* chg = nullptr;               // [n] ... this one
};

// Workgroup size of the per-site kernels launched by the calling thread (a
// tuning: results do not depend on it; 64..1024, a multiple of 64).
extern thread_local uint32_t fl_site_block;

// keep_unchanged: sites whose neighbourhood's elevations the previous
// iteration did not change keep its receivers (valid when that iteration's
// tree had no lake joins).
// mark: flag the previous basins that the changed receivers touch (w.dirty,
// cleared by the previous iteration's k_tau_scatter).
void fl_launch_receivers(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, bool keep_unchanged,
                         bool mark);
void fl_launch_sub_init(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w);
void fl_launch_sub_jump(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, bool flip);
void fl_launch_lake_join(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w);
// Marks the dirty basins from the final receivers after a lake join (the
// receivers marked those of their own changes; k_tau_scatter clears the
// marks for the next iteration).
void fl_launch_dirty(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w);
// Children, heavy children and chain pointers of the iteration's active
// sites: every site of the terrains still iterating (`full`, or no previous
// tree), else those of the dirty basins (tid, t_active: per-site terrain and
// per-terrain mask of a union, or null).
void fl_launch_children(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, bool full, const uint32_t* tid,
                        const uint8_t* t_active);
void fl_launch_chain_jump(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, uint32_t round);
void fl_launch_chain_count(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w);
// Bucket offsets (scan), chain placement and the chain-major permutation;
// offsets are written up to round `rounds` (rounds past the deepest light
// depth are empty).
void fl_launch_chain_layout(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, uint32_t rounds);
// Launch shape of one sweep round; the round's chain and position bounds are
// read by the kernels from the bucket offsets (boff/bnoff at 2l..2l+2), so a
// round can be enqueued before its bounds are known on the host.  Kernels
// loop over their range, so any non-zero grid is correct.
struct fl_round_grid {
  uint32_t fold_blocks;  // k_da_fold / k_tau_fold, 64 threads each
  uint64_t positions;    // profiler units (0 when not known at launch)
};
// Round l of the drainage-area sweep; also writes the response-time inputs
// of its positions (and, with clamp, the max-slope recurrence's inputs).
void fl_launch_da_round(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, uint32_t l,
                        const fl_round_grid& g, bool clamp);
// Rounds lsplit..min(deepest light depth, cap) of the drainage-area sweep
// (deepest first), then of the response-time sweep (shallowest first), in
// one launch of one workgroup of FL_TAIL_WAVES wavefronts each: for the
// deep rounds, which hold few chains.
enum { FL_TAIL_WAVES = 16 };
void fl_launch_da_tail(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, uint32_t lsplit, uint32_t cap,
                       uint64_t units, bool clamp);
void fl_launch_tau_tail(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, uint32_t lsplit, uint32_t cap,
                        uint64_t units, bool clamp);
void fl_launch_clear_iter(hipStream_t s, const fl_dev_work& w, uint32_t nb);
void fl_launch_tau_round(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, uint32_t l,
                         const fl_round_grid& g, bool clamp);
// Back to site order, one thread per site.
void fl_launch_tau_scatter(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, bool clamp,
                           const uint32_t* tid, const uint8_t* active, uint32_t* tchanged);
void fl_launch_fill_u32(hipStream_t s, uint32_t* p, uint64_t count, uint32_t v);

// Gather of whole terrains of a union into a smaller union (in the same
// order, each keeping its internal layout): new terrain q takes the sites
// [new_off[q], new_off[q+1]) from old sites starting at old_off[q], and its
// entries from old entry old_eoff[q] to new entry new_eoff[q].
struct fl_compact_map {
  uint32_t count;
  const uint32_t* new_off;   // [count + 1]
  const uint32_t* old_off;   // [count]
  const uint32_t* new_eoff;  // [count]
  const uint32_t* old_eoff;  // [count]
};
// Copies the graph, parameters, visit ranks and order, current elevations
// and drainage areas (shifting site and entry indices) and writes new_tid,
// for the new sites [v0, v1): the terrains of the map, all from one source
// union, which become terrains tid0, tid0 + 1, ... of the new union.
void fl_launch_compact(hipStream_t s, const fl_compact_map& m, const fl_dev_terrain& ot, const fl_dev_work& ow,
                       const fl_dev_terrain& nt, const fl_dev_work& nw, uint32_t* new_tid, uint32_t v0, uint32_t v1,
                       uint32_t tid0);
// Verification: the reference's BFS traversal of the basin of `outlet`.
void fl_launch_basin_order(hipStream_t s, const fl_dev_terrain& t, const fl_dev_work& w, uint32_t outlet,
                           uint32_t* order, uint32_t* len_out);

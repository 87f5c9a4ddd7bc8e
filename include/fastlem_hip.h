/*
 * fastlem_hip.h -- C ABI of the MI355X (gfx950) uplift-erosion solve.
 *
 * This library replaces the body of fastlem's TerrainGenerator::generate()
 * (src/lem/generator.rs:118-210 of fastlem 0.1.4): receivers, lake labelling
 * and joining, per-outlet drainage basins, drainage area, response time and
 * elevation, iterated until no elevation changes or max_iteration.  The
 * caller keeps the model build, parameter validation (generator.rs:91-116),
 * outlet resolution (generator.rs:120-132) and the elevation jitter
 * (generator.rs:134-138) and passes their results in.
 *
 * Conventions
 *   - Every function returns int32_t: FL_OK (0) or a negative FL_ERR_* code;
 *     fl_last_error() returns the calling thread's last message.
 *   - Host pointers are borrowed for the duration of the call only.
 *   - Indices are uint32_t; a terrain, or a batch or served union of
 *     terrains, may hold up to 2^31-1 sites and 2^32-2 directed adjacency
 *     entries (FL_ERR_TOO_LARGE beyond).
 *   - Arithmetic is IEEE binary64 evaluated exactly as the reference writes
 *     it (no contraction), so results are bit-identical to the reference
 *     given the same inputs.
 *   - A context is bound to one HIP device and one stream; calls on one
 *     context must not overlap.  Different contexts are independent.
 */
#ifndef FASTLEM_HIP_H
#define FASTLEM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FL_OK = 0,
  FL_ERR_INVALID = -1,     /* bad argument (null pointer, index out of range) */
  FL_ERR_HIP = -2,         /* HIP runtime failure; message has the HIP error */
  FL_ERR_OUT_OF_MEMORY = -3,
  FL_ERR_TOO_LARGE = -4,   /* sizes beyond the uint32 index range */
  FL_ERR_NAN_LENGTH = -5,  /* NaN edge length: the reference heap panics
                              (stream_tree.rs:36 unwraps partial_cmp) */
  FL_ERR_NO_DEVICE = -6,
  FL_ERR_BUDGET = -7,      /* request exceeds the server's HBM budget */
};

#define FL_UNBOUNDED_ITERATIONS 0xffffffffu /* max_iteration = None */

/* One terrain as generate() sees it through the Model trait
 * (src/core/traits.rs:13-20) plus the resolved per-site parameters
 * (src/core/parameters.rs:24-30). */
typedef struct fl_terrain_desc {
  uint64_t n;                    /* model.num() (model.rs:42-44) */
  uint64_t nnz;                  /* total adjacency entries = row_ptr[n] */
  const uint32_t *row_ptr;       /* [n+1] CSR row offsets */
  const uint32_t *col;           /* [nnz] neighbors_of(i) site indices, in order */
  const double *length;          /* [nnz] neighbors_of(i) edge attributes */
  const double *area;            /* [n] model.areas() */
  const double *erodibility;     /* [n] TopographicalParameters::erodibility */
  const double *uplift_rate;     /* [n] TopographicalParameters::uplift_rate */
  const uint8_t *has_max_slope;  /* [n] max_slope.is_some(); NULL = all None */
  const double *max_slope;       /* [n] max_slope in radians (tan taken here) */
  const double *init_elevation;  /* [n] base_elevation + jitter (generator.rs:134-138) */
  const uint32_t *outlets;       /* [n_outlets] resolved outlets, generator.rs:120-132 order */
  uint64_t n_outlets;
  uint32_t max_iteration;        /* FL_UNBOUNDED_ITERATIONS for None */
  uint32_t reserved;
} fl_terrain_desc;

typedef struct fl_ctx fl_ctx;         /* device + stream + scratch */
typedef struct fl_terrain fl_terrain; /* one terrain resident in HBM */

const char *fl_last_error(void);
const char *fl_version(void);

/* Context on HIP device `device`. */
int32_t fl_ctx_create(int32_t device, fl_ctx **out);
int32_t fl_ctx_destroy(fl_ctx *ctx);

/* Drop-in for generator.rs:118-210: upload, iterate, download.
 * elevations_out: [n]; iterations_out (may be NULL): outer iterations run. */
int32_t fl_generate_f64(fl_ctx *ctx, const fl_terrain_desc *desc, double *elevations_out,
                        uint32_t *iterations_out);

/* Allocation-free form of fl_generate_f64 (same arguments, same results).
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (tight serving loops, code that must not call hipMalloc).
 * Every device buffer of the call (the terrain's arrays and the visit order's
 * scratch) lives in a static device array of FL_NOALLOC_SCRATCH_BYTES that
 * each GPU running the library holds until the process ends; the call makes
 * no device or pinned-host allocation and frees nothing (host heap memory is
 * still used for staging).  A terrain that needs more than the array holds
 * (fl_request_device_bytes; about 120,000 sites at 6 adjacency entries per
 * site) is refused with FL_ERR_TOO_LARGE before anything runs.  All calls on
 * one GPU share its array and would race on it: the library serialises them
 * within the process.  The context's device must be the one it was created
 * on (fl_ctx_create takes the array's device address). */
#define FL_NOALLOC_SCRATCH_BYTES (64ull << 20)
int32_t fl_generate_f64_noalloc(fl_ctx *ctx, const fl_terrain_desc *desc, double *elevations_out,
                                uint32_t *iterations_out);

/* Device-resident form: upload once, solve many times. */
int32_t fl_terrain_create(fl_ctx *ctx, const fl_terrain_desc *desc, fl_terrain **out);
int32_t fl_terrain_destroy(fl_terrain *t);
/* Reset elevations to init_elevation and iterate (generator.rs:140-210). */
int32_t fl_terrain_solve(fl_ctx *ctx, fl_terrain *t, uint32_t max_iteration, uint32_t *iterations_out);
int32_t fl_terrain_download(fl_ctx *ctx, fl_terrain *t, double *elevations_out);
/* Device pointer of the current elevations (f64[n]), for callers that keep
 * the result in HBM. */
int32_t fl_terrain_device_elevations(fl_terrain *t, double **dev_ptr);
/* Bytes of HBM the terrain holds. */
int32_t fl_terrain_device_bytes(const fl_terrain *t, uint64_t *bytes);

/* Batched form: `count` independent terrains packed into one object and
 * iterated together, one launch sequence per iteration for all of them.
 * Terrain g stops when an iteration changes none of its elevations or after
 * descs[g].max_iteration iterations, and is then frozen while the others go
 * on, so every terrain's elevations and iteration count equal its solo run.
 * The handle is an fl_terrain (destroy with fl_terrain_destroy). */
int32_t fl_batch_create(fl_ctx *ctx, const fl_terrain_desc *descs, uint32_t count, fl_terrain **out);
int32_t fl_batch_solve(fl_ctx *ctx, fl_terrain *batch, uint32_t *iterations_out /* [count] or NULL */);
int32_t fl_batch_download(fl_ctx *ctx, fl_terrain *batch, uint32_t index, double *elevations_out);
int32_t fl_batch_count(const fl_terrain *batch, uint32_t *count);
/* A batch solve gathers its still-running terrains into a smaller union when
 * they hold less than half of the current one and at least min_idle_sites
 * sites belong to terrains that have stopped (default 65536; UINT32_MAX never
 * compacts).  Results do not depend on it. */
int32_t fl_ctx_set_batch_compaction(fl_ctx *ctx, uint32_t min_idle_sites);
/* Mode 1 (the default): iterations after the first are enqueued whole, the
 * next one while the device still runs the current one, their
 * pointer-jumping and sweep round counts taken from the last confirmed
 * iteration (an iteration whose assumptions did not hold is redone
 * synchronously from its input, which three elevation buffers keep intact).
 * Mode 0 synchronises inside every iteration instead; 2 speculates a single
 * round of each kind, so that most iterations take the redo path (for
 * testing it); 3 is mode 1 with each iteration replayed from a captured HIP
 * graph.  Results do not depend on the mode. */
int32_t fl_ctx_set_pipelining(fl_ctx *ctx, int32_t enable);
/* Which basins an iteration recomputes.  A new elevation depends on the
 * drainage tree alone, so by default (0) an iteration recomputes only the
 * basins of the previous iteration's tree in which some receiver changed
 * (or that a changed receiver leaves or enters); the others keep their
 * elevations bit for bit.  1 recomputes every basin in every iteration, as
 * the reference does (for measurement).  Results do not depend on it. */
int32_t fl_ctx_set_recompute_all(fl_ctx *ctx, int32_t all);
/* The heavy child of a site continues its chain; it is chosen by the previous
 * iteration's drainage areas, or by exact subtree sizes where those describe
 * another tree.  An iteration whose decomposition needs more than `rounds`
 * sweep rounds is decomposed again with subtree sizes (0, the default:
 * 2 log2(n) + 8 for n sites; sizes keep the rounds within log2(n) + 1).
 * Results do not depend on it. */
int32_t fl_ctx_set_rehint_rounds(fl_ctx *ctx, uint32_t rounds);

/* Launch tuning.  None of these change any result (every setting gives the
 * same bits); they shape the work:
 *   site_block        workgroup size of the per-site kernels (64..1024, a
 *                     multiple of 64);
 *   long_chain        chains of at least this many sites fold in a wavefront
 *                     with LDS-staged tiles, shorter ones in one lane (the
 *                     switch between the two forms of a sweep round, >= 1);
 *   tail_units        sweep rounds of at most this many fold units run
 *                     together, one launch per sweep (levels per launch; 0:
 *                     one launch per round);
 *   pipelining        0, 1 or 3 as fl_ctx_set_pipelining (3: HIP graphs);
 *   compact_min_idle  batches: as fl_ctx_set_batch_compaction.
 * Without fl_ctx_set_tuning a solve takes the library's table, chosen by
 * fl_autotune per workload class (site count, one terrain or a batch);
 * fl_ctx_set_pipelining / fl_ctx_set_batch_compaction override those fields. */
typedef struct fl_tuning {
  uint32_t site_block, long_chain, tail_units, pipelining, compact_min_idle;
  uint32_t reserved[3];
} fl_tuning;
/* A fixed tuning for every later call on the context; NULL returns to the
 * table. */
int32_t fl_ctx_set_tuning(fl_ctx *ctx, const fl_tuning *tuning);
/* The tuning a solve of `sites` sites (a batch when terrains > 1) takes on
 * this context. */
int32_t fl_ctx_tuning_for(fl_ctx *ctx, uint64_t sites, uint32_t terrains, fl_tuning *out);
/* Autotuner: coordinate descent over the tunings on a resident terrain or
 * batch, each candidate timed over `repeats` full solves (the visit order
 * cached) and checked to give the same elevations and iteration counts bit
 * for bit.  *best: the fastest; *best_ms and *table_ms: its solve time and
 * that of the table's tuning; *candidates: tunings tried.  The context's
 * tuning is left as it was. */
int32_t fl_autotune(fl_ctx *ctx, fl_terrain *t, uint32_t repeats, fl_tuning *best, double *best_ms,
                    double *table_ms, uint32_t *candidates);

/* Serving: a request loop on one GPU with continuous batching.  A request
 * is admitted (first come, first served) as soon as the HBM it needs fits in
 * hbm_budget_bytes (fl_request_device_bytes plus a second copy of its
 * arrays, held while it joins a batch); a preparation worker validates,
 * uploads and orders it; it then joins a lane at the lane's next iteration
 * boundary.  A lane iterates a union of terrains in shared launches
 * (fl_batch_* semantics: every terrain stops at its own convergence or
 * max_iteration, so its result equals its solo fl_generate_f64), and a
 * terrain leaves as soon as it stops.  `slots` streams are used in all (at
 * most 4: one lane, the rest preparing).  The caller keeps the description's
 * arrays and elevations_out alive until fl_server_wait returns for that
 * ticket.  fl_server_destroy drains the queue. */
typedef struct fl_server fl_server;
int32_t fl_server_create(int32_t device, uint32_t slots, uint64_t hbm_budget_bytes, fl_server **out);
int32_t fl_server_submit(fl_server *server, const fl_terrain_desc *desc, double *elevations_out, uint64_t *ticket);
/* Blocks until the request finished; returns its status.  queue_ms: time
 * from submission to admission; latency_ms: submission to completion. */
int32_t fl_server_wait(fl_server *server, uint64_t ticket, uint32_t *iterations_out, double *queue_ms,
                       double *latency_ms);
int32_t fl_server_stats(fl_server *server, uint64_t *peak_bytes, uint32_t *peak_running);
int32_t fl_server_destroy(fl_server *server);
/* What the server did so far: terrains that joined an idle lane / a lane
 * already iterating others, the most terrains one lane iterated together,
 * and wall time summed over workers (preparation, lane iterations, joins,
 * lanes waiting for prepared terrains). */
typedef struct fl_server_counts {
  uint64_t joins_idle, joins_busy, steps;
  uint32_t max_lane_terrains, lanes, prep_workers, reserved;
  double prep_ms, step_ms, join_ms, idle_ms;
} fl_server_counts;
int32_t fl_server_counters(fl_server *server, fl_server_counts *out);

/* Statistics of the last fl_terrain_solve on a terrain. */
typedef struct fl_solve_stats {
  uint32_t iterations;      /* outer iterations run */
  uint32_t lake_iterations; /* iterations that joined lakes */
  uint32_t host_syncs;      /* host-device synchronisations */
  uint32_t order_from_forest; /* 1: visit order from the spanning forest,
                                 0: exact heap simulation (ties present) */
  double visit_order_ms;    /* wall time of the lake-traversal visit order */
  double solve_ms;          /* wall time of the solve */
  double outlet_phase_ms;   /* visit order parts: outlet phase (host, runs */
  double forest_ms;         /*   concurrently with) the GPU spanning forest, */
  double ranks_ms;          /*   then the ranks from the forest (device) */
  uint32_t compactions;     /* batch: times the running terrains were gathered */
  uint32_t redone;          /* pipelined iterations redone synchronously */
  /* fl_terrain_create / fl_batch_create: validation, relabelling (the
   * locality order and the relabelled arrays), allocation and upload */
  double create_validate_ms, create_relabel_ms, create_upload_ms;
  uint32_t create_on_device;  /* 1: graphs checked and relabelled on the
                                 device, 0: on host threads (rows longer
                                 than 64 entries, or a header error) */
  uint32_t create_clusters;   /* device path: clusters of the locality order */
  uint64_t recomputed_sites;  /* sites whose basin was recomputed, summed over
                                 the solve's iterations (n x iterations when
                                 every basin is; see fl_ctx_set_recompute_all) */
  uint32_t first_sweep_rounds; /* sweep rounds (light depth of the heavy-chain
                                  decomposition, + 1) of the first iteration */
  uint32_t max_sweep_rounds;   /*   and the most of any iteration */
  uint32_t rehinted;           /* iterations decomposed again with subtree
                                  sizes as the heavy-child hint (the previous
                                  drainage areas gave a far deeper one) */
} fl_solve_stats;
int32_t fl_terrain_last_stats(const fl_terrain *t, fl_solve_stats *out);

/* Forget cached per-terrain host work (the lake-traversal visit order), so
 * the next solve recomputes it as a fresh generate() call would. */
int32_t fl_terrain_invalidate_cache(fl_terrain *t);

/* Kernel timing: when enabled, every launch on the context's stream is
 * bracketed by HIP events and its duration accumulated per kernel name. */
int32_t fl_ctx_set_profiling(fl_ctx *ctx, int32_t enable);
/* Accumulated launches, milliseconds and work items (units: sites for the
 * per-site kernels, chain positions for the sweep rounds; may be NULL) of
 * kernel `name` since the last reset; fl_ctx_reset_profile clears all
 * counters. */
int32_t fl_ctx_kernel_time(fl_ctx *ctx, const char *name, uint64_t *launches, double *total_ms, uint64_t *units);
int32_t fl_ctx_reset_profile(fl_ctx *ctx);

/* Teacher-forced stages (verification entry points): run one stage on the
 * device from the given state and return its output.
 *   receivers:   stream_tree.rs:109-137            -> next[n]
 *   stream tree: stream_tree.rs:72-99 (incl. lakes) -> next[n], subroot[n]
 *                (subroot = pre-join chain end), returns has_lake in *flag
 *   iteration:   generator.rs:141-204 from elev     -> elev updated in place,
 *                next, drainage area, response time; *flag = changed */
int32_t fl_stage_receivers(fl_ctx *ctx, fl_terrain *t, const double *elev, uint32_t *next_out);
int32_t fl_stage_stream_tree(fl_ctx *ctx, fl_terrain *t, const double *elev, uint32_t *next_out,
                             uint32_t *subroot_out, int32_t *flag);
int32_t fl_stage_iteration(fl_ctx *ctx, fl_terrain *t, double *elev, uint32_t *next_out, double *area_out,
                           double *response_out, int32_t *flag);
/* The reference's traversal of one basin (drainage_basin.rs:13-36) on the
 * stream tree of elev: order_out[0..*len_out) (room for n entries).  The
 * sweeps never need it (they follow heavy-path chains); it verifies the tree
 * order. */
int32_t fl_stage_basin_order(fl_ctx *ctx, fl_terrain *t, const double *elev, uint32_t outlet, uint32_t *order_out,
                             uint32_t *len_out);

/* The lake traversal's visit order (stream_tree.rs:184-240) as the solve
 * uses it: rank_out[i] = position of caller site i in the order in which
 * its terrain's traversal first pops it (outlets first), UINT32_MAX when it
 * is never reached.  For a batch, i runs over the terrains' sites in order
 * and positions are within each terrain. */
int32_t fl_stage_visit_order(fl_ctx *ctx, fl_terrain *t, uint32_t *rank_out);

/* HBM a generate() call of this terrain holds at its peak: its device
 * arrays plus the visit order's scratch (the serving loop's admission
 * estimate).  No device is needed. */
int32_t fl_request_device_bytes(const fl_terrain_desc *desc, uint64_t *bytes);

/* Natural-neighbour (Sibson) interpolation of a terrain's elevations:
 * Terrain2D::get_elevation (src/models/surface/terrain.rs:36-38) through
 * TerrainInterpolator2D (src/models/surface/interpolator.rs:11-27).
 * fl_interp_create takes the sites and their Delaunay triangles (ntri x 3
 * site indices, either orientation) once; fl_interp_query evaluates m
 * points on the device: out[k] = the interpolated value of values[] (one per
 * site) at (qx[k], qy[k]) with found[k] = 1, or NaN with found[k] = 0 outside
 * the sites' convex hull (the reference returns None there).  A query equal
 * to a site returns that site's value. */
typedef struct fl_interp fl_interp;
int32_t fl_interp_create(fl_ctx *ctx, uint64_t n, const double *xs, const double *ys, uint64_t ntri,
                         const uint32_t *triangles, fl_interp **out);
int32_t fl_interp_query(fl_ctx *ctx, fl_interp *interp, const double *values, uint64_t m, const double *qx,
                        const double *qy, double *out, uint8_t *found);
int32_t fl_interp_destroy(fl_interp *interp);

/* ---- host setup helpers (one-time, CPU) ---- */

/* StdRng (rand 0.8.5 ChaCha12) restatement used for the jitter of
 * generator.rs:134-138: out[i] = base[i] + gen::<f64>() * f64::EPSILON with
 * StdRng::seed_from_u64(seed). */
void fl_jitter_elevations(uint64_t seed, size_t n, const double *base, double *out);
/* from_random_sites (builder.rs:38-52): StdRng::from_seed([0;32]),
 * gen_range(x) then gen_range(y) per site. */
void fl_random_sites(size_t n, double xmin, double ymin, double xmax, double ymax, double *xs, double *ys);
/* ChaCha block function (rounds = 12 for StdRng). */
void fl_chacha_block(int rounds, const uint32_t in[16], uint32_t out[16]);
/* Voronoi cells clipped to the box: areas (builder.rs:236-247) and
 * centroids (the Lloyd step of builder.rs:155-200). */
int fl_voronoi_cells(size_t n, const double *xs, const double *ys, size_t ntri, const int64_t *tris, double xmin,
                     double ymin, double xmax, double ymax, double *area_out, double *cx_out, double *cy_out);
/* Adjacency from triangles with the a<b rule (builder.rs:249-268).  Call with
 * col == NULL to size: row_ptr[n] = entry count. */
int fl_build_adjacency(size_t n, const double *xs, const double *ys, size_t ntri, const int64_t *tris,
                       uint32_t *row_ptr, uint32_t *col, double *len);
/* Visit order of the lake-removal heap (stream_tree.rs:184-240):
 * rank_out[i] = first-pop position of site i, UINT32_MAX if never reached. */
int fl_lake_visit_order(size_t n, const uint32_t *row_ptr, const uint32_t *col, const double *len,
                        const uint32_t *outlets, size_t n_outlets, uint32_t *rank_out, uint64_t *visited_out);

/* Pop order of the outlets in the traversal's first phase (all keys 0):
 * exact std BinaryHeap tie handling, driven only by how many unvisited
 * neighbours each popped outlet pushes (stream_tree.rs:185-237). */
int fl_outlet_pop_order(size_t n, const uint32_t *row_ptr, const uint32_t *col, const uint32_t *outlets, size_t k,
                        uint32_t *order_out);
/* Visit ranks from a minimum spanning forest of the graph with the outlets
 * contracted (edges sorted by increasing length; lengths positive and
 * pairwise distinct): outlets 0..k-1 in outlet_order, then Prim order. */
int fl_ranks_from_msf(size_t n, const uint32_t *outlet_order, size_t k, const uint32_t *msf_a, const uint32_t *msf_b,
                      size_t m, uint32_t *rank_out);

#ifdef __cplusplus
}
#endif

#endif /* FASTLEM_HIP_H */

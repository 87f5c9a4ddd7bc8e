/*
 * fl_oracle.c -- CPU restatement of fastlem's TerrainGenerator::generate()
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the
 * single-core CPU baseline ("port").  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the checker.
 * The product library (fastlem_amd/) never links or calls it.
 *
 * It restates, line by line and with the reference's own data structures,
 * the uplift-erosion loop of fastlem 0.1.4 (all paths relative to the
 * reference crate root):
 *
 *   src/lem/generator.rs:90-213      generate(): outlets, jitter, outer loop
 *   src/lem/stream_tree.rs:72-99     StreamTree::construct
 *   src/lem/stream_tree.rs:101-107   create_outlet_table
 *   src/lem/stream_tree.rs:109-137   construct_initial_stream_tree (receivers)
 *   src/lem/stream_tree.rs:139-173   find_roots_with_lakes
 *   src/lem/stream_tree.rs:175-243   remove_lakes_from_stream_tree (lazy Prim)
 *   src/lem/stream_tree.rs:15-44     RidgeElement ordering (reversed cmp)
 *   src/lem/drainage_basin.rs:13-46  DrainageBasin BFS and its iterators
 *
 * Data structures kept as in the reference so that the CPU timing is that of
 * the reference algorithm:
 *   - adjacency as one heap block per site holding (usize, f64) pairs, the
 *     layout of terrain-graph 1.0.1's Vec<Vec<(usize, T)>>;
 *   - a binary heap with Rust std BinaryHeap push/pop mechanics
 *     (sift_up stops on `elem <= parent`; pop moves the last element to the
 *     root and runs sift_down_to_bottom, which prefers the right child when
 *     `left <= right`, then sift_up);
 *   - one growable vector per drainage basin per iteration;
 *   - has_edge as a linear scan of the row returning the first match.
 *
 * Third-party semantics restated from their published behaviour (no Rust
 * toolchain exists here to run them): terrain-graph insertion-ordered rows
 * and first-match has_edge; std BinaryHeap sift rules; f64::powf(x, 0.5)
 * lowered by LLVM in optimised builds to (x == -inf ? +inf : fabs(sqrt(x)));
 * f64::max returning the non-NaN operand (C fmax).
 *
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math (see oracle/Makefile).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define FLO_NONE ((size_t)-1)

enum {
  FLO_OK = 0,
  FLO_ERR_ALLOC = -1,
  FLO_ERR_NAN_LENGTH = -2,  /* RidgeElement::cmp unwraps partial_cmp: panics */
  FLO_ERR_BAD_INDEX = -3,
};

/* ---------------------------------------------------------------- graph --- */

typedef struct {
  size_t index;
  double attr;
} flo_edge; /* (usize, f64) as returned by neighbors_of */

typedef struct {
  flo_edge *items;
  size_t len;
} flo_row;

typedef struct {
  size_t n;
  flo_row *rows;
} flo_graph;

/* Build the vector-of-vectors graph from a CSR that lists neighbors_of(i) in
 * order.  This mirrors holding the reference's graph; it is setup, not part
 * of the generate() loop. */
static int graph_from_csr(flo_graph *g, size_t n, const uint32_t *row_ptr,
                          const uint32_t *col, const double *len) {
  g->n = n;
  g->rows = (flo_row *)calloc(n ? n : 1, sizeof(flo_row));
  if (!g->rows) return FLO_ERR_ALLOC;
  for (size_t i = 0; i < n; i++) {
    size_t b = row_ptr[i], e = row_ptr[i + 1];
    g->rows[i].len = e - b;
    g->rows[i].items = (flo_edge *)malloc((e - b ? e - b : 1) * sizeof(flo_edge));
    if (!g->rows[i].items) return FLO_ERR_ALLOC;
    for (size_t k = b; k < e; k++) {
      if (col[k] >= n) return FLO_ERR_BAD_INDEX;
      g->rows[i].items[k - b].index = col[k];
      g->rows[i].items[k - b].attr = len[k];
    }
  }
  return FLO_OK;
}

static void graph_free(flo_graph *g) {
  if (!g->rows) return;
  for (size_t i = 0; i < g->n; i++) free(g->rows[i].items);
  free(g->rows);
  g->rows = NULL;
}

/* terrain-graph has_edge: first match in row a, else (false, default) */
static inline int graph_has_edge(const flo_graph *g, size_t a, size_t b, double *attr) {
  const flo_row *r = &g->rows[a];
  for (size_t k = 0; k < r->len; k++) {
    if (r->items[k].index == b) {
      *attr = r->items[k].attr;
      return 1;
    }
  }
  *attr = 0.0;
  return 0;
}

/* --------------------------------------------------- Rust BinaryHeap ------ */

typedef struct {
  size_t index;
  double dist;
} ridge_t; /* RidgeElement, stream_tree.rs:15-18 */

/* RidgeElement::cmp is other.dist.partial_cmp(self.dist) (stream_tree.rs:36):
 * `a <= b` holds exactly when a.dist >= b.dist.  The max-heap therefore pops
 * the smallest distance first. */
static inline int ridge_le(const ridge_t *a, const ridge_t *b) { return a->dist >= b->dist; }

typedef struct {
  ridge_t *data;
  size_t len, cap;
} rheap;

static size_t rheap_sift_up(rheap *h, size_t start, size_t pos) {
  ridge_t elem = h->data[pos];
  while (pos > start) {
    size_t parent = (pos - 1) / 2;
    if (ridge_le(&elem, &h->data[parent])) break;
    h->data[pos] = h->data[parent];
    pos = parent;
  }
  h->data[pos] = elem;
  return pos;
}

static void rheap_sift_down_to_bottom(rheap *h, size_t pos) {
  size_t end = h->len, start = pos;
  ridge_t elem = h->data[pos];
  size_t child = 2 * pos + 1;
  size_t lim = end >= 2 ? end - 2 : 0;
  while (child <= lim && end >= 2) {
    child += ridge_le(&h->data[child], &h->data[child + 1]) ? 1 : 0;
    h->data[pos] = h->data[child];
    pos = child;
    child = 2 * pos + 1;
  }
  if (child == end - 1) {
    h->data[pos] = h->data[child];
    pos = child;
  }
  h->data[pos] = elem;
  rheap_sift_up(h, start, pos);
}

static int rheap_push(rheap *h, ridge_t x) {
  if (h->len == h->cap) {
    size_t nc = h->cap ? h->cap * 2 : 16;
    ridge_t *nd = (ridge_t *)realloc(h->data, nc * sizeof(ridge_t));
    if (!nd) return FLO_ERR_ALLOC;
    h->data = nd;
    h->cap = nc;
  }
  size_t old_len = h->len;
  h->data[h->len++] = x;
  rheap_sift_up(h, 0, old_len);
  return FLO_OK;
}

static int rheap_pop(rheap *h, ridge_t *out) {
  if (h->len == 0) return 0;
  ridge_t item = h->data[--h->len];
  if (h->len > 0) {
    ridge_t t = h->data[0];
    h->data[0] = item;
    item = t;
    rheap_sift_down_to_bottom(h, 0);
  }
  *out = item;
  return 1;
}

/* ------------------------------------------------------------ vec ------- */

typedef struct {
  size_t *data;
  size_t len, cap;
} svec;

static int svec_push(svec *v, size_t x) {
  if (v->len == v->cap) {
    size_t nc = v->cap ? v->cap * 2 : 4; /* Rust RawVec: max(2*cap, 4) */
    size_t *nd = (size_t *)realloc(v->data, nc * sizeof(size_t));
    if (!nd) return FLO_ERR_ALLOC;
    v->data = nd;
    v->cap = nc;
  }
  v->data[v->len++] = x;
  return FLO_OK;
}

/* ------------------------------------------------------ stream tree ----- */

/* stream_tree.rs:101-107 */
static void create_outlet_table(size_t n, const size_t *outlets, size_t n_outlets, uint8_t *is_outlet) {
  memset(is_outlet, 0, n);
  for (size_t k = 0; k < n_outlets; k++) is_outlet[outlets[k]] = 1;
}

/* stream_tree.rs:109-137 */
static void construct_initial_stream_tree(const flo_graph *g, const double *elev,
                                          const uint8_t *is_outlet, size_t *next) {
  size_t n = g->n;
  for (size_t i = 0; i < n; i++) next[i] = i;
  for (size_t i = 0; i < n; i++) {
    if (is_outlet[i]) continue;
    double steepest_slope = 0.0;
    const flo_row *r = &g->rows[i];
    for (size_t k = 0; k < r->len; k++) {
      size_t j = r->items[k].index;
      if (elev[i] > elev[j]) {
        double distance = r->items[k].attr;
        double down_hill_slope = (elev[i] - elev[j]) / distance;
        if (down_hill_slope > steepest_slope) {
          steepest_slope = down_hill_slope;
          next[i] = j;
        }
      }
    }
  }
}

/* stream_tree.rs:139-173; subroot uses FLO_NONE for Option::None */
static int find_roots_with_lakes(size_t n, const uint8_t *is_outlet, const size_t *next, size_t *subroot) {
  for (size_t i = 0; i < n; i++) subroot[i] = is_outlet[i] ? i : FLO_NONE;
  int has_lake = 0;
  for (size_t i = 0; i < n; i++) {
    if (subroot[i] != FLO_NONE) continue;
    size_t iv = i;
    while (subroot[iv] == FLO_NONE && iv != next[iv]) iv = next[iv];
    size_t ir;
    if (subroot[iv] == FLO_NONE) {
      has_lake = 1;
      ir = iv;
    } else {
      ir = subroot[iv];
    }
    iv = i;
    while (subroot[iv] == FLO_NONE && iv != next[iv]) {
      subroot[iv] = ir;
      iv = next[iv];
    }
    subroot[iv] = ir;
  }
  return has_lake;
}

/* stream_tree.rs:175-243 */
static int remove_lakes_from_stream_tree(const flo_graph *g, const size_t *outlets, size_t n_outlets,
                                         const size_t *subroot, size_t *next /* in/out */) {
  size_t n = g->n;
  int rc = FLO_OK;
  size_t *root = (size_t *)malloc((n ? n : 1) * sizeof(size_t));
  uint8_t *visited = (uint8_t *)calloc(n ? n : 1, 1);
  rheap heap = {0};
  if (!root || !visited) {
    rc = FLO_ERR_ALLOC;
    goto done;
  }
  heap.cap = n ? n : 1; /* BinaryHeap::with_capacity(num) */
  heap.data = (ridge_t *)malloc(heap.cap * sizeof(ridge_t));
  if (!heap.data) {
    rc = FLO_ERR_ALLOC;
    goto done;
  }
  for (size_t i = 0; i < n; i++) root[i] = FLO_NONE;
  for (size_t k = 0; k < n_outlets; k++) {
    size_t o = outlets[k];
    root[o] = o;
    ridge_t e = {o, 0.0};
    if ((rc = rheap_push(&heap, e))) goto done;
  }
  ridge_t el;
  while (rheap_pop(&heap, &el)) {
    size_t i = el.index;
    if (visited[i]) continue;
    const flo_row *r = &g->rows[i];
    for (size_t s = 0; s < r->len; s++) {
      size_t j = r->items[s].index;
      if (visited[j]) continue;
      if (root[subroot[j]] == FLO_NONE) {
        size_t k = j, nk = i;
        for (;;) {
          if (next[k] != k) {
            size_t tmp = next[k];
            next[k] = nk;
            nk = k;
            k = tmp;
          } else {
            break;
          }
        }
        next[k] = nk;
        root[subroot[j]] = root[subroot[i]];
      }
      double distance = r->items[s].attr;
      if (isnan(distance)) { /* partial_cmp(..).unwrap() panics on NaN */
        rc = FLO_ERR_NAN_LENGTH;
        goto done;
      }
      ridge_t e = {j, distance};
      if ((rc = rheap_push(&heap, e))) goto done;
    }
    root[i] = root[subroot[i]];
    visited[i] = 1;
  }
done:
  free(root);
  free(visited);
  free(heap.data);
  return rc;
}

/* stream_tree.rs:72-99.  Returns has_lake (>=0) or an error (<0).
 * subroot_out (optional) receives the pre-join subroots. */
static int stream_tree_construct(const flo_graph *g, const double *elev, const uint8_t *is_outlet,
                                 const size_t *outlets, size_t n_outlets, size_t *next, size_t *subroot) {
  construct_initial_stream_tree(g, elev, is_outlet, next);
  int has_lake = find_roots_with_lakes(g->n, is_outlet, next, subroot);
  if (!has_lake) return 0;
  int rc = remove_lakes_from_stream_tree(g, outlets, n_outlets, subroot, next);
  return rc < 0 ? rc : 1;
}

/* drainage_basin.rs:13-36 */
static int drainage_basin_construct(const flo_graph *g, const size_t *next, size_t outlet, svec *trav) {
  trav->len = 0;
  int rc = svec_push(trav, outlet);
  if (rc) return rc;
  size_t i = 0;
  for (;;) {
    size_t it = trav->data[i];
    const flo_row *r = &g->rows[it];
    for (size_t k = 0; k < r->len; k++) {
      size_t jt = r->items[k].index;
      if (next[jt] == it)
        if ((rc = svec_push(trav, jt))) return rc;
    }
    i += 1;
    if (i >= trav->len) break;
  }
  return FLO_OK;
}

/* f64::powf(x, 0.5) as lowered by LLVM in optimised Rust builds */
static inline double rust_powf_half(double x) { return (x == -INFINITY) ? INFINITY : fabs(sqrt(x)); }

/* f64::max: returns the non-NaN operand */
static inline double rust_max(double a, double b) { return fmax(a, b); }

/* ---------------------------------------------------------- public API --- */

typedef struct {
  uint32_t iterations;      /* outer iterations executed */
  uint32_t lake_iterations; /* iterations in which has_lake was true */
  double seconds_tree;      /* filled by the timing wrapper only */
} flo_stats;

/* One outer iteration of generator.rs:140-210 on an already-built graph.
 * Returns changed (0/1) or an error (<0).  Optional outputs expose the
 * intermediate state for teacher-forced tests. */
static int one_iteration(const flo_graph *g, const double *area, const double *erod, const double *uplift,
                         const uint8_t *has_max_slope, const double *max_slope, const size_t *outlets,
                         size_t n_outlets, const uint8_t *is_outlet, double *elev, size_t *next,
                         size_t *subroot, double *da, double *rt, int *had_lake) {
  size_t n = g->n;
  const double m_exp = 0.5; /* DEFAULT_M_EXP, generator.rs:16 */
  (void)m_exp;
  int hl = stream_tree_construct(g, elev, is_outlet, outlets, n_outlets, next, subroot);
  if (hl < 0) return hl;
  if (had_lake) *had_lake = hl;
  memcpy(da, area, n * sizeof(double));
  for (size_t i = 0; i < n; i++) rt[i] = 0.0;
  int changed = 0;
  svec trav = {0};
  for (size_t q = 0; q < n_outlets; q++) {
    size_t outlet = outlets[q];
    int rc = drainage_basin_construct(g, next, outlet, &trav);
    if (rc) {
      free(trav.data);
      return rc;
    }
    /* drainage areas: for_each_downstream (generator.rs:154-159) */
    for (size_t t = trav.len; t-- > 0;) {
      size_t i = trav.data[t];
      size_t j = next[i];
      if (j != i) da[j] += da[i];
    }
    /* response times: for_each_upstream (generator.rs:162-174) */
    for (size_t t = 0; t < trav.len; t++) {
      size_t i = trav.data[t];
      size_t j = next[i];
      double edge;
      double distance = graph_has_edge(g, i, j, &edge) ? edge : 1.0;
      double celerity = erod[i] * rust_powf_half(da[i]);
      rt[i] += rt[j] + 1.0 / celerity * distance;
    }
    /* elevations (generator.rs:177-203) */
    for (size_t t = 0; t < trav.len; t++) {
      size_t i = trav.data[t];
      double new_elevation = elev[outlet] + uplift[i] * rust_max(rt[i] - rt[outlet], 0.0);
      if (has_max_slope && has_max_slope[i]) {
        size_t j = next[i];
        double edge;
        double distance = graph_has_edge(g, i, j, &edge) ? edge : 1.0;
        double ms = tan(max_slope[i]);
        double slope = (new_elevation - elev[j]) / distance;
        if (slope > ms) new_elevation = elev[j] + ms * distance;
      }
      changed |= new_elevation != elev[i];
      elev[i] = new_elevation;
    }
  }
  free(trav.data);
  return changed;
}

typedef struct {
  flo_graph g;
  size_t *outlets;
  size_t n_outlets;
  uint8_t *is_outlet;
  size_t *next, *subroot;
  double *da, *rt;
} flo_state;

static void state_free(flo_state *s) {
  graph_free(&s->g);
  free(s->outlets);
  free(s->is_outlet);
  free(s->next);
  free(s->subroot);
  free(s->da);
  free(s->rt);
}

static int state_init(flo_state *s, size_t n, const uint32_t *row_ptr, const uint32_t *col, const double *len,
                      const uint32_t *outlets, size_t n_outlets) {
  memset(s, 0, sizeof(*s));
  int rc = graph_from_csr(&s->g, n, row_ptr, col, len);
  if (rc) return rc;
  size_t nn = n ? n : 1;
  s->outlets = (size_t *)malloc((n_outlets ? n_outlets : 1) * sizeof(size_t));
  s->is_outlet = (uint8_t *)malloc(nn);
  s->next = (size_t *)malloc(nn * sizeof(size_t));
  s->subroot = (size_t *)malloc(nn * sizeof(size_t));
  s->da = (double *)malloc(nn * sizeof(double));
  s->rt = (double *)malloc(nn * sizeof(double));
  if (!s->outlets || !s->is_outlet || !s->next || !s->subroot || !s->da || !s->rt) return FLO_ERR_ALLOC;
  for (size_t k = 0; k < n_outlets; k++) {
    if (outlets[k] >= n) return FLO_ERR_BAD_INDEX;
    s->outlets[k] = outlets[k];
  }
  s->n_outlets = n_outlets;
  create_outlet_table(n, s->outlets, n_outlets, s->is_outlet);
  return FLO_OK;
}

/* generator.rs:140-210: iterate from init_elev (the jittered elevations of
 * generator.rs:134-138, supplied by the caller) until no elevation changes
 * or max_iteration iterations ran.  Outlets are already resolved
 * (generator.rs:120-132). */
int flo_generate(size_t n, const uint32_t *row_ptr, const uint32_t *col, const double *len,
                 const double *area, const double *erod, const double *uplift, const uint8_t *has_max_slope,
                 const double *max_slope, const double *init_elev, const uint32_t *outlets, size_t n_outlets,
                 uint32_t max_iteration, double *elev_out, flo_stats *stats) {
  flo_state s;
  int rc = state_init(&s, n, row_ptr, col, len, outlets, n_outlets);
  if (rc) {
    state_free(&s);
    return rc;
  }
  memcpy(elev_out, init_elev, n * sizeof(double));
  uint32_t it = 0, lakes = 0;
  for (uint32_t step = 0; step < max_iteration; step++) {
    int had_lake = 0;
    int changed = one_iteration(&s.g, area, erod, uplift, has_max_slope, max_slope, s.outlets, s.n_outlets,
                                s.is_outlet, elev_out, s.next, s.subroot, s.da, s.rt, &had_lake);
    if (changed < 0) {
      rc = changed;
      break;
    }
    it++;
    lakes += had_lake ? 1 : 0;
    if (!changed) break;
  }
  if (stats) {
    stats->iterations = it;
    stats->lake_iterations = lakes;
  }
  state_free(&s);
  return rc;
}

/* ---- per-stage entry points for teacher-forced tests (uint32 indices) ---- */

static void copy_out_idx(const size_t *src, uint32_t *dst, size_t n) {
  for (size_t i = 0; i < n; i++) dst[i] = (uint32_t)src[i];
}

/* Receivers only (stream_tree.rs:109-137). */
int flo_receivers(size_t n, const uint32_t *row_ptr, const uint32_t *col, const double *len, const double *elev,
                  const uint32_t *outlets, size_t n_outlets, uint32_t *next_out) {
  flo_state s;
  int rc = state_init(&s, n, row_ptr, col, len, outlets, n_outlets);
  if (!rc) {
    construct_initial_stream_tree(&s.g, elev, s.is_outlet, s.next);
    copy_out_idx(s.next, next_out, n);
  }
  state_free(&s);
  return rc;
}

/* find_roots_with_lakes on a given receiver array.  Returns has_lake. */
int flo_find_roots(size_t n, const uint32_t *next_in, const uint32_t *outlets, size_t n_outlets,
                   uint32_t *subroot_out) {
  uint8_t *is_outlet = (uint8_t *)malloc(n ? n : 1);
  size_t *next = (size_t *)calloc(n ? n : 1, sizeof(size_t));
  size_t *sub = (size_t *)malloc((n ? n : 1) * sizeof(size_t));
  size_t *outs = (size_t *)malloc((n_outlets ? n_outlets : 1) * sizeof(size_t));
  int rc = FLO_ERR_ALLOC;
  if (is_outlet && next && sub && outs) {
    for (size_t k = 0; k < n_outlets; k++) outs[k] = outlets[k];
    create_outlet_table(n, outs, n_outlets, is_outlet);
    for (size_t i = 0; i < n; i++) next[i] = next_in[i];
    rc = find_roots_with_lakes(n, is_outlet, next, sub);
    copy_out_idx(sub, subroot_out, n);
  }
  free(is_outlet);
  free(next);
  free(sub);
  free(outs);
  return rc;
}

/* Full StreamTree::construct (receivers, labels, lake removal).
 * Returns has_lake (0/1) or an error. */
int flo_stream_tree(size_t n, const uint32_t *row_ptr, const uint32_t *col, const double *len,
                    const double *elev, const uint32_t *outlets, size_t n_outlets, uint32_t *next_out,
                    uint32_t *subroot_out) {
  flo_state s;
  int rc = state_init(&s, n, row_ptr, col, len, outlets, n_outlets);
  if (!rc) {
    rc = stream_tree_construct(&s.g, elev, s.is_outlet, s.outlets, s.n_outlets, s.next, s.subroot);
    if (rc >= 0) {
      copy_out_idx(s.next, next_out, n);
      if (subroot_out) copy_out_idx(s.subroot, subroot_out, n);
    }
  }
  state_free(&s);
  return rc;
}

/* Lake removal alone on given (next, subroot) (stream_tree.rs:175-243). */
int flo_remove_lakes(size_t n, const uint32_t *row_ptr, const uint32_t *col, const double *len,
                     const uint32_t *outlets, size_t n_outlets, const uint32_t *subroot_in,
                     uint32_t *next_inout) {
  flo_state s;
  int rc = state_init(&s, n, row_ptr, col, len, outlets, n_outlets);
  if (!rc) {
    for (size_t i = 0; i < n; i++) {
      s.next[i] = next_inout[i];
      s.subroot[i] = subroot_in[i];
    }
    rc = remove_lakes_from_stream_tree(&s.g, s.outlets, s.n_outlets, s.subroot, s.next);
    if (!rc) copy_out_idx(s.next, next_inout, n);
  }
  state_free(&s);
  return rc;
}

/* The pop sequence of the lake-removal heap: order_out[k] is the k-th site
 * visited (first pop).  Returns the number of visited sites.  This exposes
 * the heap mechanics of stream_tree.rs:184-240 for tests; lake flipping does
 * not influence it. */
long flo_visit_order(size_t n, const uint32_t *row_ptr, const uint32_t *col, const double *len,
                     const uint32_t *outlets, size_t n_outlets, uint32_t *order_out) {
  flo_state s;
  long cnt = 0;
  int rc = state_init(&s, n, row_ptr, col, len, outlets, n_outlets);
  if (rc) {
    state_free(&s);
    return rc;
  }
  uint8_t *visited = (uint8_t *)calloc(n ? n : 1, 1);
  rheap heap = {0};
  for (size_t k = 0; k < n_outlets && rc == 0; k++) {
    ridge_t e = {s.outlets[k], 0.0};
    rc = rheap_push(&heap, e);
  }
  ridge_t el;
  while (rc == 0 && rheap_pop(&heap, &el)) {
    size_t i = el.index;
    if (visited[i]) continue;
    const flo_row *r = &s.g.rows[i];
    for (size_t q = 0; q < r->len; q++) {
      size_t j = r->items[q].index;
      if (visited[j]) continue;
      ridge_t e = {j, r->items[q].attr};
      if ((rc = rheap_push(&heap, e))) break;
    }
    visited[i] = 1;
    order_out[cnt++] = (uint32_t)i;
  }
  free(visited);
  free(heap.data);
  state_free(&s);
  return rc ? rc : cnt;
}

/* BFS traversal of one basin (drainage_basin.rs:13-36).  Returns its length;
 * order_out must hold n entries. */
long flo_basin_order(size_t n, const uint32_t *row_ptr, const uint32_t *col, const double *len,
                     const uint32_t *next_in, uint32_t outlet, uint32_t *order_out) {
  flo_graph g;
  int rc = graph_from_csr(&g, n, row_ptr, col, len);
  size_t *next = (size_t *)malloc((n ? n : 1) * sizeof(size_t));
  svec trav = {0};
  long out = rc;
  if (!rc && next) {
    for (size_t i = 0; i < n; i++) next[i] = next_in[i];
    rc = drainage_basin_construct(&g, next, outlet, &trav);
    if (!rc) {
      for (size_t t = 0; t < trav.len; t++) order_out[t] = (uint32_t)trav.data[t];
      out = (long)trav.len;
    } else {
      out = rc;
    }
  }
  free(trav.data);
  free(next);
  graph_free(&g);
  return out;
}

/* One full outer iteration (generator.rs:141-204) from elev (updated in
 * place), exposing next, pre-join subroot, drainage area and response time.
 * Returns changed (0/1) or an error; *had_lake receives has_lake. */
int flo_iteration(size_t n, const uint32_t *row_ptr, const uint32_t *col, const double *len, const double *area,
                  const double *erod, const double *uplift, const uint8_t *has_max_slope, const double *max_slope,
                  const uint32_t *outlets, size_t n_outlets, double *elev, uint32_t *next_out,
                  uint32_t *subroot_out, double *da_out, double *rt_out, int *had_lake) {
  flo_state s;
  int rc = state_init(&s, n, row_ptr, col, len, outlets, n_outlets);
  if (!rc) {
    rc = one_iteration(&s.g, area, erod, uplift, has_max_slope, max_slope, s.outlets, s.n_outlets, s.is_outlet,
                       elev, s.next, s.subroot, s.da, s.rt, had_lake);
    if (rc >= 0) {
      if (next_out) copy_out_idx(s.next, next_out, n);
      if (subroot_out) copy_out_idx(s.subroot, subroot_out, n);
      if (da_out) memcpy(da_out, s.da, n * sizeof(double));
      if (rt_out) memcpy(rt_out, s.rt, n * sizeof(double));
    }
  }
  state_free(&s);
  return rc;
}

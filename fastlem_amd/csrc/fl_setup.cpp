// fl_setup.cpp -- one-time host setup for the uplift-erosion solve.
//
// Everything here runs once per model (or once per generate() call for the
// jitter) on the host; none of it is inside the iteration loop:
//   * a restatement of rand 0.8.5's StdRng (ChaCha12, rand_chacha 0.3) with
//     SeedableRng::seed_from_u64 / from_seed, Rng::gen::<f64>() and
//     Rng::gen_range(f64..f64), used by
//       src/models/surface/builder.rs:38-52 (from_random_sites) and
//       src/lem/generator.rs:134-138 (elevation jitter);
//   * clipped Voronoi cell areas and centroids (the cell areas of
//     src/models/surface/builder.rs:236-247 and the Lloyd step of :155-200);
//   * the adjacency build of src/models/surface/builder.rs:249-268 (edge
//     (a,b) added only when a<b, insertion-ordered rows);
//   * the static visit order of the lake-removal heap
//     (src/lem/stream_tree.rs:184-240), which depends only on the graph and
//     the outlet order.
//
// Built with -ffp-contract=off: every floating-point expression is evaluated
// as written, as rustc does.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "fl_internal.h"

// ------------------------------------------------------------ ChaCha12 ----

namespace {

inline uint32_t rotl32(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }

inline void quarter(uint32_t* x, int a, int b, int c, int d) {
  x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl32(x[d], 16);
  x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl32(x[b], 12);
  x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl32(x[d], 8);
  x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl32(x[b], 7);
}

}  // namespace

// One ChaCha block: `rounds` rounds (12 for StdRng, 20 for RFC 8439) over the
// 16-word input state; out = permuted state + input state.
extern "C" void fl_chacha_block(int rounds, const uint32_t in[16], uint32_t out[16]) {
  uint32_t x[16];
  std::memcpy(x, in, sizeof(x));
  for (int r = 0; r < rounds; r += 2) {
    quarter(x, 0, 4, 8, 12);
    quarter(x, 1, 5, 9, 13);
    quarter(x, 2, 6, 10, 14);
    quarter(x, 3, 7, 11, 15);
    quarter(x, 0, 5, 10, 15);
    quarter(x, 1, 6, 11, 12);
    quarter(x, 2, 7, 8, 13);
    quarter(x, 3, 4, 9, 14);
  }
  for (int i = 0; i < 16; i++) out[i] = x[i] + in[i];
}

struct fl_rng {
  uint32_t key[8];
  uint64_t counter;   // next block counter
  uint32_t buf[64];   // four blocks, as rand_chacha buffers them
  uint32_t index;     // next word in buf (64 = empty)
};

static void rng_refill(fl_rng* r) {
  for (int b = 0; b < 4; b++) {
    uint32_t st[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
    for (int k = 0; k < 8; k++) st[4 + k] = r->key[k];
    st[12] = (uint32_t)(r->counter & 0xffffffffu);
    st[13] = (uint32_t)(r->counter >> 32);
    st[14] = 0;  // stream id 0
    st[15] = 0;
    fl_chacha_block(12, st, r->buf + 16 * b);
    r->counter++;
  }
  r->index = 0;
}

extern "C" void fl_rng_from_seed(fl_rng* r, const uint8_t seed[32]) {
  for (int k = 0; k < 8; k++)
    r->key[k] = (uint32_t)seed[4 * k] | ((uint32_t)seed[4 * k + 1] << 8) |
                ((uint32_t)seed[4 * k + 2] << 16) | ((uint32_t)seed[4 * k + 3] << 24);
  r->counter = 0;
  r->index = 64;
}

// rand_core 0.6 SeedableRng::seed_from_u64: PCG32 output words fill the seed.
extern "C" void fl_rng_seed_from_u64(fl_rng* r, uint64_t state) {
  const uint64_t MUL = 6364136223846793005ull;
  const uint64_t INC = 11634580027462260723ull;
  uint8_t seed[32];
  for (int c = 0; c < 8; c++) {
    state = state * MUL + INC;
    uint32_t xorshifted = (uint32_t)(((state >> 18) ^ state) >> 27);
    uint32_t rot = (uint32_t)(state >> 59);
    uint32_t x = (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
    seed[4 * c + 0] = (uint8_t)(x);
    seed[4 * c + 1] = (uint8_t)(x >> 8);
    seed[4 * c + 2] = (uint8_t)(x >> 16);
    seed[4 * c + 3] = (uint8_t)(x >> 24);
  }
  fl_rng_from_seed(r, seed);
}

extern "C" uint64_t fl_rng_next_u64(fl_rng* r) {
  // BlockRng::next_u64 with only 64-bit draws: the index stays even.
  if (r->index >= 64) rng_refill(r);
  uint64_t lo = r->buf[r->index], hi = r->buf[r->index + 1];
  r->index += 2;
  return (hi << 32) | lo;
}

// Standard distribution for f64: 53 high bits times 2^-53.
extern "C" double fl_rng_gen_f64(fl_rng* r) {
  uint64_t v = fl_rng_next_u64(r) >> 11;
  return (double)v * (1.0 / 9007199254740992.0);
}

// UniformFloat::<f64>::sample_single(low, high): [1,2) mantissa trick,
// value0_1 * scale + low, rejected while not < high.
extern "C" double fl_rng_gen_range_f64(fl_rng* r, double low, double high) {
  double scale = high - low;
  for (;;) {
    uint64_t bits = (fl_rng_next_u64(r) >> 12) | 0x3ff0000000000000ull;
    double v12;
    std::memcpy(&v12, &bits, sizeof(v12));
    double v01 = v12 - 1.0;
    double res = v01 * scale + low;
    if (res < high) return res;
  }
}

extern "C" size_t fl_rng_state_size(void) { return sizeof(fl_rng); }

// Fill out[i] = base[i] + gen::<f64>() * f64::EPSILON with a StdRng seeded by
// seed_from_u64(seed) (generator.rs:134-138 uses seed 0).
extern "C" void fl_jitter_elevations(uint64_t seed, size_t n, const double* base, double* out) {
  fl_rng r;
  fl_rng_seed_from_u64(&r, seed);
  const double eps = 2.220446049250313e-16;  // f64::EPSILON
  for (size_t i = 0; i < n; i++) out[i] = base[i] + fl_rng_gen_f64(&r) * eps;
}

// n sites drawn as from_random_sites(n, min, max): StdRng::from_seed([0;32]),
// x then y per site (builder.rs:38-52).
extern "C" void fl_random_sites(size_t n, double xmin, double ymin, double xmax, double ymax, double* xs,
                                double* ys) {
  fl_rng r;
  uint8_t seed[32] = {0};
  fl_rng_from_seed(&r, seed);
  for (size_t i = 0; i < n; i++) {
    xs[i] = fl_rng_gen_range_f64(&r, xmin, xmax);
    ys[i] = fl_rng_gen_range_f64(&r, ymin, ymax);
  }
}

// ------------------------------------------------------- Voronoi cells ----

namespace {

struct P2 {
  double x, y;
};

// Keep the part of polygon `in` with (p - m) . nrm <= 0.
void clip_halfplane(const std::vector<P2>& in, std::vector<P2>& out, P2 m, P2 nrm) {
  out.clear();
  size_t k = in.size();
  for (size_t a = 0; a < k; a++) {
    const P2& p = in[a];
    const P2& q = in[(a + 1) % k];
    double fp = (p.x - m.x) * nrm.x + (p.y - m.y) * nrm.y;
    double fq = (q.x - m.x) * nrm.x + (q.y - m.y) * nrm.y;
    if (fp <= 0) out.push_back(p);
    if ((fp < 0 && fq > 0) || (fp > 0 && fq < 0)) {
      double t = fp / (fp - fq);
      out.push_back(P2{p.x + (q.x - p.x) * t, p.y + (q.y - p.y) * t});
    }
  }
}

}  // namespace

// Voronoi cell of every site clipped to the box [xmin,xmax]x[ymin,ymax]:
// cell i = box intersected with the bisector half-planes of i's Delaunay
// neighbours (tris: ntri triangles of site indices).  Writes the shoelace
// area and the polygon centroid (the Lloyd target).
extern "C" int fl_voronoi_cells(size_t n, const double* xs, const double* ys, size_t ntri, const int64_t* tris,
                                double xmin, double ymin, double xmax, double ymax, double* area_out,
                                double* cx_out, double* cy_out) {
  std::vector<uint32_t> deg(n + 1, 0);
  for (size_t t = 0; t < ntri; t++)
    for (int e = 0; e < 3; e++) {
      int64_t a = tris[3 * t + e];
      if (a < 0 || (size_t)a >= n) return FL_ERR_INVALID;
      deg[a] += 2;
    }
  std::vector<uint64_t> off(n + 1, 0);
  for (size_t i = 0; i < n; i++) off[i + 1] = off[i] + deg[i];
  std::vector<uint32_t> nb(off[n]);
  std::vector<uint64_t> fill(off.begin(), off.end() - 1);
  for (size_t t = 0; t < ntri; t++)
    for (int e = 0; e < 3; e++) {
      int64_t a = tris[3 * t + e], b = tris[3 * t + (e + 1) % 3], c = tris[3 * t + (e + 2) % 3];
      nb[fill[a]++] = (uint32_t)b;
      nb[fill[a]++] = (uint32_t)c;
    }
  std::vector<P2> poly, tmp;
  for (size_t i = 0; i < n; i++) {
    poly.assign({P2{xmin, ymin}, P2{xmax, ymin}, P2{xmax, ymax}, P2{xmin, ymax}});
    P2 s{xs[i], ys[i]};
    for (uint64_t k = off[i]; k < off[i + 1] && !poly.empty(); k++) {
      uint32_t j = nb[k];
      P2 o{xs[j], ys[j]};
      P2 m{(s.x + o.x) * 0.5, (s.y + o.y) * 0.5};
      P2 nrm{o.x - s.x, o.y - s.y};
      clip_halfplane(poly, tmp, m, nrm);
      poly.swap(tmp);
    }
    double a2 = 0, cx = 0, cy = 0;
    size_t k = poly.size();
    for (size_t a = 0; a < k; a++) {
      const P2& p = poly[a];
      const P2& q = poly[(a + 1) % k];
      double cr = p.x * q.y - q.x * p.y;
      a2 += cr;
      cx += (p.x + q.x) * cr;
      cy += (p.y + q.y) * cr;
    }
    area_out[i] = std::fabs(a2) / 2.0;
    if (a2 != 0) {
      cx_out[i] = cx / (3.0 * a2);
      cy_out[i] = cy / (3.0 * a2);
    } else {
      cx_out[i] = s.x;
      cy_out[i] = s.y;
    }
  }
  return FL_OK;
}

// ------------------------------------------------------------ adjacency ----

// builder.rs:249-268: for each triangle (a,b,c) in order add (a,b) if a<b,
// (b,c) if b<c, (c,a) if c<a; add_edge appends to both rows, so rows keep
// insertion order.  Edge attribute = Site2D::distance (sites.rs:27-29).
// row_ptr must hold n+1 entries; col/len at least 6*ntri... call with
// col == NULL first to get the entry count in row_ptr[n].
extern "C" int fl_build_adjacency(size_t n, const double* xs, const double* ys, size_t ntri, const int64_t* tris,
                                  uint32_t* row_ptr, uint32_t* col, double* len) {
  std::vector<uint64_t> deg(n, 0);
  auto each_edge = [&](auto&& fn) {
    for (size_t t = 0; t < ntri; t++) {
      int64_t a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
      if (a < b) fn(a, b);
      if (b < c) fn(b, c);
      if (c < a) fn(c, a);
    }
  };
  for (size_t t = 0; t < 3 * ntri; t++)
    if (tris[t] < 0 || (size_t)tris[t] >= n) return FL_ERR_INVALID;
  each_edge([&](int64_t a, int64_t b) {
    deg[a]++;
    deg[b]++;
  });
  uint64_t total = 0;
  row_ptr[0] = 0;
  for (size_t i = 0; i < n; i++) {
    total += deg[i];
    if (total > 0xffffffffull) return FL_ERR_TOO_LARGE;
    row_ptr[i + 1] = (uint32_t)total;
  }
  if (!col) return FL_OK;
  std::vector<uint32_t> fill(row_ptr, row_ptr + n);
  each_edge([&](int64_t a, int64_t b) {
    double dx = xs[a] - xs[b], dy = ys[a] - ys[b];
    double d = std::sqrt(dx * dx + dy * dy);
    col[fill[a]] = (uint32_t)b;
    len[fill[a]++] = d;
    col[fill[b]] = (uint32_t)a;
    len[fill[b]++] = d;
  });
  return FL_OK;
}

// ---------------------------------------------------- lake visit order ----

namespace {

struct Ridge {
  uint32_t index;
  double dist;
};

// RidgeElement ordering (stream_tree.rs:34-38): a <= b iff a.dist >= b.dist,
// so the max-heap yields the shortest edge first.
inline bool le(const Ridge& a, const Ridge& b) { return a.dist >= b.dist; }

// std::collections::BinaryHeap mechanics: push = append + sift_up (stop when
// elem <= parent); pop = take root, move last to root, sift_down_to_bottom
// (descend to the larger child, right on ties) then sift_up.
struct RustHeap {
  std::vector<Ridge> d;
  size_t sift_up(size_t start, size_t pos) {
    Ridge e = d[pos];
    while (pos > start) {
      size_t parent = (pos - 1) / 2;
      if (le(e, d[parent])) break;
      d[pos] = d[parent];
      pos = parent;
    }
    d[pos] = e;
    return pos;
  }
  void push(Ridge x) {
    d.push_back(x);
    sift_up(0, d.size() - 1);
  }
  bool pop(Ridge* out) {
    if (d.empty()) return false;
    Ridge item = d.back();
    d.pop_back();
    if (!d.empty()) {
      std::swap(item, d[0]);
      size_t end = d.size(), pos = 0;
      Ridge e = d[0];
      size_t child = 1;
      while (end >= 2 && child <= end - 2) {
        child += le(d[child], d[child + 1]) ? 1 : 0;
        d[pos] = d[child];
        pos = child;
        child = 2 * pos + 1;
      }
      if (child == end - 1) {
        d[pos] = d[child];
        pos = child;
      }
      d[pos] = e;
      sift_up(0, pos);
    }
    *out = item;
    return true;
  }
};

}  // namespace

// Exact simulation of the traversal for the terrain occupying sites
// [s0, s0 + nl); its ranks start at rank_base.
int fl_lake_visit_order_range(uint32_t s0, uint32_t nl, const uint32_t* row_ptr, const uint32_t* col,
                              const double* len, const uint32_t* outlets, size_t k, uint32_t rank_base,
                              uint32_t* rank_out) {
  RustHeap h;
  h.d.reserve(nl);
  std::vector<uint8_t> visited(nl, 0);
  for (uint32_t i = 0; i < nl; i++) rank_out[s0 + i] = UINT32_MAX;
  for (size_t q = 0; q < k; q++) {
    if (outlets[q] < s0 || outlets[q] - s0 >= nl) return FL_ERR_INVALID;
    h.push(Ridge{outlets[q], 0.0});
  }
  uint32_t cnt = 0;
  Ridge el;
  while (h.pop(&el)) {
    uint32_t i = el.index;
    if (visited[i - s0]) continue;
    for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; e++) {
      uint32_t j = col[e];
      if (visited[j - s0]) continue;
      if (std::isnan(len[e])) return FL_ERR_NAN_LENGTH;
      h.push(Ridge{j, len[e]});
    }
    visited[i - s0] = 1;
    rank_out[i] = rank_base + cnt++;
  }
  return FL_OK;
}

// Visit order of the lake-removal traversal: sites in the order in which the
// heap first pops them.  rank_out[i] = position of i in that order, or
// UINT32_MAX when i is never reached (no path to an outlet).  The flips the
// reference performs during the traversal never change what is pushed, so
// this order is a function of (graph, outlet order) alone.
extern "C" int fl_lake_visit_order(size_t n, const uint32_t* row_ptr, const uint32_t* col, const double* len,
                                   const uint32_t* outlets, size_t n_outlets, uint32_t* rank_out,
                                   uint64_t* visited_out) {
  int rc = fl_lake_visit_order_range(0, (uint32_t)n, row_ptr, col, len, outlets, n_outlets, 0, rank_out);
  if (rc != FL_OK) return rc;
  if (visited_out) {
    uint64_t cnt = 0;
    for (size_t i = 0; i < n; i++) cnt += rank_out[i] != UINT32_MAX;
    *visited_out = cnt;
  }
  return FL_OK;
}

// ------------------------------------------- visit order, fast form ----
//
// The lake traversal (stream_tree.rs:184-240) first pops every outlet (all
// pushed with key 0.0, so they share the maximum priority), then behaves as
// Prim's algorithm from the outlet set.  With positive, pairwise distinct
// edge lengths the second phase has a unique order, computable from a
// minimum spanning forest; the first phase depends on the heap's handling of
// equal keys, which is simulated exactly below.

// Outlet phase.  Elements are either Z (key 0: an outlet) or P (positive
// key: a pushed edge).  Z outranks P, so Z elements occupy a root-connected
// set of heap positions; pushes append P elements that never enter it, and a
// pop shifts the Z elements on the sift-down path up by one level (on Z|Z
// the path takes the right child, on Z|P the Z child, and it leaves the Z
// set at a P|P pair).  The element moved from the end, if it is Z, settles in
// the vacated lowest Z position of that path.  None of this depends on the
// values of P elements, so only their number (the unvisited neighbours of
// each popped outlet) is needed.
//
// While the end element is P ("pure" pops) the process has a closed form: a
// pop outputs the root and refills it from the right subtree while that
// holds Z elements, then from the left one, so the outputs follow the
// right-first preorder of the Z set, and the position a pop vacates (the end
// of its path) is the first remaining one in right-first postorder.  One
// depth-first walk therefore yields every pure pop; the walk is replayed
// until a pop would find a Z element at the heap's end (the count of pushes
// decides where the end is), and from the state there -- positions vacated
// so far freed, the remaining values laid out in preorder -- the pops are
// simulated one by one again until the end element is P.  After 16 such
// returns the simulation goes on pop by pop.
namespace {

struct OutletHeap {
  static constexpr uint32_t NONE = 0xffffffffu;
  uint64_t k = 0;
  std::vector<uint32_t> z;     // Z element at each heap position, NONE elsewhere (2k + 3 entries)
  std::vector<uint32_t> val;   // walk: Z elements in right-first preorder
  std::vector<uint32_t> vac;   // walk: right-first postorder index of each position
  std::vector<uint32_t> post;  // walk: positions in right-first postorder
  uint32_t pc = 0, qc = 0;

  // Heap positions hold at most k Z elements, all in [0, k); the depth of the
  // recursion is the heap's, at most 64.
  void walk(uint64_t h) {
    val[pc++] = z[h];
    if (z[2 * h + 2] != NONE) walk(2 * h + 2);
    if (z[2 * h + 1] != NONE) walk(2 * h + 1);
    vac[h] = qc;
    post[qc++] = (uint32_t)h;
  }
  void place(uint64_t h, const uint32_t* v) {
    z[h] = v[pc++];
    if (z[2 * h + 2] != NONE) place(2 * h + 2, v);
    if (z[2 * h + 1] != NONE) place(2 * h + 1, v);
  }
  // One pop of a heap of `len` elements (len updated); returns the popped Z
  // element.
  uint32_t pop(uint64_t& len) {
    const uint32_t result = z[0];
    const uint64_t last = len - 1;
    uint32_t last_z = NONE;
    if (last < k) {
      last_z = z[last];
      z[last] = NONE;
    }
    len = last;
    uint64_t h = 0;
    for (;;) {
      const uint32_t a = z[2 * h + 1], b = z[2 * h + 2];
      if ((a & b) == NONE) break;  // neither child holds a Z element
      z[h] = b != NONE ? b : a;
      h = b != NONE ? 2 * h + 2 : 2 * h + 1;
    }
    z[h] = last_z;
    return result;
  }
};

// The simulation over k outlets named ids[0..k) in push order (their
// listed order).  pushes(id) returns how many entries the popped outlet id
// pushes (its neighbours not popped before it) and records it as popped;
// ahead(id) fetches its adjacency early.
template <class Pushes, class Ahead>
int outlet_pop_core(const uint32_t* ids, size_t k, Pushes pushes, Ahead ahead, uint32_t* order_out) {
  const uint32_t NONE = OutletHeap::NONE;
  OutletHeap H;
  H.k = k;
  H.z.assign(2 * k + 3, NONE);
  std::copy(ids, ids + k, H.z.begin());
  uint64_t len = k;
  size_t t = 0;
  int returns = 0;
  auto one_pop = [&]() -> bool {
    for (uint64_t q = 1; q <= 2; q++)
      if (H.z[q] != NONE) ahead(H.z[q]);
    const uint32_t r = H.pop(len);
    if (r == NONE) return false;  // cannot happen with distinct outlets
    order_out[t++] = r;
    len += pushes(r);
    return true;
  };
  while (t < k) {
    // a Z element at the end moves into the path: simulate
    while (t < k && len - 1 < k && H.z[len - 1] != NONE)
      if (!one_pop()) return FL_ERR_INVALID;
    if (t == k) break;
    if (returns == 16) {
      while (t < k)
        if (!one_pop()) return FL_ERR_INVALID;
      break;
    }
    returns++;
    const size_t m = k - t;
    H.val.resize(m);
    H.post.resize(m);
    H.vac.resize(k);
    H.pc = H.qc = 0;
    if (H.z[0] == NONE) return FL_ERR_INVALID;
    H.walk(0);
    if (H.pc != m) return FL_ERR_INVALID;
    size_t r = 0;
    for (; r < m; r++) {
      const uint64_t last = len - 1;
      if (last < k && H.z[last] != NONE && H.vac[last] >= r) break;  // still a Z element there
      if (r + 4 < m) ahead(H.val[r + 4]);
      const uint32_t res = H.val[r];
      order_out[t + r] = res;
      len = last + pushes(res);
    }
    if (r == m) break;
    for (size_t q = 0; q < r; q++) H.z[H.post[q]] = NONE;
    H.pc = 0;
    H.place(0, H.val.data() + r);
    t += r;
  }
  return FL_OK;
}

}  // namespace

int fl_outlet_pop_order_range(uint32_t s0, uint32_t nl, const uint32_t* row_ptr, const uint32_t* col,
                              const uint32_t* outlets, size_t k, uint32_t* order_out) {
  for (size_t q = 0; q < k; q++)
    if (outlets[q] < s0 || outlets[q] - s0 >= nl) return FL_ERR_INVALID;
  std::vector<uint8_t> visited(nl, 0);
  return outlet_pop_core(
      outlets, k,
      [&](uint32_t v) {
        visited[v - s0] = 1;
        uint64_t pushes = 0;
        for (uint32_t e = row_ptr[v]; e < row_ptr[v + 1]; e++) pushes += visited[col[e] - s0] ? 0 : 1;
        return pushes;
      },
      [&](uint32_t v) { __builtin_prefetch(col + row_ptr[v]); },
      order_out);
}

int fl_outlet_pop_order_compact(size_t k, const uint32_t* deg, const uint32_t* nb_ptr, const uint32_t* nb,
                                uint32_t* order_out) {
  std::vector<uint32_t> ids(k);
  for (size_t q = 0; q < k; q++) ids[q] = (uint32_t)q;
  std::vector<uint8_t> popped(k, 0);
  return outlet_pop_core(
      ids.data(), k,
      [&](uint32_t q) {
        popped[q] = 1;
        uint64_t before = 0;
        for (uint32_t e = nb_ptr[q]; e < nb_ptr[q + 1]; e++) before += popped[nb[e]];
        return (uint64_t)deg[q] - before;
      },
      [&](uint32_t q) { __builtin_prefetch(nb + nb_ptr[q]); },
      order_out);
}

extern "C" int fl_outlet_pop_order(size_t n, const uint32_t* row_ptr, const uint32_t* col, const uint32_t* outlets,
                                   size_t k, uint32_t* order_out) {
  return fl_outlet_pop_order_range(0, (uint32_t)n, row_ptr, col, outlets, k, order_out);
}

// Prim phase from a minimum spanning forest of the graph in which each
// group's outlets are contracted into one vertex S_g (groups are independent
// terrains: no edge joins two groups).  msf_a/msf_b: the forest's edges
// sorted by increasing length.  Kruskal's merges build the reconstruction
// tree; Prim from S_g completes the side of each merge edge that contains
// the entry point before crossing to the other side, which it enters at the
// edge's endpoint there (msf_a lies on the first-merged side, msf_b on the
// other).  Group g's non-outlet sites reached from S_g get ranks
// rank_off[g] + k_g + (Prim order); outlets and unreached sites are left
// UINT32_MAX.  `outlets` may list each group's outlets in any order.
//
// The two top-down passes over the reconstruction tree run on `threads`
// threads: the subtrees holding more than `cut` sites (recorded while the
// tree is built) are visited by the calling thread in creation order
// reversed, the smaller subtrees below them by depth-first tasks.
namespace {

template <class F>
void parallel_range(size_t n, unsigned threads, F f) {
  if (threads <= 1 || n < 65536) {
    f((size_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  for (unsigned t = 0; t < threads; t++)
    th.emplace_back([&, t] { f(n * t / threads, n * (t + 1) / threads); });
  for (auto& x : th) x.join();
}

}  // namespace

namespace {

const uint32_t kNone = 0xffffffffu;
struct KrtNode {
  uint32_t ch0, ch1;
  uint32_t cnt;   // leaves below
  uint32_t size;  // non-outlet leaves below
};
struct KrtDown {
  uint32_t lo;     // first leaf position of the subtree
  uint32_t entry;  // site where Prim enters the subtree
  uint32_t off;    // rank of the first site Prim reaches in it
};
// union-find root record: parent, the component's current reconstruction
// node and its counts (kept at the root, so a merge reads no child node)
struct KrtRoot {
  uint32_t parent, top, cnt, size;
};

}  // namespace

// The ranks' working arrays for n sites (reconstruction-tree nodes for up to
// n forest edges), set up from the outlets alone.
struct fl_krt_work {
  size_t n = 0;
  std::vector<uint32_t> leaf;
  std::unique_ptr<KrtRoot[]> uf;
  std::unique_ptr<KrtNode[]> node;
  std::unique_ptr<KrtDown[]> down;
};

fl_krt_work* fl_krt_prepare(size_t n, size_t groups, const uint32_t* outlet_off, const uint32_t* outlets,
                            unsigned threads, int* rc) {
  if (threads == 0) threads = 1;
  std::unique_ptr<fl_krt_work> w(new fl_krt_work());
  w->n = n;
  w->leaf.resize(n);
  w->uf.reset(new KrtRoot[n]);
  w->node.reset(new KrtNode[2 * n]);
  w->down.reset(new KrtDown[2 * n]);
  uint32_t* leaf = w->leaf.data();
  KrtRoot* uf = w->uf.get();
  KrtNode* node = w->node.get();
  KrtDown* down = w->down.get();
  parallel_range(n, threads, [&](size_t a, size_t b) {
    for (size_t v = a; v < b; v++) leaf[v] = (uint32_t)v;
  });
  for (size_t g = 0; g < groups; g++)
    for (uint32_t q = outlet_off[g]; q < outlet_off[g + 1]; q++) {
      if (outlets[q] >= n) {
        *rc = FL_ERR_INVALID;
        return nullptr;
      }
      leaf[outlets[q]] = outlets[outlet_off[g]];
    }
  parallel_range(n, threads, [&](size_t a, size_t b) {
    for (size_t v = a; v < b; v++) {
      uf[v] = {leaf[v], (uint32_t)v, 1u, 1u};
      node[v] = {kNone, kNone, 1u, 1u};
      down[v] = {kNone, kNone, 0};
      down[n + v] = {kNone, kNone, 0};
    }
  });
  for (size_t g = 0; g < groups; g++)
    for (uint32_t q = outlet_off[g]; q < outlet_off[g + 1]; q++) {
      node[outlets[q]].size = 0;
      uf[outlets[q]].size = 0;
    }
  *rc = FL_OK;
  return w.release();
}

void fl_krt_release(fl_krt_work* w) { delete w; }

namespace {

// Ranks of one forest on the arrays of `work` (prepared for the same sites
// and outlets; consumed).
int krt_ranks_run(fl_krt_work* work, size_t groups, const uint32_t* outlet_off, const uint32_t* outlets,
                  const uint32_t* rank_off, const uint32_t* msf_a, const uint32_t* msf_b, size_t m,
                  uint32_t* rank_out, unsigned threads) {
  const uint32_t NONE = kNone;
  const size_t n = work->n;
  if (m > n) return FL_ERR_INVALID;  // a forest on n sites has fewer than n edges
  const size_t total = n + m;
  if (threads == 0) threads = 1;
  using Node = KrtNode;
  using Down = KrtDown;
  using Root = KrtRoot;
  const uint32_t* leaf = work->leaf.data();
  Root* uf = work->uf.get();
  Node* node = work->node.get();
  Down* down = work->down.get();
  parallel_range(n, threads, [&](size_t a, size_t b) {
    for (size_t v = a; v < b; v++) rank_out[v] = NONE;
  });
  auto find = [&](uint32_t v) {
    while (uf[v].parent != v) {
      uf[v].parent = uf[uf[v].parent].parent;
      v = uf[v].parent;
    }
    return v;
  };
  const uint32_t cut = (uint32_t)std::max<size_t>(4096, total / (64 * (size_t)threads));
  std::vector<uint32_t> big;  // reconstruction nodes with more than `cut` leaves, ascending
  // the union-find records of the edges a few steps ahead are fetched early,
  // and half as far ahead their parents' records: the loop is bound by their
  // cache misses
  constexpr size_t kAhead = 32;
  for (size_t i = 0; i < m; i++) {
    uint32_t a = msf_a[i], b = msf_b[i];
    if (a >= n || b >= n) return FL_ERR_INVALID;
    if (i + kAhead < m) {
      const uint32_t pa = msf_a[i + kAhead], pb = msf_b[i + kAhead];
      if (pa < n) __builtin_prefetch(&uf[pa]);
      if (pb < n) __builtin_prefetch(&uf[pb]);
    }
    if (i + kAhead / 2 < m) {
      const uint32_t qa = msf_a[i + kAhead / 2], qb = msf_b[i + kAhead / 2];
      if (qa < n) __builtin_prefetch(&uf[uf[qa].parent]);
      if (qb < n) __builtin_prefetch(&uf[uf[qb].parent]);
    }
    uint32_t ra = find(a), rb = find(b);
    if (ra == rb) return FL_ERR_INVALID;  // not a forest
    Root& A = uf[ra];
    Root& B = uf[rb];
    Node& x = node[n + i];
    x.ch0 = A.top;  // the side of msf_a first
    x.ch1 = B.top;
    x.cnt = A.cnt + B.cnt;
    x.size = A.size + B.size;
    Root& keep = A.cnt >= B.cnt ? A : B;  // union by size
    (A.cnt >= B.cnt ? B : A).parent = A.cnt >= B.cnt ? ra : rb;
    keep.top = (uint32_t)(n + i);
    keep.cnt = x.cnt;
    keep.size = x.size;
    if (x.cnt > cut) big.push_back((uint32_t)(n + i));
  }
  std::vector<uint32_t> roots;
  for (size_t g = 0; g < groups; g++) {
    if (outlet_off[g + 1] == outlet_off[g]) continue;
    uint32_t s_leaf = outlets[outlet_off[g]];
    uint32_t root = uf[find(s_leaf)].top;
    down[root] = {0, s_leaf, rank_off[g] + (outlet_off[g + 1] - outlet_off[g])};
    roots.push_back(root);
  }
  // one top-down pass: `visit(x)` fills the children of reached node x
  auto pass = [&](auto&& visit, bool by_lo) {
    auto reached = [&](uint32_t x) { return (by_lo ? down[x].lo : down[x].entry) != NONE; };
    std::vector<uint32_t> tasks;
    for (size_t q = big.size(); q-- > 0;) {
      uint32_t x = big[q];
      // the sequential walk down the large nodes: records 16 and 8 nodes
      // ahead (and their children's) are fetched early
      if (q >= 16) {
        const uint32_t y = big[q - 16];
        __builtin_prefetch(&node[y]);
        __builtin_prefetch(&down[y]);
      }
      if (q >= 8) {
        const uint32_t y = big[q - 8];
        const Node& Y = node[y];
        __builtin_prefetch(&node[Y.ch0]);
        __builtin_prefetch(&down[Y.ch0]);
        __builtin_prefetch(&node[Y.ch1]);
        __builtin_prefetch(&down[Y.ch1]);
        __builtin_prefetch(&leaf[msf_a[y - n]]);
        __builtin_prefetch(&leaf[msf_b[y - n]]);
      }
      if (!reached(x)) continue;
      visit(x);
      for (uint32_t c : {node[x].ch0, node[x].ch1})
        if (c >= n && node[c].cnt <= cut) tasks.push_back(c);
    }
    for (uint32_t r : roots)
      if (r >= n && node[r].cnt <= cut) tasks.push_back(r);
    std::atomic<size_t> next{0};
    auto worker = [&] {
      std::vector<uint32_t> st;
      for (;;) {
        size_t q = next.fetch_add(1);
        if (q >= tasks.size()) break;
        st.push_back(tasks[q]);
        while (!st.empty()) {
          uint32_t x = st.back();
          st.pop_back();
          visit(x);
          if (node[x].ch0 >= n) st.push_back(node[x].ch0);
          if (node[x].ch1 >= n) st.push_back(node[x].ch1);
        }
      }
    };
    if (threads <= 1 || tasks.size() < 2) {
      worker();
      return;
    }
    std::vector<std::thread> th;
    for (unsigned t = 0; t < threads; t++) th.emplace_back(worker);
    for (auto& x : th) x.join();
  };
  // leaf intervals (ch0's leaves first)
  pass(
      [&](uint32_t x) {
        const Node& X = node[x];
        down[X.ch0].lo = down[x].lo;
        down[X.ch1].lo = down[x].lo + node[X.ch0].cnt;
      },
      true);
  // entry points and first ranks
  pass(
      [&](uint32_t x) {
        const Down D = down[x];
        const Node& X = node[x];
        size_t i = x - n;
        uint32_t l = down[leaf[D.entry]].lo;
        const Down& C0 = down[X.ch0];
        bool in0 = l >= C0.lo && l - C0.lo < node[X.ch0].cnt;
        uint32_t first = in0 ? X.ch0 : X.ch1, second = in0 ? X.ch1 : X.ch0;
        uint32_t enter = in0 ? msf_b[i] : msf_a[i];
        down[first].entry = D.entry;
        down[first].off = D.off;
        down[second].entry = leaf[enter];
        down[second].off = D.off + node[first].size;
      },
      false);
  parallel_range(n, threads, [&](size_t a, size_t b) {
    for (size_t v = a; v < b; v++)
      if (node[v].size && down[v].entry != NONE) rank_out[v] = down[v].off;
  });
  return FL_OK;
}

int krt_ranks_one(size_t n, size_t groups, const uint32_t* outlet_off, const uint32_t* outlets,
                  const uint32_t* rank_off, const uint32_t* msf_a, const uint32_t* msf_b, size_t m,
                  uint32_t* rank_out, unsigned threads) {
  int rc = FL_OK;
  std::unique_ptr<fl_krt_work> work(fl_krt_prepare(n, groups, outlet_off, outlets, threads, &rc));
  if (!work) return rc;
  return krt_ranks_run(work.get(), groups, outlet_off, outlets, rank_off, msf_a, msf_b, m, rank_out, threads);
}

}  // namespace

int fl_krt_ranks_prepared(fl_krt_work* work, size_t groups, const uint32_t* outlet_off, const uint32_t* outlets,
                          const uint32_t* rank_off, const uint32_t* msf_a, const uint32_t* msf_b, size_t m,
                          uint32_t* rank_out, unsigned threads) {
  return krt_ranks_run(work, groups, outlet_off, outlets, rank_off, msf_a, msf_b, m, rank_out, threads);
}

int fl_krt_ranks(size_t n, size_t groups, const uint32_t* outlet_off, const uint32_t* outlets,
                 const uint32_t* rank_off, const uint32_t* site_off, const uint32_t* msf_a, const uint32_t* msf_b,
                 size_t m, uint32_t* rank_out, unsigned threads) {
  if (groups <= 1 || !site_off)
    return krt_ranks_one(n, groups, outlet_off, outlets, rank_off, msf_a, msf_b, m, rank_out, threads);
  // independent groups (terrains of a batch): each group's forest edges, in
  // their sorted order, on a thread of its own, the largest groups first
  std::vector<uint32_t> eoff(groups + 1, 0);
  std::vector<uint32_t> group_of_edge(m);
  for (size_t i = 0; i < m; i++) {
    uint32_t a = msf_a[i];
    if (a >= n) return FL_ERR_INVALID;
    size_t g = (size_t)(std::upper_bound(site_off, site_off + groups + 1, a) - site_off) - 1;
    group_of_edge[i] = (uint32_t)g;
    eoff[g + 1]++;
  }
  for (size_t g = 0; g < groups; g++) eoff[g + 1] += eoff[g];
  std::vector<uint32_t> ea(m), eb(m), cur(eoff.begin(), eoff.end() - 1);
  for (size_t i = 0; i < m; i++) {
    uint32_t q = cur[group_of_edge[i]]++;
    ea[q] = msf_a[i];
    eb[q] = msf_b[i];
  }
  std::vector<uint32_t> order(groups);
  for (size_t g = 0; g < groups; g++) order[g] = (uint32_t)g;
  std::sort(order.begin(), order.end(),
            [&](uint32_t x, uint32_t y) { return site_off[x + 1] - site_off[x] > site_off[y + 1] - site_off[y]; });
  std::atomic<size_t> next{0};
  std::atomic<int> status{FL_OK};
  auto worker = [&] {
    for (;;) {
      size_t q = next.fetch_add(1);
      if (q >= groups) break;
      uint32_t g = order[q];
      uint32_t s0 = site_off[g], ng = site_off[g + 1] - s0;
      // the group's sites relabelled to [0, ng)
      std::vector<uint32_t> la(eoff[g + 1] - eoff[g]), lb(la.size()), lo_out(outlet_off[g + 1] - outlet_off[g]);
      for (size_t i = 0; i < la.size(); i++) {
        la[i] = ea[eoff[g] + i] - s0;
        lb[i] = eb[eoff[g] + i] - s0;
        if (la[i] >= ng || lb[i] >= ng) {
          status = FL_ERR_INVALID;
          return;
        }
      }
      for (size_t i = 0; i < lo_out.size(); i++) lo_out[i] = outlets[outlet_off[g] + i] - s0;
      uint32_t ooff[2] = {0, (uint32_t)lo_out.size()}, roff[1] = {rank_off[g]};
      int rc = krt_ranks_one(ng, 1, ooff, lo_out.data(), roff, la.data(), lb.data(), la.size(), rank_out + s0, 1);
      if (rc != FL_OK) status = rc;
    }
  };
  if (threads <= 1) {
    worker();
  } else {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < threads; t++) th.emplace_back(worker);
    for (auto& x : th) x.join();
  }
  return status.load();
}

unsigned fl_host_threads() {
  unsigned h = std::thread::hardware_concurrency();
  return h == 0 ? 1 : (h > 16 ? 16 : h);
}

// As fl_krt_ranks, with the outlets ranked by their pop order (outlet_order,
// the order of fl_outlet_pop_order).
extern "C" int fl_ranks_from_msf_groups(size_t n, size_t groups, const uint32_t* outlet_off,
                                        const uint32_t* outlet_order, const uint32_t* rank_off,
                                        const uint32_t* msf_a, const uint32_t* msf_b, size_t m,
                                        uint32_t* rank_out) {
  int rc = fl_krt_ranks(n, groups, outlet_off, outlet_order, rank_off, nullptr, msf_a, msf_b, m, rank_out,
                        fl_host_threads());
  if (rc != FL_OK) return rc;
  for (size_t g = 0; g < groups; g++)
    for (uint32_t q = outlet_off[g]; q < outlet_off[g + 1]; q++)
      rank_out[outlet_order[q]] = rank_off[g] + (q - outlet_off[g]);
  return FL_OK;
}

// One group: all outlets contracted into one vertex.
extern "C" int fl_ranks_from_msf(size_t n, const uint32_t* outlet_order, size_t k, const uint32_t* msf_a,
                                 const uint32_t* msf_b, size_t m, uint32_t* rank_out) {
  uint32_t off[2] = {0, (uint32_t)k}, roff[1] = {0};
  return fl_ranks_from_msf_groups(n, 1, off, outlet_order, roff, msf_a, msf_b, m, rank_out);
}

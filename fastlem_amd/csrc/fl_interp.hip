// fl_interp.hip -- natural-neighbour (Sibson) interpolation of a solved
// terrain: Terrain2D::get_elevation (src/models/surface/terrain.rs:36-38)
// through TerrainInterpolator2D (src/models/surface/interpolator.rs:11-27),
// which builds the crate naturalneighbor 1.2.2's interpolator over the sites
// and returns None for a query outside their convex hull.
//
// Built once per terrain (host): the sites' Delaunay triangles oriented
// counter-clockwise, the neighbour across each triangle edge, and a uniform
// grid giving a triangle near every cell.  Queried in batches (device, one
// thread per query point q):
//   1. location: a visibility walk from the grid cell's triangle, crossing
//      an edge that has q strictly on its outer side until none does (it
//      ends on Delaunay triangulations); leaving through a hull edge means q
//      lies outside the hull: no value;
//   2. natural neighbours: the triangles whose circumcircle strictly holds q
//      (the Bowyer-Watson cavity of q, connected, grown from q's triangle);
//   3. Sibson weights by Watson's decomposition: for each cavity triangle T
//      with circumcentre c_T and each vertex a of T (then b, c following
//      it), the region q's cell takes from a's cell inside T's corner at a
//      is the triangle (c_T, cc(q, a, b), cc(q, c, a)), signed; a vertex's
//      weight is the sum of its regions over the cavity divided by their
//      total (the area of q's cell), and the value is sum(weight * value).
// A query equal to a site returns that site's value.  All arithmetic is
// FP64, in coordinates relative to q.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "fl_internal.h"

#define FL_NONE 0xffffffffu

namespace {

// cavity triangles per query held by one thread (a Delaunay cavity holds a
// handful; a larger one reports an error instead of a value)
constexpr uint32_t kMaxCavity = 48;
constexpr uint32_t kMaxWalk = 1u << 20;

struct InterpDev {
  uint32_t n, ntri;
  const double* x;
  const double* y;
  const uint32_t* tri;  // [3 ntri], counter-clockwise
  const uint32_t* nbr;  // [3 ntri]: across edge (v_i, v_{i+1}), FL_NONE on the hull
  const uint32_t* grid; // [gw gh]
  double gx0, gy0, gix, giy;  // grid origin and inverse cell sizes
  uint32_t gw, gh;
};

// twice the signed area of (a, b, c): > 0 counter-clockwise
__device__ __forceinline__ double orient(double ax, double ay, double bx, double by, double cx, double cy) {
  return (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
}

// Circumcentre of (0, a, b) (coordinates relative to q); d == 0 when the
// three points are collinear.
__device__ __forceinline__ void cc0(double ax, double ay, double bx, double by, double& ux, double& uy, double& d) {
  d = 2.0 * (ax * by - ay * bx);
  const double a2 = ax * ax + ay * ay, b2 = bx * bx + by * by;
  ux = (by * a2 - ay * b2) / d;
  uy = (ax * b2 - bx * a2) / d;
}

// Circumcentre of (a, b, c) (relative coordinates): translated to a.
__device__ __forceinline__ void cc3(double ax, double ay, double bx, double by, double cx, double cy, double& ux,
                                    double& uy, double& d) {
  cc0(bx - ax, by - ay, cx - ax, cy - ay, ux, uy, d);
  ux += ax;
  uy += ay;
}

// q strictly inside the circumcircle of counter-clockwise (a, b, c)
// (relative coordinates, q at the origin).
__device__ __forceinline__ bool in_circle(double ax, double ay, double bx, double by, double cx, double cy) {
  const double a2 = ax * ax + ay * ay, b2 = bx * bx + by * by, c2 = cx * cx + cy * cy;
  const double det = ax * (by * c2 - b2 * cy) - ay * (bx * c2 - b2 * cx) + a2 * (bx * cy - by * cx);
  return det > 0.0;
}

// Sibson value at (qx, qy) from the cavity; false when a new circumcentre
// is degenerate (q collinear with a cavity edge).
__device__ bool watson(const InterpDev& I, const double* __restrict__ val, const uint32_t* cav, uint32_t nc, double qx,
                       double qy, double& out) {
  double num = 0.0, den = 0.0;
  for (uint32_t k = 0; k < nc; ++k) {
    const uint32_t t = cav[k];
    uint32_t v[3] = {I.tri[3 * t], I.tri[3 * t + 1], I.tri[3 * t + 2]};
    double px[3], py[3];
    for (int i = 0; i < 3; ++i) {
      px[i] = I.x[v[i]] - qx;
      py[i] = I.y[v[i]] - qy;
    }
    double ctx, cty, dt;
    cc3(px[0], py[0], px[1], py[1], px[2], py[2], ctx, cty, dt);
    (void)dt;  // a cavity triangle is never degenerate
    // circumcentres of (q, v_i, v_{i+1})
    double gx[3], gy[3];
    for (int i = 0; i < 3; ++i) {
      double d;
      cc0(px[i], py[i], px[(i + 1) % 3], py[(i + 1) % 3], gx[i], gy[i], d);
      if (d == 0.0) return false;
    }
    for (int i = 0; i < 3; ++i) {
      // vertex a = v_i: g(q, a, b) = gx[i], g(q, c, a) = gx[i + 2]
      const int j = (i + 2) % 3;
      const double a = 0.5 * orient(ctx, cty, gx[j], gy[j], gx[i], gy[i]);
      num += a * val[v[i]];
      den += a;
    }
  }
  out = num / den;
  return true;
}

// status bits: 1 a walk did not end, 2 a cavity overflowed
__global__ void k_nn_query(InterpDev I, const double* __restrict__ val, uint32_t m, const double* __restrict__ qxs,
                           const double* __restrict__ qys, double* __restrict__ out, uint8_t* __restrict__ found,
                           uint32_t* __restrict__ status) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  double qx = qxs[k], qy = qys[k];
  out[k] = __builtin_nan("");
  found[k] = 0;
  if (!__builtin_isfinite(qx) || !__builtin_isfinite(qy) || I.ntri == 0) return;
  // 1. location
  int cx = (int)__builtin_floor((qx - I.gx0) * I.gix), cy = (int)__builtin_floor((qy - I.gy0) * I.giy);
  cx = cx < 0 ? 0 : (cx >= (int)I.gw ? (int)I.gw - 1 : cx);
  cy = cy < 0 ? 0 : (cy >= (int)I.gh ? (int)I.gh - 1 : cy);
  uint32_t t = I.grid[(uint32_t)cy * I.gw + (uint32_t)cx];
  uint32_t start = k % 3u;  // which edge is tried first: varies, so walks do not repeat a pattern
  uint32_t steps = 0;
  for (;; ++steps) {
    if (steps >= kMaxWalk) {
      atomicOr(status, 1u);
      return;
    }
    uint32_t cross = 3;
    for (uint32_t e = 0; e < 3; ++e) {
      const uint32_t i = (start + e) % 3u;
      const uint32_t a = I.tri[3 * t + i], b = I.tri[3 * t + (i + 1) % 3];
      if (orient(I.x[a], I.y[a], I.x[b], I.y[b], qx, qy) < 0.0) {
        cross = i;
        break;
      }
    }
    if (cross == 3) break;  // q in (or on the boundary of) triangle t
    const uint32_t u = I.nbr[3 * t + cross];
    if (u == FL_NONE) return;  // beyond a hull edge: outside the convex hull
    t = u;
    start = (start + 1) % 3u;
  }
  // a site itself
  for (uint32_t i = 0; i < 3; ++i) {
    const uint32_t v = I.tri[3 * t + i];
    if (I.x[v] == qx && I.y[v] == qy) {
      out[k] = val[v];
      found[k] = 1;
      return;
    }
  }
  // 2. cavity (breadth-first over triangle neighbours)
  uint32_t cav[kMaxCavity];
  uint32_t nc = 1;
  cav[0] = t;
  for (uint32_t h = 0; h < nc; ++h) {
    const uint32_t s = cav[h];
    for (uint32_t e = 0; e < 3; ++e) {
      const uint32_t u = I.nbr[3 * s + e];
      if (u == FL_NONE) continue;
      bool seen = false;
      for (uint32_t z = 0; z < nc && !seen; ++z) seen = cav[z] == u;
      if (seen) continue;
      const uint32_t a = I.tri[3 * u], b = I.tri[3 * u + 1], c = I.tri[3 * u + 2];
      if (!in_circle(I.x[a] - qx, I.y[a] - qy, I.x[b] - qx, I.y[b] - qy, I.x[c] - qx, I.y[c] - qy)) continue;
      if (nc == kMaxCavity) {
        atomicOr(status, 2u);
        return;
      }
      cav[nc++] = u;
    }
  }
  // 3. Sibson weights; q collinear with a cavity edge makes a new
  // circumcentre degenerate: the value is then taken at q moved off the
  // line by 2^-40 of the local edge scale (the interpolant is continuous)
  double v;
  if (!watson(I, val, cav, nc, qx, qy, v)) {
    const uint32_t a = I.tri[3 * t], b = I.tri[3 * t + 1];
    const double ex = I.x[b] - I.x[a], ey = I.y[b] - I.y[a];
    const double h = 0x1p-40 * __builtin_sqrt(ex * ex + ey * ey);
    if (!watson(I, val, cav, nc, qx + h * 0.6, qy + h * 0.8, v)) return;
  }
  out[k] = v;
  found[k] = 1;
}

}  // namespace

struct fl_interp {
  fl_ctx* ctx = nullptr;
  int device = 0;
  hipStream_t stream = nullptr;
  InterpDev dev{};
  std::vector<void*> allocs;
};

namespace {

int32_t ierr(int32_t code, const std::string& msg) {
  fl_internal_set_error(msg.c_str());
  return code;
}

#define IHIP(call)                                                                                       \
  do {                                                                                                   \
    hipError_t e_ = (call);                                                                              \
    if (e_ != hipSuccess)                                                                                \
      return ierr(e_ == hipErrorOutOfMemory ? FL_ERR_OUT_OF_MEMORY : FL_ERR_HIP,                         \
                  std::string(#call) + ": " + hipGetErrorString(e_));                                    \
  } while (0)

#define FLCHK_I(call)     \
  do {                    \
    int32_t r_ = (call);  \
    if (r_ != FL_OK) return r_; \
  } while (0)

template <typename T>
int32_t ialloc(fl_interp* P, T** p, uint64_t count) {
  IHIP(hipMalloc((void**)p, std::max<uint64_t>(count, 1) * sizeof(T)));
  P->allocs.push_back(*p);
  return FL_OK;
}

}  // namespace

extern "C" int32_t fl_interp_destroy(fl_interp* P) {
  if (!P) return FL_OK;
  hipSetDevice(P->device);
  for (void* p : P->allocs) hipFree(p);
  delete P;
  return FL_OK;
}

extern "C" int32_t fl_interp_create(fl_ctx* ctx, uint64_t n, const double* xs, const double* ys, uint64_t ntri,
                                    const uint32_t* triangles, fl_interp** out) {
  if (!ctx || !out || (n && (!xs || !ys)) || (ntri && !triangles)) return ierr(FL_ERR_INVALID, "null argument");
  if (n >= FL_NONE || ntri >= FL_NONE / 3) return ierr(FL_ERR_TOO_LARGE, "too many sites or triangles");
  hipStream_t s = nullptr;
  int device = 0;
  int32_t rc = fl_ctx_stream(ctx, &s, &device);
  if (rc != FL_OK) return rc;
  IHIP(hipSetDevice(device));
  // triangles counter-clockwise; the neighbour across each edge
  std::vector<uint32_t> tri(triangles, triangles + 3 * ntri), nbr(3 * ntri, FL_NONE);
  for (uint64_t t = 0; t < ntri; t++) {
    uint32_t* v = &tri[3 * t];
    for (int i = 0; i < 3; i++)
      if (v[i] >= n) return ierr(FL_ERR_INVALID, "triangle vertex out of range");
    const double o = (xs[v[1]] - xs[v[0]]) * (ys[v[2]] - ys[v[0]]) - (ys[v[1]] - ys[v[0]]) * (xs[v[2]] - xs[v[0]]);
    if (o == 0.0) return ierr(FL_ERR_INVALID, "degenerate (zero-area) triangle");
    if (o < 0.0) std::swap(v[1], v[2]);
  }
  std::unordered_map<uint64_t, uint64_t> open;  // directed edge (a, b) -> slot 3 t + i
  open.reserve(3 * ntri);
  for (uint64_t t = 0; t < ntri; t++)
    for (int i = 0; i < 3; i++) {
      const uint64_t a = tri[3 * t + i], b = tri[3 * t + (i + 1) % 3];
      auto it = open.find((b << 32) | a);  // the same edge, the other way round, in a neighbour
      if (it != open.end()) {
        if (nbr[it->second] != FL_NONE) return ierr(FL_ERR_INVALID, "an edge is shared by more than two triangles");
        nbr[it->second] = (uint32_t)t;
        nbr[3 * t + i] = (uint32_t)(it->second / 3);
      } else if (!open.emplace((a << 32) | b, 3 * t + i).second) {
        return ierr(FL_ERR_INVALID, "two triangles hold the same directed edge");
      }
    }
  // grid: about one cell per two triangles, a triangle near every cell
  double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
  for (uint64_t i = 0; i < n; i++) {
    x0 = std::min(x0, xs[i]);
    x1 = std::max(x1, xs[i]);
    y0 = std::min(y0, ys[i]);
    y1 = std::max(y1, ys[i]);
  }
  const double w = std::max(x1 - x0, 1e-300), h = std::max(y1 - y0, 1e-300);
  const double cells = std::max<double>(1.0, (double)ntri / 2.0);
  uint32_t gw = (uint32_t)std::max(1.0, std::min(65536.0, std::round(std::sqrt(cells * w / h))));
  uint32_t gh = (uint32_t)std::max(1.0, std::min(65536.0, std::round(cells / gw)));
  std::vector<uint32_t> grid((size_t)gw * gh, FL_NONE);
  const double gix = gw / w, giy = gh / h;
  auto cell_of = [&](double x, double y) {
    int cx = (int)std::floor((x - x0) * gix), cy = (int)std::floor((y - y0) * giy);
    cx = std::min(std::max(cx, 0), (int)gw - 1);
    cy = std::min(std::max(cy, 0), (int)gh - 1);
    return (size_t)cy * gw + cx;
  };
  for (uint64_t t = 0; t < ntri; t++) {
    const uint32_t* v = &tri[3 * t];
    grid[cell_of((xs[v[0]] + xs[v[1]] + xs[v[2]]) / 3.0, (ys[v[0]] + ys[v[1]] + ys[v[2]]) / 3.0)] = (uint32_t)t;
  }
  // empty cells take a neighbouring cell's triangle (breadth-first fill)
  if (ntri) {
    std::vector<size_t> queue;
    queue.reserve(grid.size());
    for (size_t c = 0; c < grid.size(); c++)
      if (grid[c] != FL_NONE) queue.push_back(c);
    for (size_t h2 = 0; h2 < queue.size(); h2++) {
      const size_t c = queue[h2];
      const size_t cx = c % gw, cy = c / gw;
      const size_t nb[4] = {cx > 0 ? c - 1 : c, cx + 1 < gw ? c + 1 : c, cy > 0 ? c - gw : c,
                            cy + 1 < gh ? c + gw : c};
      for (size_t d : nb)
        if (grid[d] == FL_NONE) {
          grid[d] = grid[c];
          queue.push_back(d);
        }
    }
  }
  std::unique_ptr<fl_interp, int32_t (*)(fl_interp*)> P(new fl_interp(), fl_interp_destroy);
  P->ctx = ctx;
  P->device = device;
  P->stream = s;
  double *dx, *dy;
  uint32_t *dtri, *dnbr, *dgrid;
  FLCHK_I(ialloc(P.get(), &dx, n));
  FLCHK_I(ialloc(P.get(), &dy, n));
  FLCHK_I(ialloc(P.get(), &dtri, 3 * ntri));
  FLCHK_I(ialloc(P.get(), &dnbr, 3 * ntri));
  FLCHK_I(ialloc(P.get(), &dgrid, grid.size()));
  IHIP(hipMemcpyAsync(dx, xs, n * 8, hipMemcpyHostToDevice, s));
  IHIP(hipMemcpyAsync(dy, ys, n * 8, hipMemcpyHostToDevice, s));
  IHIP(hipMemcpyAsync(dtri, tri.data(), tri.size() * 4, hipMemcpyHostToDevice, s));
  IHIP(hipMemcpyAsync(dnbr, nbr.data(), nbr.size() * 4, hipMemcpyHostToDevice, s));
  IHIP(hipMemcpyAsync(dgrid, grid.data(), grid.size() * 4, hipMemcpyHostToDevice, s));
  IHIP(hipStreamSynchronize(s));  // the host arrays above must outlive the copies
  P->dev = InterpDev{(uint32_t)n, (uint32_t)ntri, dx, dy, dtri, dnbr, dgrid, x0, y0, gix, giy, gw, gh};
  *out = P.release();
  return FL_OK;
}

extern "C" int32_t fl_interp_query(fl_ctx* ctx, fl_interp* P, const double* values, uint64_t m, const double* qx,
                                   const double* qy, double* out, uint8_t* found) {
  if (!ctx || !P || (P->dev.n && !values) || (m && (!qx || !qy || !out || !found)))
    return ierr(FL_ERR_INVALID, "null argument");
  if (m >= FL_NONE) return ierr(FL_ERR_TOO_LARGE, "too many query points");
  if (m == 0) return FL_OK;
  hipStream_t s = nullptr;
  int device = 0;
  int32_t rc = fl_ctx_stream(ctx, &s, &device);
  if (rc != FL_OK) return rc;
  IHIP(hipSetDevice(device));
  double *dval = nullptr, *dq = nullptr, *dout = nullptr;
  uint8_t* dfound = nullptr;
  uint32_t* dstatus = nullptr;
  const uint64_t n = P->dev.n;
  IHIP(hipMallocAsync((void**)&dval, std::max<uint64_t>(n, 1) * 8, s));
  IHIP(hipMallocAsync((void**)&dq, 2 * m * 8, s));
  IHIP(hipMallocAsync((void**)&dout, m * 8, s));
  IHIP(hipMallocAsync((void**)&dfound, m + 8, s));
  IHIP(hipMallocAsync((void**)&dstatus, 8, s));
  uint32_t status = 0;
  hipError_t e = hipSuccess;
  auto go = [&]() -> hipError_t {
    hipError_t r;
    if ((r = hipMemcpyAsync(dval, values, n * 8, hipMemcpyHostToDevice, s)) != hipSuccess) return r;
    if ((r = hipMemcpyAsync(dq, qx, m * 8, hipMemcpyHostToDevice, s)) != hipSuccess) return r;
    if ((r = hipMemcpyAsync(dq + m, qy, m * 8, hipMemcpyHostToDevice, s)) != hipSuccess) return r;
    if ((r = hipMemcpyAsync(dstatus, &status, 4, hipMemcpyHostToDevice, s)) != hipSuccess) return r;
    hipLaunchKernelGGL(k_nn_query, dim3((uint32_t)((m + 127) / 128)), dim3(128), 0, s, P->dev, dval, (uint32_t)m, dq,
                       dq + m, dout, dfound, dstatus);
    if ((r = hipGetLastError()) != hipSuccess) return r;
    if ((r = hipMemcpyAsync(out, dout, m * 8, hipMemcpyDeviceToHost, s)) != hipSuccess) return r;
    if ((r = hipMemcpyAsync(found, dfound, m, hipMemcpyDeviceToHost, s)) != hipSuccess) return r;
    if ((r = hipMemcpyAsync(&status, dstatus, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return r;
    return hipStreamSynchronize(s);
  };
  e = go();
  for (void* p : {(void*)dval, (void*)dq, (void*)dout, (void*)dfound, (void*)dstatus}) (void)hipFreeAsync(p, s);
  if (e != hipSuccess) return ierr(FL_ERR_HIP, std::string("interpolation: ") + hipGetErrorString(e));
  if (status & 1u) return ierr(FL_ERR_INVALID, "a point-location walk did not end (not a Delaunay triangulation?)");
  if (status & 2u) return ierr(FL_ERR_INVALID, "a query's natural-neighbour cavity exceeds the kernel's limit");
  return FL_OK;
}

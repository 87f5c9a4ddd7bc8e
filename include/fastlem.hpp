// fastlem.hpp -- C++ host API over the C ABI of fastlem_hip.h.
//
// Mirrors the parts of fastlem 0.1.4's public API that generate() consumes,
// with the same names, argument meaning and error behaviour:
//   fastlem::core::parameters::TopographicalParameters   src/core/parameters.rs:24-69
//   fastlem::core::traits::Model (num/areas/graph/default_outlets)
//                                                         src/core/traits.rs:13-20
//   fastlem::lem::generator::{TerrainGenerator, GenerationError}
//                                                         src/lem/generator.rs:19-213
//   fastlem::models::surface::{Site2D, TerrainModel2D, Terrain2D}
//                                                         src/models/surface/{sites,model,terrain}.rs
// generate() validates and resolves outlets on the host exactly as
// generator.rs:91-132, draws the jitter of generator.rs:134-138 with the
// StdRng restatement, and runs the loop of generator.rs:140-210 on the GPU.
// A device failure throws std::runtime_error carrying the C ABI message.
#pragma once

#include <cstdint>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "fastlem_hip.h"

namespace fastlem {

struct Site2D {
  double x = 0.0, y = 0.0;
};

// src/core/parameters.rs:24-69 (defaults: base 0, erodibility 1, uplift 1,
// not an outlet, no maximum slope).
class TopographicalParameters {
 public:
  TopographicalParameters set_base_elevation(double v) const { auto p = *this; p.base_elevation = v; return p; }
  TopographicalParameters set_erodibility(double v) const { auto p = *this; p.erodibility = v; return p; }
  TopographicalParameters set_uplift_rate(double v) const { auto p = *this; p.uplift_rate = v; return p; }
  TopographicalParameters set_is_outlet(bool v) const { auto p = *this; p.is_outlet = v; return p; }
  TopographicalParameters set_max_slope(std::optional<double> v) const { auto p = *this; p.max_slope = v; return p; }

  double base_elevation = 0.0;
  double erodibility = 1.0;
  double uplift_rate = 1.0;
  bool is_outlet = false;
  std::optional<double> max_slope;  // radians
};

// The graph as generate() reads it: row i lists neighbors_of(i) in order.
struct Graph {
  std::vector<uint32_t> row_ptr;  // n+1
  std::vector<uint32_t> col;
  std::vector<double> length;
  size_t order() const { return row_ptr.empty() ? 0 : row_ptr.size() - 1; }
};

// The calling thread's device context (HIP device 0), created on first use
// and destroyed when the thread ends.  A context is one stream with its own
// host slots and calls on one context must not overlap, so generate() calls
// made at once from several threads each run on a context of their own (and
// their work overlaps on the device), as separate reference generators run
// independently.
inline fl_ctx* default_context() {
  struct Holder {
    fl_ctx* ctx = nullptr;
    ~Holder() {
      if (ctx) fl_ctx_destroy(ctx);
    }
  };
  thread_local Holder h;
  if (!h.ctx) {
    fl_ctx* c = nullptr;
    if (fl_ctx_create(0, &c) != FL_OK) throw std::runtime_error(fl_last_error());
    h.ctx = c;
  }
  return h.ctx;
}

// src/models/surface/terrain.rs:8-39: sites, elevations and get_elevation,
// the natural-neighbour interpolation of the elevations
// (src/models/surface/interpolator.rs:11-27) on the GPU.  The interpolator
// is built on the first query from the model's Delaunay triangles (the
// reference builds it inside generate()); copies share it.
class Terrain2D {
 public:
  Terrain2D(std::vector<Site2D> sites, std::vector<double> elevations,
            std::shared_ptr<const std::vector<uint32_t>> triangles = nullptr)
      : sites_(std::move(sites)), elevations_(std::move(elevations)), triangles_(std::move(triangles)) {}
  const std::vector<Site2D>& sites() const { return sites_; }
  const std::vector<double>& elevations() const { return elevations_; }

  // None (std::nullopt) outside the sites' convex hull.
  std::optional<double> get_elevation(const Site2D& site) const {
    double v = 0.0;
    uint8_t found = 0;
    if (fl_interp_query(default_context(), interpolator(), elevations_.data(), 1, &site.x, &site.y, &v, &found) !=
        FL_OK)
      throw std::runtime_error(fl_last_error());
    return found ? std::optional<double>(v) : std::nullopt;
  }
  // get_elevation for many points in one launch: NaN where it is None.
  std::vector<double> get_elevations(const std::vector<Site2D>& points) const {
    std::vector<double> xs(points.size()), ys(points.size()), out(points.size());
    std::vector<uint8_t> found(points.size());
    for (size_t k = 0; k < points.size(); k++) {
      xs[k] = points[k].x;
      ys[k] = points[k].y;
    }
    if (fl_interp_query(default_context(), interpolator(), elevations_.data(), points.size(), xs.data(), ys.data(),
                        out.data(), found.data()) != FL_OK)
      throw std::runtime_error(fl_last_error());
    return out;
  }

 private:
  fl_interp* interpolator() const {
    if (!interp_) {
      if (!triangles_) throw std::runtime_error("get_elevation needs the model's triangles (TerrainModel2D::set_triangles)");
      std::vector<double> xs(sites_.size()), ys(sites_.size());
      for (size_t i = 0; i < sites_.size(); i++) {
        xs[i] = sites_[i].x;
        ys[i] = sites_[i].y;
      }
      fl_interp* h = nullptr;
      if (fl_interp_create(default_context(), sites_.size(), xs.data(), ys.data(), triangles_->size() / 3,
                           triangles_->data(), &h) != FL_OK)
        throw std::runtime_error(fl_last_error());
      interp_ = std::shared_ptr<fl_interp>(h, [](fl_interp* p) { fl_interp_destroy(p); });
    }
    return interp_.get();
  }
  std::vector<Site2D> sites_;
  std::vector<double> elevations_;
  std::shared_ptr<const std::vector<uint32_t>> triangles_;
  mutable std::shared_ptr<fl_interp> interp_;
};

// src/models/surface/model.rs:18-69 (Model implementation).
class TerrainModel2D {
 public:
  TerrainModel2D(std::vector<Site2D> sites, std::vector<double> areas, Graph graph,
                 std::vector<uint32_t> default_outlets)
      : sites_(std::move(sites)), areas_(std::move(areas)), graph_(std::move(graph)),
        default_outlets_(std::move(default_outlets)) {}
  size_t num() const { return graph_.order(); }
  const std::vector<Site2D>& sites() const { return sites_; }
  const std::vector<double>& areas() const { return areas_; }
  const std::vector<uint32_t>& default_outlets() const { return default_outlets_; }
  const Graph& graph() const { return graph_; }
  // The sites' Delaunay triangles (3 site indices each), for get_elevation.
  TerrainModel2D set_triangles(std::vector<uint32_t> triangles) const {
    auto m = *this;
    m.triangles_ = std::make_shared<const std::vector<uint32_t>>(std::move(triangles));
    return m;
  }
  // src/models/surface/model.rs:62-68
  Terrain2D create_terrain_from_result(const std::vector<double>& elevations) const {
    return Terrain2D(sites_, elevations, triangles_);
  }

 private:
  std::shared_ptr<const std::vector<uint32_t>> triangles_;
  std::vector<Site2D> sites_;
  std::vector<double> areas_;
  Graph graph_;
  std::vector<uint32_t> default_outlets_;
};

// src/lem/generator.rs:19-26
enum class GenerationError { InvalidNumberOfParameters, ParametersNotSet, ModelNotSet };

inline const char* to_string(GenerationError e) {
  switch (e) {
    case GenerationError::InvalidNumberOfParameters:
      return "The number of topographical parameters must be equal to the number of sites";
    case GenerationError::ParametersNotSet:
      return "You must set topographical parameters before generating terrain";
    case GenerationError::ModelNotSet:
      return "You must set `TerrainModel` before generating terrain";
  }
  return "";
}

// Result<T, GenerationError>
template <typename T>
class Result {
 public:
  static Result ok(T v) { return Result(std::move(v)); }
  static Result err(GenerationError e) { return Result(e); }
  bool is_ok() const { return value_.has_value(); }
  const T& unwrap() const {
    if (!value_) throw std::runtime_error(std::string("called unwrap on an error: ") + to_string(*error_));
    return *value_;
  }
  GenerationError error() const { return *error_; }

 private:
  explicit Result(T v) : value_(std::move(v)) {}
  explicit Result(GenerationError e) : error_(e) {}
  std::optional<T> value_;
  std::optional<GenerationError> error_;
};

// src/lem/generator.rs:37-213
template <typename M = TerrainModel2D>
class TerrainGenerator {
 public:
  TerrainGenerator set_model(M model) const {
    auto g = *this;
    g.model_ = std::move(model);
    return g;
  }
  TerrainGenerator set_parameters(std::vector<TopographicalParameters> parameters) const {
    auto g = *this;
    g.parameters_ = std::move(parameters);
    return g;
  }
  TerrainGenerator set_max_iteration(uint32_t max_iteration) const {
    auto g = *this;
    g.max_iteration_ = max_iteration;
    return g;
  }

  // Outer iterations run by the last generate() on this object.
  uint32_t last_iterations() const { return last_iterations_; }

  auto generate() -> Result<decltype(std::declval<M>().create_terrain_from_result({}))> {
    using T = decltype(std::declval<M>().create_terrain_from_result({}));
    if (!model_) return Result<T>::err(GenerationError::ModelNotSet);
    const M& model = *model_;
    size_t num = model.num();
    if (!parameters_) return Result<T>::err(GenerationError::ParametersNotSet);
    const auto& params = *parameters_;
    if (params.size() != num) return Result<T>::err(GenerationError::InvalidNumberOfParameters);
    // generator.rs:120-132
    std::vector<uint32_t> outlets;
    for (size_t i = 0; i < num; i++)
      if (params[i].is_outlet) outlets.push_back((uint32_t)i);
    if (outlets.empty()) outlets = model.default_outlets();
    // generator.rs:134-138
    std::vector<double> base(num), erod(num), uplift(num), init(num), max_slope(num, 0.0);
    std::vector<uint8_t> has_max_slope(num, 0);
    bool any_ms = false;
    for (size_t i = 0; i < num; i++) {
      base[i] = params[i].base_elevation;
      erod[i] = params[i].erodibility;
      uplift[i] = params[i].uplift_rate;
      if (params[i].max_slope) {
        has_max_slope[i] = 1;
        max_slope[i] = *params[i].max_slope;
        any_ms = true;
      }
    }
    fl_jitter_elevations(0, num, base.data(), init.data());
    const Graph& g = model.graph();
    fl_terrain_desc d{};
    d.n = num;
    d.nnz = g.col.size();
    d.row_ptr = g.row_ptr.data();
    d.col = g.col.data();
    d.length = g.length.data();
    d.area = model.areas().data();
    d.erodibility = erod.data();
    d.uplift_rate = uplift.data();
    d.has_max_slope = any_ms ? has_max_slope.data() : nullptr;
    d.max_slope = any_ms ? max_slope.data() : nullptr;
    d.init_elevation = init.data();
    d.outlets = outlets.data();
    d.n_outlets = outlets.size();
    d.max_iteration = max_iteration_ ? *max_iteration_ : FL_UNBOUNDED_ITERATIONS;
    std::vector<double> out(num);
    uint32_t iters = 0;
    if (fl_generate_f64(default_context(), &d, out.data(), &iters) != FL_OK) throw std::runtime_error(fl_last_error());
    last_iterations_ = iters;
    return Result<T>::ok(model.create_terrain_from_result(out));
  }

 private:
  std::optional<M> model_;
  std::optional<std::vector<TopographicalParameters>> parameters_;
  std::optional<uint32_t> max_iteration_;
  uint32_t last_iterations_ = 0;
};

}  // namespace fastlem

// Exercises the C++ host API (include/fastlem.hpp) the way the reference's
// own tests use TerrainGenerator (tests/landscape_evolution.rs:8-54).
//   generate_example errors          -> the three GenerationError cases
//   generate_example run IN OUT [K]  -> IN: n, nnz, row_ptr, col, length,
//                                       area, is_outlet(u8), erodibility,
//                                       max_slope (NaN = None) as raw arrays;
//                                       OUT: iterations (u32) + elevations;
//                                       K: set_max_iteration(K)
//   generate_example run2 IN OUT      -> the same generate() from two threads
//                                       at once; OUT holds both results
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <thread>
#include <vector>

#include "fastlem.hpp"

using namespace fastlem;

template <typename T>
static std::vector<T> read_vec(std::ifstream& f, size_t count) {
  std::vector<T> v(count);
  f.read(reinterpret_cast<char*>(v.data()), count * sizeof(T));
  return v;
}

int main(int argc, char** argv) {
  if (argc >= 2 && std::strcmp(argv[1], "errors") == 0) {
    Graph g;
    g.row_ptr = {0, 1, 2};
    g.col = {1, 0};
    g.length = {1.0, 1.0};
    TerrainModel2D model({{0, 0}, {1, 0}}, {1.0, 1.0}, g, {0});
    int ok = 0;
    ok += TerrainGenerator<>().generate().error() == GenerationError::ModelNotSet;
    ok += TerrainGenerator<>().set_model(model).generate().error() == GenerationError::ParametersNotSet;
    ok += TerrainGenerator<>()
              .set_model(model)
              .set_parameters({TopographicalParameters()})
              .generate()
              .error() == GenerationError::InvalidNumberOfParameters;
    TopographicalParameters p;
    ok += p.erodibility == 1.0 && p.uplift_rate == 1.0 && !p.max_slope && !p.is_outlet;
    std::cout << "errors " << ok << "/4" << std::endl;
    return ok == 4 ? 0 : 1;
  }
  if (argc >= 4 && (std::strcmp(argv[1], "run") == 0 || std::strcmp(argv[1], "run2") == 0)) {
    const bool two = std::strcmp(argv[1], "run2") == 0;
    std::ifstream f(argv[2], std::ios::binary);
    uint64_t hdr[2];
    f.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
    size_t n = hdr[0], nnz = hdr[1];
    Graph g;
    g.row_ptr = read_vec<uint32_t>(f, n + 1);
    g.col = read_vec<uint32_t>(f, nnz);
    g.length = read_vec<double>(f, nnz);
    auto area = read_vec<double>(f, n);
    auto is_outlet = read_vec<uint8_t>(f, n);
    auto erod = read_vec<double>(f, n);
    auto ms = read_vec<double>(f, n);
    std::vector<Site2D> sites(n);
    TerrainModel2D model(sites, area, g, {});
    std::vector<TopographicalParameters> params;
    for (size_t i = 0; i < n; i++)
      params.push_back(TopographicalParameters()
                           .set_erodibility(erod[i])
                           .set_is_outlet(is_outlet[i] != 0)
                           .set_max_slope(std::isnan(ms[i]) ? std::optional<double>() : std::optional<double>(ms[i])));
    auto gen = TerrainGenerator<>().set_model(model).set_parameters(params);
    if (argc >= 5) gen = gen.set_max_iteration((uint32_t)std::strtoul(argv[4], nullptr, 10));
    std::ofstream o(argv[3], std::ios::binary);
    if (!two) {
      auto terrain = gen.generate().unwrap();
      uint32_t it = gen.last_iterations();
      o.write(reinterpret_cast<const char*>(&it), sizeof(it));
      o.write(reinterpret_cast<const char*>(terrain.elevations().data()), n * sizeof(double));
      return 0;
    }
    // two generators (copies) run at once on two threads
    auto g2 = gen;
    std::vector<double> e1, e2;
    std::thread t1([&] { e1 = gen.generate().unwrap().elevations(); });
    std::thread t2([&] { e2 = g2.generate().unwrap().elevations(); });
    t1.join();
    t2.join();
    for (auto* g : {&gen, &g2}) {
      uint32_t it = g->last_iterations();
      o.write(reinterpret_cast<const char*>(&it), sizeof(it));
    }
    o.write(reinterpret_cast<const char*>(e1.data()), n * sizeof(double));
    o.write(reinterpret_cast<const char*>(e2.data()), n * sizeof(double));
    return 0;
  }
  std::cerr << "usage: generate_example errors | run IN OUT" << std::endl;
  return 2;
}

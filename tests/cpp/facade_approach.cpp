// The reference's idiom for adapting to a target metric
// (src/aniso_test.cpp:15-27), written against the C++ facade:
//
//   mesh.add_tag(VERT, "metric", ...); mesh.add_tag(VERT, "target_metric", ...);
//   while (approach_metric(&mesh, opts)) adapt(&mesh, opts);
//
//   facade_approach <dim> <n> <ncomps> <metric.f64> <target.f64> <out_prefix>
//
// writes <out_prefix>.counts (int32 nents per dimension, then the number of
// approach steps), .verts (int32), .coords and .metric (float64).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "Omega_h_amd.hpp"

static std::vector<double> read_doubles(std::string const& path, std::size_t n) {
  std::vector<double> v(n);
  std::ifstream f(path, std::ios::binary);
  f.read(reinterpret_cast<char*>(v.data()), std::streamsize(n * 8));
  if (!f) throw std::runtime_error("cannot read " + path);
  return v;
}

template <class T>
static void write_file(std::string const& path, T const* data, std::size_t n) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<char const*>(data), std::streamsize(n * sizeof(T)));
  if (!f) throw std::runtime_error("cannot write " + path);
}

// a host vector as a device array, built the way a caller builds one
// (HostWrite, then write())
static Omega_h::Reals to_reals(std::vector<double> const& v) {
  Omega_h::HostWrite<Omega_h::Real> h(Omega_h::LO(v.size()));
  for (std::size_t i = 0; i < v.size(); ++i) h[Omega_h::LO(i)] = v[i];
  return Omega_h::Reals(h.write());
}

// a device array written through its host mirror
template <class T>
static void write_read(std::string const& path, Omega_h::Read<T> const& r) {
  Omega_h::HostRead<T> h(r);
  write_file(path, h.data(), std::size_t(h.size()));
}

int main(int argc, char** argv) {
  if (argc != 7) {
    std::cerr << "usage: facade_approach <dim> <n> <ncomps> <metric.f64> <target.f64> <out>\n";
    return 2;
  }
  int dim = std::atoi(argv[1]);
  int n = std::atoi(argv[2]);
  int ncomps = std::atoi(argv[3]);
  std::string prefix = argv[6];
  try {
    Omega_h::Library lib(&argc, &argv);
    Omega_h::Mesh mesh = Omega_h::build_box(lib.world(), OMEGA_H_SIMPLEX, 1.0, 1.0,
        dim == 3 ? 1.0 : 0.0, n, n, dim == 3 ? n : 0);
    std::size_t nv = std::size_t(mesh.nverts());
    mesh.add_tag(Omega_h::VERT, "metric", ncomps,
        to_reals(read_doubles(argv[4], nv * ncomps)));
    mesh.add_tag(Omega_h::VERT, "target_metric", ncomps,
        to_reals(read_doubles(argv[5], nv * ncomps)));
    Omega_h::AdaptOpts opts(&mesh);
    opts.verbosity = Omega_h::SILENT;
    int steps = 0;
    while (Omega_h::approach_metric(&mesh, opts)) {
      Omega_h::adapt(&mesh, opts);
      ++steps;
    }
    std::vector<int> counts;
    for (int d = 0; d <= dim; ++d) counts.push_back(mesh.nents(d));
    counts.push_back(steps);
    write_file(prefix + ".counts", counts.data(), counts.size());
    auto ev2v = mesh.ask_elem_verts();
    write_read(prefix + ".verts", ev2v);
    auto coords = mesh.coords();
    write_read(prefix + ".coords", coords);
    auto metric = mesh.get_array<Omega_h::Real>(Omega_h::VERT, "metric");
    write_read(prefix + ".metric", metric);
    std::printf("approached in %d steps: %d elements\n", steps, mesh.nelems());
  } catch (Omega_h::exception const& e) {
    std::cerr << "Omega_h::exception: " << e.what() << "\n";
    return 1;
  }
  return 0;
}

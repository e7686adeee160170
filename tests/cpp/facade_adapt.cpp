// Drives adapt() through the C++ facade the way an Omega_h caller does
// (src/aniso_test.cpp:15-27): build_box, add the vertex metric tag, adapt,
// read the result back.  Used by tests/test_facade.py.
//
//   facade_adapt <dim> <n> <ncomps> <metric.f64> <out_prefix> [every]
//
// With "every", the call site assigns every field of the reference's
// AdaptOpts (src/Omega_h_adapt.hpp:50-77) and of its TransferOpts
// (:30-42) explicitly, with should_prevent_coarsen_flip on.
//
// writes <out_prefix>.counts (int32 nents per dimension, then the adapt()
// return value), .verts (int32 element vertices), .coords and .quality
// (float64).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "Omega_h_amd.hpp"

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
  if (argc != 6 && argc != 7) {
    std::cerr << "usage: facade_adapt <dim> <n> <ncomps> <metric.f64> <out_prefix> [every]\n";
    return 2;
  }
  int dim = std::atoi(argv[1]);
  int n = std::atoi(argv[2]);
  int ncomps = std::atoi(argv[3]);
  std::string prefix = argv[5];
  bool every = argc == 7 && std::string(argv[6]) == "every";
  try {
    Omega_h::Library lib(&argc, &argv);
    Omega_h::Mesh mesh = Omega_h::build_box(lib.world(), OMEGA_H_SIMPLEX, 1.0, 1.0,
        dim == 3 ? 1.0 : 0.0, n, n, dim == 3 ? n : 0);
    std::vector<double> metric(std::size_t(mesh.nverts()) * ncomps);
    std::ifstream f(argv[4], std::ios::binary);
    f.read(reinterpret_cast<char*>(metric.data()), std::streamsize(metric.size() * 8));
    if (!f) throw std::runtime_error("cannot read the metric file");
    mesh.add_tag(Omega_h::VERT, "metric", ncomps,
        to_reals(metric));
    Omega_h::AdaptOpts opts(&mesh);
    opts.verbosity = Omega_h::SILENT;
    if (every) {
      Omega_h::AdaptOpts d(dim);
      opts.min_length_desired = d.min_length_desired;
      opts.max_length_desired = d.max_length_desired;
      opts.max_length_allowed = d.max_length_allowed;
      opts.min_quality_allowed = d.min_quality_allowed;
      opts.min_quality_desired = d.min_quality_desired;
      opts.nsliver_layers = d.nsliver_layers;
      opts.verbosity = Omega_h::SILENT;
      opts.length_histogram_min = 0.0;
      opts.length_histogram_max = 3.0;
      opts.nlength_histogram_bins = 10;
      opts.nquality_histogram_bins = 10;
      opts.should_refine = true;
      opts.should_coarsen = true;
      opts.should_swap = true;
      opts.should_coarsen_slivers = true;
      opts.should_prevent_coarsen_flip = true;
      opts.xfer_opts.type_map.clear();
      opts.xfer_opts.integral_map.clear();
      opts.xfer_opts.velocity_density_map.clear();
      opts.xfer_opts.velocity_momentum_map.clear();
      opts.xfer_opts.integral_diffuse_map.clear();
    }
    bool did = Omega_h::adapt(&mesh, opts);
    std::vector<int> counts;
    for (int d = 0; d <= dim; ++d) counts.push_back(mesh.nents(d));
    counts.push_back(did ? 1 : 0);
    write_file(prefix + ".counts", counts.data(), counts.size());
    auto ev2v = mesh.ask_elem_verts();
    write_read(prefix + ".verts", ev2v);
    auto coords = mesh.coords();
    write_read(prefix + ".coords", coords);
    auto q = mesh.get_array<Omega_h::Real>(dim, "quality");
    write_read(prefix + ".quality", q);
    std::printf("adapted: %d elements, min quality %.6f\n", mesh.nelems(),
        mesh.min_quality());
  } catch (Omega_h::exception const& e) {
    std::cerr << "Omega_h::exception: " << e.what() << "\n";
    return 1;
  }
  return 0;
}

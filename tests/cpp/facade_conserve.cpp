// The reference's conservation tests (src/2d_conserve_test.cpp,
// src/3d_conserve_test.cpp) written against the C++ facade: a conserved
// density ("mass") and a momentum velocity ("momentum") carried through
// adapt() with the transfer maps of TransferOpts.
//
//   facade_conserve <dim> <n> <metric.f64> <out_prefix>
//
// writes <out_prefix>.counts (int32 nents per dimension), .verts (int32),
// .coords, .density and .velocity (float64).
#include <cmath>
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
  if (argc != 5) {
    std::cerr << "usage: facade_conserve <dim> <n> <metric.f64> <out>\n";
    return 2;
  }
  int dim = std::atoi(argv[1]);
  int n = std::atoi(argv[2]);
  std::string prefix = argv[4];
  using namespace Omega_h;
  try {
    Library lib(&argc, &argv);
    Mesh mesh = build_box(lib.world(), OMEGA_H_SIMPLEX, 1.0, 1.0, dim == 3 ? 1.0 : 0.0, n, n,
        dim == 3 ? n : 0);
    std::size_t nv = std::size_t(mesh.nverts());
    mesh.add_tag(VERT, "metric", 1, to_reals(read_doubles(argv[3], nv)));
    HostRead<Real> coords(mesh.coords());
    HostRead<LO> ev2v(mesh.ask_elem_verts());
    HostWrite<Real> density(mesh.nelems());
    for (LO e = 0; e < mesh.nelems(); ++e) {
      Real y = 0;
      for (int k = 0; k <= dim; ++k) y += coords[ev2v[e * (dim + 1) + k] * dim + 1];
      density[e] = (y / (dim + 1) > 0.5) ? 1e-4 : 1.0;
    }
    mesh.add_tag(dim, "density", 1, Reals(density.write()));
    HostWrite<Real> velocity(mesh.nverts() * dim);
    for (LO v = 0; v < mesh.nverts(); ++v) {
      Real x2 = coords[v * dim] - 0.3;
      Real w = (x2 > 0 ? 1.0 : (x2 < 0 ? -1.0 : 0.0)) * std::sqrt(std::abs(x2));
      for (int k = 0; k < dim; ++k) velocity[v * dim + k] = (k == 0) ? w : 0.0;
    }
    mesh.add_tag(VERT, "velocity", dim, Reals(velocity.write()));
    AdaptOpts opts(&mesh);
    opts.verbosity = SILENT;
    opts.xfer_opts.type_map["density"] = OMEGA_H_CONSERVE;
    opts.xfer_opts.integral_map["density"] = "mass";
    opts.xfer_opts.type_map["velocity"] = OMEGA_H_MOMENTUM_VELOCITY;
    opts.xfer_opts.velocity_density_map["velocity"] = "density";
    opts.xfer_opts.velocity_momentum_map["velocity"] = "momentum";
    opts.xfer_opts.integral_diffuse_map["mass"] =
        VarCompareOpts{VarCompareOpts::RELATIVE, 0.9, 0.0};
    opts.xfer_opts.integral_diffuse_map["momentum"] =
        VarCompareOpts{VarCompareOpts::RELATIVE, 0.02, 1e-6};
    adapt(&mesh, opts);
    std::vector<int> counts;
    for (int d = 0; d <= dim; ++d) counts.push_back(mesh.nents(d));
    write_file(prefix + ".counts", counts.data(), counts.size());
    auto verts = mesh.ask_elem_verts();
    write_read(prefix + ".verts", verts);
    auto x = mesh.coords();
    write_read(prefix + ".coords", x);
    auto rho = mesh.get_array<Real>(dim, "density");
    write_read(prefix + ".density", rho);
    auto vel = mesh.get_array<Real>(VERT, "velocity");
    write_read(prefix + ".velocity", vel);
    std::printf("adapted to %d elements\n", mesh.nelems());
  } catch (Omega_h::exception const& e) {
    std::cerr << "Omega_h::exception: " << e.what() << "\n";
    return 1;
  }
  return 0;
}

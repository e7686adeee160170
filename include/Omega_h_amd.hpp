// C++ drop-in facade of the MI355X adapt() path (header only, over the C ABI
// of omega_h_amd.h).
//
// It keeps the names and call shapes Omega_h callers use on this path:
//   Omega_h::Library, Omega_h::Mesh           src/Omega_h_library.hpp,
//                                             src/Omega_h_mesh.hpp:36-170
//   Omega_h::Read<T> / Write<T> / HostWrite    src/Omega_h_array.hpp:21-204
//   Omega_h::AdaptOpts, Omega_h::adapt         src/Omega_h_adapt.hpp:44-82
//   Omega_h::build_box                         src/Omega_h_build.hpp:24
//   refine_by_size, coarsen_by_size, coarsen_slivers, swap_edges
// so that a call site such as src/aniso_test.cpp:15-27 compiles against it
// by changing the include.  The mesh lives on the GPU behind an osh_mesh
// handle; Read<T>/Write<T> live in device memory too (tags are read and set
// by device-to-device copies), and HostRead/HostWrite are their host
// mirrors.  Errors throw Omega_h::exception,
// the reference's behaviour under Omega_h_THROW (src/Omega_h_fail.hpp:33-83).
#ifndef OMEGA_H_AMD_HPP
#define OMEGA_H_AMD_HPP

#include <cstdint>
#include <initializer_list>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "omega_h_amd.h"

// C enums of src/Omega_h_defines.h, at global scope as there
enum Omega_h_Family { OMEGA_H_SIMPLEX = 0 };
// transfer types of tags carried through adapt() (src/Omega_h_defines.hpp:29-37,
// same order)
enum Omega_h_Transfer {
  OMEGA_H_INHERIT,
  OMEGA_H_LINEAR_INTERP,
  OMEGA_H_METRIC,
  OMEGA_H_DENSITY,
  OMEGA_H_CONSERVE,
  OMEGA_H_MOMENTUM_VELOCITY,
  OMEGA_H_POINTWISE,
};
// how a metric is made isotropic
enum Omega_h_Isotropy { OMEGA_H_ANISOTROPIC = 0, OMEGA_H_ISO_LENGTH = 1, OMEGA_H_ISO_SIZE = 2 };
// how a distributed mesh is partitioned; one rank holds the whole mesh, so
// every choice is the same here
enum Omega_h_Parting { OMEGA_H_ELEM_BASED = 0, OMEGA_H_GHOSTED = 1, OMEGA_H_VERT_BASED = 2 };

namespace Omega_h {

using Int = std::int32_t;
using LO = std::int32_t;
using GO = std::int64_t;
using I8 = std::int8_t;
using Real = double;

enum { VERT = 0, EDGE = 1, FACE = 2, REGION = 3 };

enum Verbosity { SILENT = 0, EACH_ADAPT = 1, EACH_REBUILD = 2, EXTRA_STATS = 3 };


struct exception : public std::runtime_error {
  explicit exception(std::string const& msg) : std::runtime_error(msg) {}
};

inline void check_status(int status) {
  if (status != 0) throw exception(osh_last_error());
}

namespace detail {
// one context (HIP stream) of a Library; arrays and meshes hold it, so it
// closes after the last of them
struct Device {
  osh_ctx* ctx = nullptr;
  explicit Device(int device) { check_status(osh_ctx_create(device, &ctx)); }
  ~Device() {
    if (ctx) osh_ctx_destroy(ctx);
  }
  Device(Device const&) = delete;
  Device& operator=(Device const&) = delete;
};
// the context new arrays are allocated on: the most recent live Library's
inline std::shared_ptr<Device>& current_device() {
  static std::shared_ptr<Device> d;
  return d;
}
inline std::shared_ptr<Device> live_device() {
  std::shared_ptr<Device> d = current_device();
  if (!d) throw exception("Omega_h arrays live on the GPU of an Omega_h::Library; none is open");
  return d;
}
// a block of device memory, shared by the arrays that view it
struct Block {
  std::shared_ptr<Device> dev;
  void* ptr = nullptr;
  long long bytes = 0;
  Block(std::shared_ptr<Device> d, long long b) : dev(std::move(d)), bytes(b) {
    if (b > 0) check_status(osh_device_alloc(dev->ctx, b, &ptr));
  }
  ~Block() {
    if (ptr) osh_device_free(dev->ctx, ptr, bytes);
  }
  Block(Block const&) = delete;
  Block& operator=(Block const&) = delete;
};
template <class T>
std::shared_ptr<Block> device_block(LO n) {
  if (n < 0) throw exception("negative array size");
  return std::make_shared<Block>(live_device(), (long long)n * (long long)sizeof(T));
}
inline void copy_bytes(Block const& b, void* dst, void const* src, long long bytes) {
  if (bytes > 0) check_status(osh_copy(b.dev->ctx, dst, src, bytes));
}
template <class T>
void fill_items(Block const& b, LO n, T value) {
  static_assert(sizeof(T) == 1 || sizeof(T) == 4 || sizeof(T) == 8, "items of 1, 4 or 8 bytes");
  if (n > 0) check_status(osh_device_fill(b.dev->ctx, b.ptr, n, int(sizeof(T)), &value));
}
}  // namespace detail

// Write<T>/Read<T> (src/Omega_h_array.hpp:21-154) live in device memory of
// the open Library's GPU, as the reference's do in a GPU build: data() is a
// device pointer, copies share the block, and set/get/last move one item.
// HostRead/HostWrite are the host mirrors for host-side loops.
template <class T>
class Write {
 public:
  Write() = default;
  explicit Write(LO size, std::string const& = "")
      : block_(detail::device_block<T>(size)), size_(size) {}
  Write(LO size, T value, std::string const& = "") : Write(size) {
    detail::fill_items(*block_, size, value);
  }
  Write(LO size, T offset, T stride, std::string const& = "") : Write(size) {
    std::vector<T> h(static_cast<std::size_t>(size));
    for (LO i = 0; i < size; ++i) h[std::size_t(i)] = T(offset + stride * T(i));
    detail::copy_bytes(*block_, data(), h.data(), bytes());
  }
  Write(std::initializer_list<T> l, std::string const& = "") : Write(LO(l.size())) {
    detail::copy_bytes(*block_, data(), l.begin(), bytes());
  }
  LO size() const { return size_; }
  bool exists() const { return bool(block_); }
  T* data() const { return block_ ? static_cast<T*>(block_->ptr) : nullptr; }
  // the context (stream) the array's block belongs to
  osh_ctx* context() const { return block_ ? block_->dev->ctx : nullptr; }
  void set(LO i, T v) const {
    check_index(i);
    detail::copy_bytes(*block_, data() + i, &v, sizeof(T));
  }
  T get(LO i) const {
    check_index(i);
    T v;
    detail::copy_bytes(*block_, &v, data() + i, sizeof(T));
    return v;
  }

 private:
  template <class U>
  friend class Read;
  void check_index(LO i) const {
    if (i < 0 || i >= size_) throw exception("array index out of range");
  }
  long long bytes() const { return (long long)size_ * (long long)sizeof(T); }
  std::shared_ptr<detail::Block> block_;
  LO size_ = 0;
};

template <class T>
class Read {
 public:
  Read() = default;
  Read(Write<T> w) : w_(std::move(w)) {}
  Read(std::initializer_list<T> l, std::string const& name = "") : w_(l, name) {}
  Read(LO size, T value, std::string const& name = "") : w_(size, value, name) {}
  Read(LO size, T offset, T stride, std::string const& name = "")
      : w_(size, offset, stride, name) {}
  LO size() const { return w_.size(); }
  bool exists() const { return w_.exists(); }
  T const* data() const { return w_.data(); }
  osh_ctx* context() const { return w_.context(); }
  T get(LO i) const { return w_.get(i); }
  T last() const { return w_.get(size() - 1); }

 private:
  Write<T> w_;
};

// host mirror of a Read (src/Omega_h_array.hpp:155-188)
template <class T>
class HostRead {
 public:
  HostRead() = default;
  HostRead(Read<T> r) : h_(std::size_t(r.size())) {
    if (r.size())
      check_status(osh_copy(r.context(), h_.data(), r.data(), (long long)h_.size() * sizeof(T)));
  }
  LO size() const { return LO(h_.size()); }
  T const& operator[](LO i) const { return h_[std::size_t(i)]; }
  T const* data() const { return h_.data(); }
  T get(LO i) const { return h_.at(std::size_t(i)); }
  T last() const { return h_.back(); }

 private:
  std::vector<T> h_;
};

// host array written on the host and sent to the device with write()
// (src/Omega_h_array.hpp:190-230)
template <class T>
class HostWrite {
 public:
  HostWrite() = default;
  explicit HostWrite(LO size, std::string const& = "") : h_(std::size_t(size)) {}
  HostWrite(LO size, T offset, T stride, std::string const& = "") : h_(std::size_t(size)) {
    for (LO i = 0; i < size; ++i) h_[std::size_t(i)] = T(offset + stride * T(i));
  }
  HostWrite(std::initializer_list<T> l, std::string const& = "") : h_(l) {}
  HostWrite(Write<T> w) : w_(w), h_(std::size_t(w.size())) {
    if (w.size())
      check_status(osh_copy(w.context(), h_.data(), w.data(), (long long)h_.size() * sizeof(T)));
  }
  // the device array holding what was written on the host (the array this
  // mirror was made from, else a new one on the open Library's GPU)
  Write<T> write() const {
    if (!w_.exists() || w_.size() != size()) w_ = Write<T>(size());
    if (size())
      check_status(osh_copy(w_.context(), w_.data(), h_.data(), (long long)h_.size() * sizeof(T)));
    return w_;
  }
  LO size() const { return LO(h_.size()); }
  T& operator[](LO i) { return h_[std::size_t(i)]; }
  T const& operator[](LO i) const { return h_[std::size_t(i)]; }
  T* data() { return h_.data(); }
  bool exists() const { return true; }
  void set(LO i, T v) { h_.at(std::size_t(i)) = v; }
  T get(LO i) const { return h_.at(std::size_t(i)); }

 private:
  mutable Write<T> w_;
  std::vector<T> h_;
};

using LOs = Read<LO>;
using GOs = Read<GO>;
using Reals = Read<Real>;
using Bytes = Read<I8>;

// up/down adjacency (src/Omega_h_adj.hpp:11-49): down has ab2b and codes,
// up adds the a2ab offsets
struct Adj {
  LOs a2ab;
  LOs ab2b;
  Bytes codes;
};
struct Graph {
  LOs a2ab;
  LOs ab2b;
};

class Library;

// stands in for the communicator of build_box(lib.world(), ...): one rank
struct Comm {
  Library* library;
  int rank() const { return 0; }
  int size() const { return 1; }
};

using CommPtr = std::shared_ptr<Comm>;

// one HIP stream on one GPU (the device is the only argument that matters);
// arrays made while it is open live on its GPU.  The reference's --osh-time
// options are the environment variables OSH_TIME, OSH_TIME_PERCENT and
// OSH_TIME_CHOP here.
class Library {
 public:
  explicit Library(int* argc = nullptr, char*** argv = nullptr, int device = 0)
      : dev_(std::make_shared<detail::Device>(device)) {
    (void)argc;
    (void)argv;
    detail::current_device() = dev_;
  }
  Library(Library const&) = delete;
  Library& operator=(Library const&) = delete;
  ~Library() {
    if (detail::current_device() == dev_) detail::current_device().reset();
  }
  osh_ctx* ctx() const { return dev_->ctx; }
  std::shared_ptr<detail::Device> const& device() const { return dev_; }
  CommPtr world() { return std::make_shared<Comm>(Comm{this}); }
  CommPtr self() { return world(); }

 private:
  std::shared_ptr<detail::Device> dev_;
};

class Mesh {
 public:
  Mesh() = default;
  explicit Mesh(Library* library) : library_(library), dev_(library->device()) {}
  // wraps a handle this object now owns
  Mesh(Library* library, osh_mesh* h)
      : library_(library), dev_(library->device()), h_(own(h)) {}
  // copies share the device arrays; passes build new ones (like assigning
  // an Omega_h::Mesh)
  Mesh(Mesh const& o) : library_(o.library_), dev_(o.dev_) {
    if (o.h_) {
      osh_mesh* c = nullptr;
      check_status(osh_mesh_copy(o.h_.get(), &c));
      h_ = own(c);
    }
  }
  Mesh& operator=(Mesh const& o) {
    if (this != &o) {
      Mesh t(o);
      library_ = t.library_;
      dev_ = t.dev_;
      h_ = std::move(t.h_);
    }
    return *this;
  }
  Mesh(Mesh&&) = default;
  Mesh& operator=(Mesh&&) = default;

  Library* library() const { return library_; }
  osh_mesh* handle() const {
    if (!h_) throw exception("mesh has no entities yet");
    return h_.get();
  }
  Omega_h_Family family() const { return OMEGA_H_SIMPLEX; }
  CommPtr comm() const { return std::make_shared<Comm>(Comm{library_}); }
  Omega_h_Parting parting() const { return OMEGA_H_ELEM_BASED; }
  void set_parting(Omega_h_Parting, Int nlayers = 0, bool verbose = false) {
    (void)nlayers;
    (void)verbose;
  }
  Int dim() const {
    int d;
    check_status(osh_mesh_dim(handle(), &d));
    return d;
  }
  LO nents(Int ent_dim) const {
    int n;
    check_status(osh_mesh_nents(handle(), ent_dim, &n));
    return n;
  }
  LO nverts() const { return nents(VERT); }
  LO nedges() const { return nents(EDGE); }
  LO nfaces() const { return nents(FACE); }
  LO nregions() const { return nents(REGION); }
  LO nelems() const { return nents(dim()); }
  GO nglobal_ents(Int ent_dim) const { return nents(ent_dim); }
  LO nents_owned(Int ent_dim) const { return nents(ent_dim); }

  Reals coords() const {
    Write<Real> w(nverts() * dim());
    check_status(osh_mesh_get_coords(handle(), w.data()));
    return w;
  }
  void set_coords(Reals const& array) {
    if (array.size() != nverts() * dim()) throw exception("coordinates have the wrong size");
    check_status(osh_mesh_set_coords(handle(), array.data()));
  }
  // one rank: the global number of an entity is its index
  Read<GO> globals(Int ent_dim) const { return Read<GO>(nents(ent_dim), GO(0), GO(1)); }
  LOs ask_verts_of(Int ent_dim) const {
    Write<LO> w(nents(ent_dim) * (ent_dim + 1));
    check_status(osh_mesh_get_verts(handle(), ent_dim, w.data()));
    return w;
  }
  LOs ask_elem_verts() const { return ask_verts_of(dim()); }
  Adj ask_down(Int from, Int to) const {
    if (to == VERT) return Adj{LOs(), ask_verts_of(from), Bytes()};
    if (to != from - 1) throw exception("ask_down: only one dimension down (or to vertices)");
    LO n = nents(from) * (from + 1);
    Write<LO> ab2b(n);
    Write<I8> codes(n);
    check_status(osh_mesh_get_down(handle(), from, ab2b.data(),
        reinterpret_cast<signed char*>(codes.data())));
    return Adj{LOs(), ab2b, codes};
  }
  Adj ask_up(Int from, Int to) const {
    int n;
    check_status(osh_mesh_up_size(handle(), from, to, &n));
    Write<LO> a2ab(nents(from) + 1), ab2b(n);
    Write<I8> codes(n);
    check_status(osh_mesh_get_up(handle(), from, to, a2ab.data(), ab2b.data(),
        reinterpret_cast<signed char*>(codes.data())));
    return Adj{a2ab, ab2b, codes};
  }
  Graph ask_star(Int ent_dim) const {
    int n;
    check_status(osh_mesh_star_size(handle(), ent_dim, &n));
    Write<LO> a2ab(nents(ent_dim) + 1), ab2b(n);
    check_status(osh_mesh_get_star(handle(), ent_dim, a2ab.data(), ab2b.data()));
    return Graph{a2ab, ab2b};
  }
  Graph ask_dual() const {
    int n;
    check_status(osh_mesh_dual_size(handle(), &n));
    Write<LO> a2ab(nelems() + 1), ab2b(n);
    check_status(osh_mesh_get_dual(handle(), a2ab.data(), ab2b.data()));
    return Graph{a2ab, ab2b};
  }
  Reals ask_sizes() const {
    Write<Real> w(nelems());
    check_status(osh_mesh_get_sizes(handle(), w.data()));
    return w;
  }
  Reals ask_lengths() const {
    Write<Real> w(nedges());
    check_status(osh_mesh_get_lengths(handle(), w.data()));
    return w;
  }
  Reals ask_qualities() const {
    Write<Real> w(nelems());
    check_status(osh_mesh_get_qualities(handle(), w.data()));
    return w;
  }
  Real min_quality() const {
    Real r;
    check_status(osh_mesh_min_quality(handle(), &r));
    return r;
  }
  Real max_length() const {
    Real r;
    check_status(osh_mesh_max_length(handle(), &r));
    return r;
  }

  // tags on this path: "coordinates", "metric", "class_id", "class_dim",
  // "length" (edges), "quality" (elements), and caller vertex tags carried
  // through adapt() ("target_metric" and any field transferred by linear
  // interpolation or as a metric)
  int vertex_tag_ncomps(std::string const& name) const {
    int nc;
    check_status(osh_mesh_vertex_tag_ncomps(handle(), name.c_str(), &nc));
    return nc;
  }
  // element tags are carried through adapt() pointwise (OMEGA_H_POINTWISE)
  int elem_tag_ncomps(std::string const& name) const {
    int nc;
    check_status(osh_mesh_elem_tag_ncomps(handle(), name.c_str(), &nc));
    return nc;
  }
  bool has_tag(Int ent_dim, std::string const& name) const {
    if (ent_dim == VERT && name != "metric" && name != "coordinates" &&
        vertex_tag_ncomps(name) >= 0)
      return true;
    if (name == "coordinates") return ent_dim == VERT;
    if (name == "metric") {
      int nc;
      check_status(osh_mesh_metric_ncomps(handle(), &nc));
      return ent_dim == VERT && nc > 0;
    }
    if (name == "class_id" || name == "class_dim") return ent_dim <= dim();
    if (name == "length") return ent_dim == EDGE;
    if (name == "quality") return ent_dim == dim();
    if (ent_dim == dim() && elem_tag_ncomps(name) >= 0) return true;
    return false;
  }
  void remove_tag(Int ent_dim, std::string const& name) {
    if (ent_dim == VERT) check_status(osh_mesh_remove_vertex_tag(handle(), name.c_str()));
    if (ent_dim == dim()) check_status(osh_mesh_remove_elem_tag(handle(), name.c_str()));
  }
  template <class T>
  void add_tag(Int ent_dim, std::string const& name, Int ncomps, Read<T> array,
      bool internal = false);
  template <class T>
  void set_tag(Int ent_dim, std::string const& name, Read<T> array, bool internal = false);
  template <class T>
  Read<T> get_array(Int ent_dim, std::string const& name) const;

 private:
  // the mesh handle keeps its context open
  std::shared_ptr<osh_mesh> own(osh_mesh* h) const {
    std::shared_ptr<detail::Device> d = dev_;
    return std::shared_ptr<osh_mesh>(h, [d](osh_mesh* m) { osh_mesh_destroy(m); });
  }
  Library* library_ = nullptr;
  std::shared_ptr<detail::Device> dev_;
  std::shared_ptr<osh_mesh> h_;
};

namespace detail {
// "metric" is the mesh's metric; other vertex tags are carried through
// adapt() as metrics when their name ends in "metric" ("target_metric"),
// otherwise by linear interpolation (AdaptOpts::xfer_opts can say otherwise)
inline void set_metric(Mesh& m, Int ent_dim, std::string const& name, Int ncomps,
    Real const* data, LO size) {
  if (ent_dim == m.dim() && ent_dim != VERT) {
    if (size != m.nelems() * ncomps) throw exception("tag array has the wrong size");
    check_status(osh_mesh_add_elem_tag(m.handle(), name.c_str(), ncomps, data));
    return;
  }
  if (ent_dim != VERT)
    throw exception("only vertex and element tags are carried through adapt()");
  if (size != m.nverts() * ncomps) throw exception("tag array has the wrong size");
  if (name == "metric") {
    check_status(osh_mesh_set_metric(m.handle(), ncomps, data));
    return;
  }
  if (name == "coordinates") {
    if (ncomps != m.dim()) throw exception("coordinates need dim components");
    check_status(osh_mesh_set_coords(m.handle(), data));
    return;
  }
  bool is_metric = name.size() >= 6 && name.compare(name.size() - 6, 6, "metric") == 0;
  check_status(osh_mesh_add_vertex_tag(m.handle(), name.c_str(), ncomps,
      is_metric ? 1 : 0, data));
}
}  // namespace detail

template <>
inline void Mesh::add_tag<Real>(Int ent_dim, std::string const& name, Int ncomps,
    Read<Real> array, bool) {
  detail::set_metric(*this, ent_dim, name, ncomps, array.data(), array.size());
}

template <>
inline void Mesh::set_tag<Real>(Int ent_dim, std::string const& name, Read<Real> array, bool) {
  int nc = (nverts() > 0) ? array.size() / nverts() : 0;
  if (ent_dim == VERT && name != "metric" && vertex_tag_ncomps(name) > 0)
    nc = vertex_tag_ncomps(name);
  detail::set_metric(*this, ent_dim, name, nc, array.data(), array.size());
}

template <>
inline Read<Real> Mesh::get_array<Real>(Int ent_dim, std::string const& name) const {
  if (name == "coordinates" && ent_dim == VERT) return coords();
  if (name == "metric" && ent_dim == VERT) {
    int nc;
    check_status(osh_mesh_metric_ncomps(handle(), &nc));
    Write<Real> w(nverts() * nc);
    check_status(osh_mesh_get_metric(handle(), w.data()));
    return w;
  }
  if (name == "length" && ent_dim == EDGE) return ask_lengths();
  if (name == "quality" && ent_dim == dim()) return ask_qualities();
  if (ent_dim == VERT) {
    int nc = vertex_tag_ncomps(name);
    if (nc >= 0) {
      Write<Real> w(nverts() * nc);
      check_status(osh_mesh_get_vertex_tag(handle(), name.c_str(), w.data()));
      return w;
    }
  }
  if (ent_dim == dim()) {
    int nc = elem_tag_ncomps(name);
    if (nc >= 0) {
      Write<Real> w(nelems() * nc);
      check_status(osh_mesh_get_elem_tag(handle(), name.c_str(), w.data()));
      return w;
    }
  }
  throw exception("no Real tag \"" + name + "\" on dimension " + std::to_string(ent_dim));
}

template <>
inline Read<LO> Mesh::get_array<LO>(Int ent_dim, std::string const& name) const {
  if (name != "class_id") throw exception("no LO tag \"" + name + "\"");
  Write<LO> ids(nents(ent_dim));
  Write<I8> dims(nents(ent_dim));
  check_status(osh_mesh_get_class(handle(), ent_dim, ids.data(),
      reinterpret_cast<signed char*>(dims.data())));
  return ids;
}

template <>
inline Read<I8> Mesh::get_array<I8>(Int ent_dim, std::string const& name) const {
  if (name != "class_dim") throw exception("no I8 tag \"" + name + "\"");
  Write<LO> ids(nents(ent_dim));
  Write<I8> dims(nents(ent_dim));
  check_status(osh_mesh_get_class(handle(), ent_dim, ids.data(),
      reinterpret_cast<signed char*>(dims.data())));
  return dims;
}

// src/Omega_h_build.hpp:24; nz == 0 builds a 2D box (as the reference does
// when z == 0)
inline Mesh build_box(CommPtr comm, Omega_h_Family family, Real x, Real y, Real z, LO nx,
    LO ny, LO nz, bool symmetric = false) {
  if (family != OMEGA_H_SIMPLEX) throw exception("only simplex meshes are supported");
  if (symmetric) throw exception("symmetric boxes are not supported");
  Library* lib = comm->library;
  int dim = (nz > 0) ? 3 : 2;
  osh_mesh* h = nullptr;
  check_status(osh_build_box(lib->ctx(), dim, x, y, z, nx, ny, nz, &h));
  return Mesh(lib, h);
}

// src/Omega_h_compare.hpp:14-21
struct VarCompareOpts {
  enum { NONE, RELATIVE, ABSOLUTE } type = RELATIVE;
  Real tolerance = 1e-6;
  Real floor = 0.0;
};

// src/Omega_h_adapt.hpp:30-42: the transfer type of each caller tag, and the
// maps of conservative and momentum transfer
struct TransferOpts {
  std::map<std::string, Omega_h_Transfer> type_map;              // "density" -> CONSERVE
  std::map<std::string, std::string> integral_map;               // "density" -> "mass"
  std::map<std::string, std::string> velocity_density_map;       // "velocity" -> "density"
  std::map<std::string, std::string> velocity_momentum_map;      // "velocity" -> "momentum"
  std::map<std::string, VarCompareOpts> integral_diffuse_map;    // "mass" -> tolerance
};

// src/Omega_h_adapt.hpp:50-77, defaults src/Omega_h_adapt.cpp:52-85
struct AdaptOpts {
  AdaptOpts() = default;
  explicit AdaptOpts(Int dim) {
    osh_adapt_opts o;
    check_status(osh_default_opts(dim, &o));
    from_c(o);
  }
  explicit AdaptOpts(Mesh* mesh) : AdaptOpts(mesh->dim()) {}
  Real min_length_desired = 0;
  Real max_length_desired = 0;
  Real max_length_allowed = 0;
  Real min_quality_allowed = 0;
  Real min_quality_desired = 0;
  Int nsliver_layers = 0;
  Verbosity verbosity = SILENT;
  bool should_refine = true;
  bool should_coarsen = true;
  bool should_swap = true;
  bool should_coarsen_slivers = true;
  Real length_histogram_min = 0.0;
  Real length_histogram_max = 3.0;
  Int nlength_histogram_bins = 10;
  Int nquality_histogram_bins = 10;
  // prevent_coarsen_flip, src/Omega_h_coarsen_flip.cpp:68-150
  bool should_prevent_coarsen_flip = false;
  // no counterpart in the reference: OSH_FP64 or OSH_MIXED cavity screens
  Int precision = OSH_FP64;
  TransferOpts xfer_opts;

  osh_adapt_opts to_c() const {
    osh_adapt_opts o;
    o.min_length_desired = min_length_desired;
    o.max_length_desired = max_length_desired;
    o.max_length_allowed = max_length_allowed;
    o.min_quality_allowed = min_quality_allowed;
    o.min_quality_desired = min_quality_desired;
    o.nsliver_layers = nsliver_layers;
    o.verbosity = int(verbosity);
    o.should_refine = should_refine;
    o.should_coarsen = should_coarsen;
    o.should_swap = should_swap;
    o.should_coarsen_slivers = should_coarsen_slivers;
    o.length_histogram_min = length_histogram_min;
    o.length_histogram_max = length_histogram_max;
    o.nlength_histogram_bins = nlength_histogram_bins;
    o.nquality_histogram_bins = nquality_histogram_bins;
    o.should_prevent_coarsen_flip = should_prevent_coarsen_flip;
    o.precision = precision;
    return o;
  }
  void from_c(osh_adapt_opts const& o) {
    min_length_desired = o.min_length_desired;
    max_length_desired = o.max_length_desired;
    max_length_allowed = o.max_length_allowed;
    min_quality_allowed = o.min_quality_allowed;
    min_quality_desired = o.min_quality_desired;
    nsliver_layers = o.nsliver_layers;
    verbosity = Verbosity(o.verbosity);
    should_refine = o.should_refine != 0;
    should_coarsen = o.should_coarsen != 0;
    should_swap = o.should_swap != 0;
    should_coarsen_slivers = o.should_coarsen_slivers != 0;
    length_histogram_min = o.length_histogram_min;
    length_histogram_max = o.length_histogram_max;
    nlength_histogram_bins = o.nlength_histogram_bins;
    nquality_histogram_bins = o.nquality_histogram_bins;
    should_prevent_coarsen_flip = o.should_prevent_coarsen_flip != 0;
    precision = o.precision;
  }
};

namespace detail {
// the diffusion of a conserved integral's error, by the integral's name
inline VarCompareOpts diffusion_of(TransferOpts const& x, std::string const& integral) {
  auto it = x.integral_diffuse_map.find(integral);
  if (it == x.integral_diffuse_map.end())
    throw exception("no integral_diffuse_map entry for \"" + integral + "\"");
  return it->second;
}
inline std::string mapped(std::map<std::string, std::string> const& m, std::string const& k,
    char const* what) {
  auto it = m.find(k);
  if (it == m.end()) throw exception(std::string("no ") + what + " entry for \"" + k + "\"");
  return it->second;
}
// give the caller's tags the transfers the options name: vertex tags are
// re-registered as linear or metric, element tags marked conserved or
// intersection-transferred, velocities marked as momentum velocities (after
// their densities)
inline void apply_transfer_types(Mesh* mesh, AdaptOpts const& opts) {
  auto const& x = opts.xfer_opts;
  for (auto const& kv : x.type_map) {
    if (kv.second == OMEGA_H_POINTWISE) {
      if (mesh->elem_tag_ncomps(kv.first) >= 0)
        check_status(osh_mesh_set_elem_transfer(mesh->handle(), kv.first.c_str(), 0, 0, 0, 0));
      continue;
    }
    if (kv.second == OMEGA_H_INHERIT) continue;  // inheritance tags are inherited already
    if (kv.second == OMEGA_H_CONSERVE || kv.second == OMEGA_H_DENSITY) {
      VarCompareOpts d;
      d.type = VarCompareOpts::NONE;
      if (kv.second == OMEGA_H_CONSERVE)
        d = diffusion_of(x, mapped(x.integral_map, kv.first, "integral_map"));
      check_status(osh_mesh_set_elem_transfer(mesh->handle(), kv.first.c_str(),
          kv.second == OMEGA_H_CONSERVE ? 1 : 2, int(d.type), d.tolerance, d.floor));
      continue;
    }
    if (kv.second == OMEGA_H_MOMENTUM_VELOCITY) continue;  // below, after the densities
    if (kv.second != OMEGA_H_LINEAR_INTERP && kv.second != OMEGA_H_METRIC)
      throw exception("transfer type of \"" + kv.first + "\" is not supported");
    int nc = mesh->vertex_tag_ncomps(kv.first);
    if (nc < 0) continue;
    Write<Real> w(mesh->nverts() * nc);
    check_status(osh_mesh_get_vertex_tag(mesh->handle(), kv.first.c_str(), w.data()));
    // the C ABI's vertex tag kinds: 0 linear interpolation, 1 metric
    check_status(osh_mesh_add_vertex_tag(mesh->handle(), kv.first.c_str(), nc,
        kv.second == OMEGA_H_METRIC ? 1 : 0, w.data()));
  }
  for (auto const& kv : x.type_map) {
    if (kv.second != OMEGA_H_MOMENTUM_VELOCITY) continue;
    auto density = mapped(x.velocity_density_map, kv.first, "velocity_density_map");
    auto momentum = mapped(x.velocity_momentum_map, kv.first, "velocity_momentum_map");
    VarCompareOpts d = diffusion_of(x, momentum);
    check_status(osh_mesh_set_momentum_velocity(mesh->handle(), kv.first.c_str(),
        density.c_str(), int(d.type), d.tolerance, d.floor));
  }
}
}  // namespace detail

// src/Omega_h_adapt.cpp:272-288: false iff the mesh already met the goals
inline bool adapt(Mesh* mesh, AdaptOpts const& opts) {
  detail::apply_transfer_types(mesh, opts);
  osh_adapt_opts o = opts.to_c();
  int did = 0;
  check_status(osh_adapt(mesh->handle(), &o, nullptr, &did));
  return did != 0;
}

namespace detail {
inline bool run_pass(int (*f)(osh_mesh*, const osh_adapt_opts*, int*), Mesh* mesh,
    AdaptOpts const& opts) {
  osh_adapt_opts o = opts.to_c();
  int did = 0;
  check_status(f(mesh->handle(), &o, &did));
  return did != 0;
}
}  // namespace detail

inline bool refine_by_size(Mesh* mesh, AdaptOpts const& opts) {
  return detail::run_pass(osh_refine_by_size, mesh, opts);
}
inline bool coarsen_by_size(Mesh* mesh, AdaptOpts const& opts) {
  return detail::run_pass(osh_coarsen_by_size, mesh, opts);
}
inline bool coarsen_slivers(Mesh* mesh, AdaptOpts const& opts) {
  return detail::run_pass(osh_coarsen_slivers, mesh, opts);
}
inline bool swap_edges(Mesh* mesh, AdaptOpts const& opts) {
  return detail::run_pass(osh_swap_edges, mesh, opts);
}

// src/Omega_h_approach.cpp:83-126 (declared src/Omega_h_adapt.hpp:90-92)
inline bool approach_metric(Mesh* mesh, AdaptOpts const& opts, Real min_step = 1e-4) {
  osh_adapt_opts o = opts.to_c();
  int did = 0;
  check_status(osh_approach_metric(mesh->handle(), &o, min_step, &did));
  return did != 0;
}

inline bool warp_to_limit(Mesh* mesh, AdaptOpts const& opts, bool exit_on_stall = false,
    Int max_niters = 40) {
  osh_adapt_opts o = opts.to_c();
  int did = 0;
  check_status(osh_warp_to_limit(mesh->handle(), &o, exit_on_stall ? 1 : 0, max_niters, &did));
  return did != 0;
}

// get_implied_metrics / get_implied_isos (src/Omega_h_metric.cpp:304-321)
// and add_implied_metric_tag / add_implied_isos_tag
// (src/Omega_h_metric_input.cpp:271-280)
inline Reals get_implied_metrics(Mesh* mesh) {
  Write<Real> w(mesh->nverts() * (mesh->dim() == 3 ? 6 : 3));
  check_status(osh_get_implied_metrics(mesh->handle(), 0, w.data()));
  return w;
}
inline Reals get_implied_isos(Mesh* mesh) {
  Write<Real> w(mesh->nverts());
  check_status(osh_get_implied_isos(mesh->handle(), 0, w.data()));
  return w;
}
inline void add_implied_metric_tag(Mesh* mesh) {
  check_status(osh_get_implied_metrics(mesh->handle(), 1, nullptr));
}
inline void add_implied_isos_tag(Mesh* mesh) {
  check_status(osh_get_implied_isos(mesh->handle(), 1, nullptr));
}

// limit_metric_gradation(mesh, values, max_rate, tol), src/Omega_h_metric.cpp:
// 199-220, for any vertex metric array (the mesh's metric is left as it was)
inline Reals limit_metric_gradation(Mesh* mesh, Reals values, Real max_rate, Real tol = 1e-2,
    bool verbose = true) {
  (void)verbose;
  LO nv = mesh->nverts();
  if (nv == 0 || values.size() % nv != 0) throw exception("metric array has the wrong size");
  int nc = values.size() / nv;
  int had_nc;
  check_status(osh_mesh_metric_ncomps(mesh->handle(), &had_nc));
  Write<Real> had(nv * had_nc);
  if (had_nc > 0) check_status(osh_mesh_get_metric(mesh->handle(), had.data()));
  check_status(osh_mesh_set_metric(mesh->handle(), nc, values.data()));
  int steps;
  check_status(osh_limit_metric_gradation(mesh->handle(), max_rate, tol, &steps));
  Write<Real> out(values.size());
  check_status(osh_mesh_get_metric(mesh->handle(), out.data()));
  if (had_nc > 0) check_status(osh_mesh_set_metric(mesh->handle(), had_nc, had.data()));
  return out;
}

// fix, fix_for_given_metric, grade_fix_adapt (src/Omega_h_fix.cpp:90-176,
// declared src/Omega_h_adapt.hpp:134-140)
inline void fix(Mesh* mesh, AdaptOpts const& adapt_opts, Omega_h_Isotropy isotropy,
    bool verbose) {
  (void)verbose;
  osh_adapt_opts o = adapt_opts.to_c();
  check_status(osh_fix(mesh->handle(), &o, int(isotropy), nullptr));
}
inline void fix_for_given_metric(Mesh* mesh, AdaptOpts const& adapt_opts, bool verbose) {
  (void)verbose;
  osh_adapt_opts o = adapt_opts.to_c();
  check_status(osh_fix_for_given_metric(mesh->handle(), &o, nullptr));
}
inline void grade_fix_adapt(Mesh* mesh, AdaptOpts const& opts, Reals target_metric,
    bool verbose) {
  (void)verbose;
  LO nv = mesh->nverts();
  if (nv == 0 || target_metric.size() % nv != 0)
    throw exception("target metric has the wrong size");
  osh_adapt_opts o = opts.to_c();
  check_status(osh_grade_fix_adapt(mesh->handle(), &o, target_metric.size() / nv,
      target_metric.data(), nullptr));
}

}  // namespace Omega_h

#endif

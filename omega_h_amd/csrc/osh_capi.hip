// The C ABI (include/omega_h_amd.h): status codes, thread-local error
// messages, handles over the device runtime.
#include <cstring>
#include <memory>
#include <string>

#include "../../include/omega_h_amd.h"
#include "osh_guard.hpp"
#include "osh_math.hpp"
#include "osh_passes.hpp"
#include "osh_swap_tables.hpp"

struct osh_ctx {
  osh::Ctx ctx;
};

struct osh_mesh {
  osh_ctx* owner;
  osh::DMesh m;
};

namespace osh {
thread_local std::string last_error_message;

void to_options(osh_adapt_opts const* o, AdaptOptions& r) {
  r.min_length_desired = o->min_length_desired;
  r.max_length_desired = o->max_length_desired;
  r.max_length_allowed = o->max_length_allowed;
  r.min_quality_allowed = o->min_quality_allowed;
  r.min_quality_desired = o->min_quality_desired;
  r.nsliver_layers = o->nsliver_layers;
  r.verbosity = o->verbosity;
  r.should_refine = o->should_refine;
  r.should_coarsen = o->should_coarsen;
  r.should_swap = o->should_swap;
  r.should_coarsen_slivers = o->should_coarsen_slivers;
  r.length_histogram_min = o->length_histogram_min;
  r.length_histogram_max = o->length_histogram_max;
  r.nlength_histogram_bins = o->nlength_histogram_bins;
  r.nquality_histogram_bins = o->nquality_histogram_bins;
  r.should_prevent_coarsen_flip = o->should_prevent_coarsen_flip;
  r.precision = o->precision;
  OSH_CHECK(r.precision == 0 || r.precision == 1, "precision must be OSH_FP64 or OSH_MIXED");
  OSH_CHECK(r.verbosity < 3 || (r.nlength_histogram_bins >= 1 && r.nquality_histogram_bins >= 1 &&
                                    r.length_histogram_max > r.length_histogram_min),
      "histogram options out of range");
}

namespace {
std::uint64_t fnv(std::uint64_t h, void const* p, std::size_t n) {
  auto b = static_cast<unsigned char const*>(p);
  for (std::size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= 1099511628211ull;
  }
  return h;
}
constexpr std::uint64_t kFnvBasis = 1469598103934665603ull;
}  // namespace

// four FNV-1a digests of a mesh: 0 the vertex lists of every dimension
// 1..dim (int32) then class_id (int32) and class_dim (int8) of every
// dimension 0..dim; 1 coordinate bits; 2 metric bits; 3 edge length bits
// then element quality bits
void mesh_digests(Ctx& c, DMesh& m, std::uint64_t* out) {
  std::uint64_t h = kFnvBasis;
  for (int d = 1; d <= m.dim; ++d) {
    auto v = to_host(c, verts_of(c, m, d));
    h = fnv(h, v.data(), v.size() * sizeof(LO));
  }
  for (int d = 0; d <= m.dim; ++d) {
    auto id = to_host(c, m.class_id[d]);
    auto cd = to_host(c, m.class_dim[d]);
    h = fnv(h, id.data(), id.size() * sizeof(LO));
    h = fnv(h, cd.data(), cd.size());
  }
  out[0] = h;
  auto x = to_host(c, m.coords);
  out[1] = fnv(kFnvBasis, x.data(), x.size() * sizeof(Real));
  auto mt = to_host(c, m.metric);
  out[2] = fnv(kFnvBasis, mt.data(), mt.size() * sizeof(Real));
  auto l = to_host(c, ask_lengths(c, m));
  auto q = to_host(c, ask_qualities(c, m));
  out[3] = fnv(fnv(kFnvBasis, l.data(), l.size() * sizeof(Real)), q.data(),
      q.size() * sizeof(Real));
}

// FNV-1a over the element vertex lists and the bits of the coordinates
std::uint64_t mesh_digest(Ctx& c, DMesh& m) {
  auto ev2v = to_host(c, verts_of(c, m, m.dim));
  auto coords = to_host(c, m.coords);
  std::uint64_t h = 1469598103934665603ull;
  auto mix = [&](void const* p, std::size_t n) {
    auto b = static_cast<unsigned char const*>(p);
    for (std::size_t i = 0; i < n; ++i) {
      h ^= b[i];
      h *= 1099511628211ull;
    }
  };
  mix(ev2v.data(), ev2v.size() * sizeof(LO));
  mix(coords.data(), coords.size() * sizeof(Real));
  return h;
}
}  // namespace osh

namespace {

using osh::guarded;
using osh::to_options;


// array arguments of the entry points may be host memory or device memory
// (a facade Read/Write, any hipMalloc'd or pool block): copies let the
// runtime tell them apart
template <class T>
void download(osh::Ctx& c, osh::DArray<T> const& d, T* out) {
  if (d.size())
    OSH_HIP(hipMemcpyAsync(out, d.data(), sizeof(T) * size_t(d.size()), hipMemcpyDefault,
        c.stream));
  c.sync();
}

// a host copy of a caller array (host or device) that host code reads
template <class T>
std::vector<T> host_copy(osh::Ctx& c, T const* p, std::size_t n) {
  std::vector<T> h(n);
  if (n) {
    OSH_HIP(hipMemcpyAsync(h.data(), p, sizeof(T) * n, hipMemcpyDefault, c.stream));
    c.sync();
  }
  return h;
}

template <class T>
__global__ void fill_items(long long n, T* p, T v) {
  long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <class T>
void fill_device(osh::Ctx& c, void* dst, long long n, void const* value) {
  T v;
  std::memcpy(&v, value, sizeof(T));
  long long blocks = (n + 255) / 256;
  OSH_CHECK(blocks <= 0x7fffffffLL, "fill: too many items");
  hipLaunchKernelGGL(fill_items<T>, dim3(unsigned(blocks)), dim3(256), 0, c.stream, n,
      static_cast<T*>(dst), v);
  OSH_HIP(hipGetLastError());
}

void check_mesh(osh_mesh* m) {
  OSH_CHECK(m != nullptr, "null mesh handle");
}

// what the passes read: a metric of 1 or dim*(dim+1)/2 components and both
// classification tags on every dimension (host checks, so a malformed mesh
// fails with a message before any kernel runs)
void check_metric(osh_mesh* m) {
  OSH_CHECK(m != nullptr, "null mesh handle");
  auto const& d = m->m;
  OSH_CHECK(d.metric.exists(), "mesh has no metric");
  OSH_CHECK(d.metric_ncomps == 1 || d.metric_ncomps == d.dim * (d.dim + 1) / 2,
      "metric must have 1 or dim*(dim+1)/2 components");
  OSH_CHECK(d.metric.size() == d.nents[0] * d.metric_ncomps, "metric has the wrong size");
}

void check_classified(osh_mesh* m) {
  OSH_CHECK(m != nullptr, "null mesh handle");
  auto const& d = m->m;
  for (int k = 0; k <= d.dim; ++k) {
    OSH_CHECK(d.class_id[k].exists() && d.class_dim[k].exists(), "mesh is not classified");
    OSH_CHECK(d.class_id[k].size() == d.nents[k] && d.class_dim[k].size() == d.nents[k],
        "classification has the wrong size");
  }
}

void check_adaptable(osh_mesh* m) {
  check_metric(m);
  check_classified(m);
}

void check_dim(osh_mesh* m, int d) {
  check_mesh(m);
  OSH_CHECK(d >= 0 && d <= m->m.dim, "entity dimension out of range");
}

__global__ void eval_math_kernel(osh::LO n, int which, double const* in, double* out) {
  osh::LO i = osh::thread_index();
  if (i >= n) return;
  double x = in[i];
  double r;
  switch (which) {
    case 0: r = osh::math::log_(x); break;
    case 1: r = osh::math::exp_(x); break;
    case 2: r = osh::math::cbrt_(x); break;
    case 3: r = osh::math::acos_(x); break;
    default: r = osh::math::cos_(x); break;
  }
  out[i] = r;
}

int* tuning_field(osh::Tuning& t, const char* name) {
  std::string n(name);
  if (n == "coarsen_lanes") return &t.coarsen_lanes;
  if (n == "side_lanes") return &t.side_lanes;
  if (n == "choose_lanes") return &t.choose_lanes;
  if (n == "first_lanes") return &t.first_lanes;
  throw osh::Error("unknown tuning setting " + n);
}

void fill_stats(std::vector<osh::PassRecord> const& log, osh_adapt_stats* stats) {
  stats->npasses = int(std::min<std::size_t>(log.size(), OSH_MAX_PASSES));
  stats->total_passes = static_cast<long long>(log.size());
  stats->total_nelems_in = 0;
  stats->total_seconds = 0.0;
  for (std::size_t i = 0; i < log.size(); ++i) {
    if (i < std::size_t(stats->npasses)) {
      stats->kind[i] = log[i].kind;
      stats->nelems_in[i] = log[i].nelems_in;
      stats->nelems_out[i] = log[i].nelems_out;
      stats->seconds[i] = log[i].seconds;
    }
    stats->total_nelems_in += log[i].nelems_in;
    stats->total_seconds += log[i].seconds;
  }
}

osh::Diffusion diffusion_of(int type, double tol, double floor) {
  OSH_CHECK(type >= 0 && type <= 2, "diffuse_type must be 0 (none), 1 or 2");
  OSH_CHECK(tol >= 0.0 && floor >= 0.0, "negative diffusion tolerance or floor");
  osh::Diffusion d;
  d.type = type;
  d.tol = tol;
  d.floor = floor;
  return d;
}

}  // namespace

extern "C" {

const char* osh_last_error(void) { return osh::last_error_message.c_str(); }

int osh_write_osh(osh_mesh* mesh, const char* path) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(path != nullptr, "null path");
    osh::write_osh(mesh->owner->ctx, mesh->m, path);
  });
}

int osh_read_osh(osh_ctx* ctx, const char* path, osh_mesh** out) {
  return guarded([&] {
    OSH_CHECK(ctx && path && out, "null argument");
    auto m = std::make_unique<osh_mesh>();
    m->owner = ctx;
    osh::read_osh(ctx->ctx, m->m, path);
    *out = m.release();
  });
}

int osh_mesh_digest(osh_mesh* mesh, unsigned long long* out) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(out != nullptr, "null output pointer");
    *out = osh::mesh_digest(mesh->owner->ctx, mesh->m);
  });
}

int osh_ctx_create(int device, osh_ctx** out) {
  return guarded([&] {
    OSH_CHECK(out != nullptr, "null output pointer");
    OSH_HIP(hipSetDevice(device));
    *out = new osh_ctx();
  });
}

int osh_ctx_destroy(osh_ctx* ctx) {
  return guarded([&] { delete ctx; });
}

int osh_ctx_synchronize(osh_ctx* ctx) {
  return guarded([&] { ctx->ctx.sync(); });
}

int osh_ctx_stream(osh_ctx* ctx, void** out) {
  return guarded([&] {
    OSH_CHECK(ctx && out, "null argument");
    *out = reinterpret_cast<void*>(ctx->ctx.stream);
  });
}

int osh_default_opts(int dim, osh_adapt_opts* out) {
  return guarded([&] {
    OSH_CHECK(dim == 2 || dim == 3, "dimension must be 2 or 3");
    osh::AdaptOptions o;
    osh::default_options(dim, o);
    out->min_length_desired = o.min_length_desired;
    out->max_length_desired = o.max_length_desired;
    out->max_length_allowed = o.max_length_allowed;
    out->min_quality_allowed = o.min_quality_allowed;
    out->min_quality_desired = o.min_quality_desired;
    out->nsliver_layers = o.nsliver_layers;
    out->verbosity = o.verbosity;
    out->should_refine = o.should_refine;
    out->should_coarsen = o.should_coarsen;
    out->should_swap = o.should_swap;
    out->should_coarsen_slivers = o.should_coarsen_slivers;
    out->length_histogram_min = o.length_histogram_min;
    out->length_histogram_max = o.length_histogram_max;
    out->nlength_histogram_bins = o.nlength_histogram_bins;
    out->nquality_histogram_bins = o.nquality_histogram_bins;
    out->should_prevent_coarsen_flip = o.should_prevent_coarsen_flip;
    out->precision = o.precision;
  });
}

static int build(osh_ctx* ctx, int dim, double x, double y, double z, int nx, int ny, int nz,
    bool full, osh_mesh** out) {
  return guarded([&] {
    OSH_CHECK(ctx && out, "null argument");
    OSH_CHECK(dim == 2 || dim == 3, "dimension must be 2 or 3");
    OSH_CHECK(nx > 0 && ny > 0 && (dim == 2 || nz > 0), "cell counts must be positive");
    std::int64_t ncells = std::int64_t(nx) * ny * (dim == 3 ? nz : 1);
    OSH_CHECK(ncells * (dim == 3 ? 6 : 2) < (std::int64_t(1) << 30), "box too large");
    auto m = std::make_unique<osh_mesh>();
    m->owner = ctx;
    osh::build_box(ctx->ctx, m->m, dim, x, y, dim == 3 ? z : 0.0, nx, ny, dim == 3 ? nz : 0,
        full);
    ctx->ctx.sync();
    *out = m.release();
  });
}

int osh_build_box(osh_ctx* ctx, int dim, double x, double y, double z, int nx, int ny, int nz,
    osh_mesh** out) {
  return build(ctx, dim, x, y, z, nx, ny, nz, true, out);
}

int osh_build_box_internal(osh_ctx* ctx, int dim, double x, double y, double z, int nx, int ny,
    int nz, osh_mesh** out) {
  return build(ctx, dim, x, y, z, nx, ny, nz, false, out);
}

int osh_build_from_elems(osh_ctx* ctx, int dim, int nelems, const int* ev2v, int nverts,
    osh_mesh** out) {
  return guarded([&] {
    OSH_CHECK(ctx && out && (ev2v || nelems == 0), "null argument");
    OSH_CHECK(dim == 2 || dim == 3, "dimension must be 2 or 3");
    auto& c = ctx->ctx;
    std::vector<int> h = host_copy(c, ev2v, std::size_t(nelems) * (dim + 1));
    for (int v : h) OSH_CHECK(v >= 0 && v < nverts, "vertex index out of range");
    auto m = std::make_unique<osh_mesh>();
    m->owner = ctx;
    osh::build_from_elems(c, m->m, dim, nelems, osh::to_device(c, h), nverts);
    c.sync();
    *out = m.release();
  });
}

int osh_mesh_destroy(osh_mesh* mesh) {
  return guarded([&] { delete mesh; });
}

int osh_mesh_copy(osh_mesh* mesh, osh_mesh** out) {
  return guarded([&] {
    check_mesh(mesh);
    auto m = std::make_unique<osh_mesh>(*mesh);
    *out = m.release();
  });
}

int osh_mesh_join(osh_ctx* ctx, int nparts, osh_mesh* const* parts, osh_mesh** out) {
  return guarded([&] {
    OSH_CHECK(ctx && parts && out && nparts >= 1, "null argument");
    std::vector<osh::DMesh const*> ps;
    for (int k = 0; k < nparts; ++k) {
      check_mesh(parts[k]);
      OSH_CHECK(parts[k]->owner == ctx, "join: every mesh must belong to the context");
      ps.push_back(&parts[k]->m);
    }
    auto m = std::make_unique<osh_mesh>();
    m->owner = ctx;
    osh::join_meshes(ctx->ctx, ps, m->m);
    *out = m.release();
  });
}

int osh_mesh_split(osh_mesh* joined, int nparts, osh_mesh** out) {
  return guarded([&] {
    check_mesh(joined);
    OSH_CHECK(out != nullptr && nparts >= 1, "null argument");
    std::vector<osh::DMesh> parts;
    osh::split_mesh(joined->owner->ctx, joined->m, nparts, parts);
    std::vector<std::unique_ptr<osh_mesh>> made;
    for (auto& p : parts) {
      made.push_back(std::make_unique<osh_mesh>());
      made.back()->owner = joined->owner;
      made.back()->m = p;
    }
    for (int k = 0; k < nparts; ++k) out[k] = made[std::size_t(k)].release();
  });
}

int osh_mesh_part_counts(osh_mesh* joined, int nparts, int* counts) {
  return guarded([&] {
    check_mesh(joined);
    OSH_CHECK(counts != nullptr && nparts >= 1, "null argument");
    std::vector<osh::LO> first[4], count[4];
    osh::part_ranges(joined->owner->ctx, joined->m, nparts, first, count);
    int dim = joined->m.dim;
    for (int k = 0; k < nparts; ++k)
      for (int d = 0; d <= dim; ++d) counts[k * (dim + 1) + d] = count[d][std::size_t(k)];
  });
}

int osh_mesh_dim(osh_mesh* mesh, int* out) {
  return guarded([&] {
    check_mesh(mesh);
    *out = mesh->m.dim;
  });
}

int osh_mesh_nents(osh_mesh* mesh, int d, int* out) {
  return guarded([&] {
    check_dim(mesh, d);
    *out = mesh->m.nents[d];
  });
}

int osh_mesh_get_verts(osh_mesh* mesh, int d, int* out) {
  return guarded([&] {
    check_dim(mesh, d);
    auto& c = mesh->owner->ctx;
    if (d == 0) {
      download(c, osh::iota(c, mesh->m.nents[0]), out);
      return;
    }
    download(c, osh::verts_of(c, mesh->m, d), out);
  });
}

int osh_mesh_get_down(osh_mesh* mesh, int d, int* out, signed char* codes) {
  return guarded([&] {
    check_dim(mesh, d);
    OSH_CHECK(d >= 1, "vertices have no downward adjacency");
    auto& c = mesh->owner->ctx;
    download(c, mesh->m.down[d], out);
    if (codes) {
      if (d >= 2)
        download(c, mesh->m.down_codes[d], reinterpret_cast<osh::I8*>(codes));
      else
        std::memset(codes, 0, std::size_t(mesh->m.nents[d]) * 2);
    }
  });
}

int osh_mesh_up_size(osh_mesh* mesh, int low, int high, int* out) {
  return guarded([&] {
    check_dim(mesh, high);
    OSH_CHECK(low >= 0 && low < high, "need low < high");
    *out = osh::ask_up(mesh->owner->ctx, mesh->m, low, high).ab2b.size();
  });
}

int osh_mesh_get_up(osh_mesh* mesh, int low, int high, int* a2ab, int* ab2b,
    signed char* codes) {
  return guarded([&] {
    check_dim(mesh, high);
    OSH_CHECK(low >= 0 && low < high, "need low < high");
    auto& c = mesh->owner->ctx;
    auto const& up = osh::ask_up(c, mesh->m, low, high);
    download(c, up.a2ab, a2ab);
    download(c, up.ab2b, ab2b);
    download(c, up.codes, reinterpret_cast<osh::I8*>(codes));
  });
}

int osh_mesh_star_size(osh_mesh* mesh, int d, int* out) {
  return guarded([&] {
    check_dim(mesh, d);
    OSH_CHECK(d <= 1, "stars of vertices and edges only");
    *out = osh::ask_star(mesh->owner->ctx, mesh->m, d).ab2b.size();
  });
}

int osh_mesh_get_star(osh_mesh* mesh, int d, int* a2ab, int* ab2b) {
  return guarded([&] {
    check_dim(mesh, d);
    OSH_CHECK(d <= 1, "stars of vertices and edges only");
    auto& c = mesh->owner->ctx;
    auto const& g = osh::ask_star(c, mesh->m, d);
    download(c, g.a2ab, a2ab);
    download(c, g.ab2b, ab2b);
  });
}

int osh_mesh_dual_size(osh_mesh* mesh, int* out) {
  return guarded([&] {
    check_mesh(mesh);
    auto& c = mesh->owner->ctx;
    *out = osh::ask_dual(c, mesh->m).ab2b.size();
    c.sync();
  });
}

int osh_mesh_get_dual(osh_mesh* mesh, int* a2ab, int* ab2b) {
  return guarded([&] {
    check_mesh(mesh);
    auto& c = mesh->owner->ctx;
    auto const& g = osh::ask_dual(c, mesh->m);
    download(c, g.a2ab, a2ab);
    download(c, g.ab2b, ab2b);
  });
}

int osh_mesh_get_sizes(osh_mesh* mesh, double* out) {
  return guarded([&] {
    check_mesh(mesh);
    auto& c = mesh->owner->ctx;
    download(c, osh::ask_sizes(c, mesh->m), out);
  });
}

int osh_mesh_get_coords(osh_mesh* mesh, double* out) {
  return guarded([&] {
    check_mesh(mesh);
    download(mesh->owner->ctx, mesh->m.coords, out);
  });
}

int osh_mesh_set_metric(osh_mesh* mesh, int ncomps, const double* metric) {
  return guarded([&] {
    check_mesh(mesh);
    int dim = mesh->m.dim;
    OSH_CHECK(ncomps == 1 || ncomps == (dim == 3 ? 6 : 3), "metric must be isotropic or full");
    auto& c = mesh->owner->ctx;
    osh::LO n = mesh->m.nents[0] * ncomps;
    osh::Reals d(c, n);
    osh::upload(c, d, metric, n);
    mesh->m.metric = d;
    mesh->m.metric_ncomps = ncomps;
    // setting the metric invalidates cached lengths and qualities
    // (src/Omega_h_mesh.cpp:167-177)
    mesh->m.length = osh::Reals();
    mesh->m.quality = osh::Reals();
    c.sync();
  });
}

int osh_mesh_metric_ncomps(osh_mesh* mesh, int* out) {
  return guarded([&] {
    check_mesh(mesh);
    *out = mesh->m.metric_ncomps;
  });
}

int osh_mesh_get_metric(osh_mesh* mesh, double* out) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(mesh->m.metric.exists(), "mesh has no metric");
    download(mesh->owner->ctx, mesh->m.metric, out);
  });
}

int osh_mesh_get_class(osh_mesh* mesh, int d, int* ids, signed char* dims) {
  return guarded([&] {
    check_dim(mesh, d);
    OSH_CHECK(mesh->m.class_id[d].exists(), "mesh is not classified");
    auto& c = mesh->owner->ctx;
    download(c, mesh->m.class_id[d], ids);
    download(c, mesh->m.class_dim[d], reinterpret_cast<osh::I8*>(dims));
  });
}

int osh_mesh_set_class(osh_mesh* mesh, int d, const int* ids, const signed char* dims) {
  return guarded([&] {
    check_dim(mesh, d);
    OSH_CHECK(ids && dims, "null argument");
    int dim = mesh->m.dim;
    osh::LO n = mesh->m.nents[d];
    auto& c = mesh->owner->ctx;
    std::vector<signed char> hd = host_copy(c, dims, std::size_t(n));
    for (osh::LO i = 0; i < n; ++i)
      OSH_CHECK(hd[std::size_t(i)] >= d && hd[std::size_t(i)] <= dim,
          "an entity classified on a model entity of lower dimension than itself");
    osh::LOs id(c, n);
    osh::Bytes cd(c, n);
    osh::upload(c, id, ids, n);
    osh::upload(c, cd, reinterpret_cast<const osh::I8*>(hd.data()), n);
    c.sync();
    mesh->m.class_id[d] = id;
    mesh->m.class_dim[d] = cd;
  });
}

int osh_mesh_get_lengths(osh_mesh* mesh, double* out) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(mesh->m.metric.exists(), "mesh has no metric");
    auto& c = mesh->owner->ctx;
    download(c, osh::ask_lengths(c, mesh->m), out);
  });
}

int osh_mesh_get_qualities(osh_mesh* mesh, double* out) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(mesh->m.metric.exists(), "mesh has no metric");
    auto& c = mesh->owner->ctx;
    download(c, osh::ask_qualities(c, mesh->m), out);
  });
}

static int pass(osh_mesh* mesh, const osh_adapt_opts* opts, int* did, int kind) {
  return guarded([&] {
    check_adaptable(mesh);
    OSH_CHECK(opts != nullptr, "null options");
    osh::AdaptOptions o;
    to_options(opts, o);
    auto& c = mesh->owner->ctx;
    bool r;
    if (kind == 0)
      r = osh::refine_by_size(c, mesh->m, o);
    else if (kind == 1)
      r = osh::coarsen_by_size(c, mesh->m, o);
    else if (kind == 2)
      r = osh::swap_edges(c, mesh->m, o);
    else
      r = osh::coarsen_slivers(c, mesh->m, o);
    c.sync();
    if (did) *did = r ? 1 : 0;
  });
}

int osh_refine_by_size(osh_mesh* m, const osh_adapt_opts* o, int* did) { return pass(m, o, did, 0); }
int osh_coarsen_by_size(osh_mesh* m, const osh_adapt_opts* o, int* did) { return pass(m, o, did, 1); }
int osh_swap_edges(osh_mesh* m, const osh_adapt_opts* o, int* did) { return pass(m, o, did, 2); }
int osh_coarsen_slivers(osh_mesh* m, const osh_adapt_opts* o, int* did) { return pass(m, o, did, 3); }

int osh_adapt(osh_mesh* mesh, const osh_adapt_opts* opts, osh_adapt_stats* stats, int* did) {
  return guarded([&] {
    check_adaptable(mesh);
    OSH_CHECK(opts != nullptr, "null options");
    osh::AdaptOptions o;
    to_options(opts, o);
    std::vector<osh::PassRecord> log;
    bool r = osh::adapt(mesh->owner->ctx, mesh->m, o, &log);
    mesh->owner->ctx.sync();
    if (stats) {
      fill_stats(log, stats);
    }
    if (did) *did = r ? 1 : 0;
  });
}

int osh_adapt_batch(int nmeshes, osh_mesh* const* meshes, const osh_adapt_opts* opts,
    osh_adapt_stats* stats, int* did) {
  return guarded([&] {
    OSH_CHECK(nmeshes >= 1 && meshes && opts && did, "null argument");
    osh_ctx* owner = nullptr;
    std::vector<osh::DMesh> ms;
    for (int i = 0; i < nmeshes; ++i) {
      check_adaptable(meshes[i]);
      if (!owner) owner = meshes[i]->owner;
      OSH_CHECK(meshes[i]->owner == owner, "batched adapt: every mesh must share one context");
      for (int j = 0; j < i; ++j)
        OSH_CHECK(meshes[j] != meshes[i], "batched adapt: a mesh listed twice");
      ms.push_back(meshes[i]->m);
    }
    osh::AdaptOptions o;
    to_options(opts, o);
    std::vector<osh::PassRecord> log;
    std::vector<int> r = osh::adapt_batch(owner->ctx, ms, o, &log);
    owner->ctx.sync();
    for (int i = 0; i < nmeshes; ++i) {
      meshes[i]->m = ms[std::size_t(i)];
      did[i] = r[std::size_t(i)];
    }
    if (stats) {
      fill_stats(log, stats);
    }
  });
}

int osh_mesh_screen_estimates(osh_mesh* mesh, double* lengths32, double* qualities32,
    signed char* unsure) {
  return guarded([&] {
    check_metric(mesh);
    OSH_CHECK(lengths32 && qualities32 && unsure, "null argument");
    auto& c = mesh->owner->ctx;
    osh::Reals l, q;
    osh::Bytes u;
    osh::screen_estimates(c, mesh->m, l, q, u);
    download(c, l, lengths32);
    download(c, q, qualities32);
    download(c, u, reinterpret_cast<osh::I8*>(unsure));
  });
}

int osh_adapt_digests(osh_mesh* mesh, const osh_adapt_opts* opts, osh_adapt_stats* stats,
    unsigned long long* digests, int* did) {
  return guarded([&] {
    check_adaptable(mesh);
    OSH_CHECK(opts != nullptr && stats != nullptr && digests != nullptr, "null argument");
    osh::AdaptOptions o;
    to_options(opts, o);
    auto& c = mesh->owner->ctx;
    std::vector<osh::PassRecord> log;
    osh::PassHook hook = [&](osh::DMesh& m) {
      if (log.size() <= OSH_MAX_PASSES) {
        std::uint64_t d[4];
        osh::mesh_digests(c, m, d);
        for (int k = 0; k < 4; ++k) digests[4 * (log.size() - 1) + k] = d[k];
      }
    };
    bool r = osh::adapt(c, mesh->m, o, &log, &hook);
    c.sync();
    fill_stats(log, stats);
    if (did) *did = r ? 1 : 0;
  });
}

int osh_mesh_digests(osh_mesh* mesh, unsigned long long* out) {
  return guarded([&] {
    check_adaptable(mesh);
    OSH_CHECK(out != nullptr, "null output pointer");
    std::uint64_t d[4];
    osh::mesh_digests(mesh->owner->ctx, mesh->m, d);
    for (int k = 0; k < 4; ++k) out[k] = d[k];
  });
}

int osh_measure_edges_metric(osh_ctx* ctx, int dim, int metric_ncomps, int n, const int* edges,
    const int* ev2v, const double* coords, const double* metric, double* out) {
  return guarded([&] {
    OSH_CHECK(ctx && ev2v && coords && metric && out, "null argument");
    OSH_CHECK(dim == 2 || dim == 3, "dimension must be 2 or 3");
    OSH_CHECK(metric_ncomps == 1 || metric_ncomps == (dim == 3 ? 6 : 3), "bad metric size");
    OSH_CHECK(n >= 0, "negative count");
    osh::launch_edge_lengths(ctx->ctx, dim, metric_ncomps == 1 ? 1 : dim, n, edges, ev2v,
        coords, metric, out);
  });
}

int osh_measure_qualities(osh_ctx* ctx, int dim, int metric_ncomps, int n, const int* elems,
    const int* ev2v, const double* coords, const double* metric, double* out) {
  return guarded([&] {
    OSH_CHECK(ctx && ev2v && coords && metric && out, "null argument");
    OSH_CHECK(dim == 2 || dim == 3, "dimension must be 2 or 3");
    OSH_CHECK(metric_ncomps == 1 || metric_ncomps == (dim == 3 ? 6 : 3), "bad metric size");
    OSH_CHECK(n >= 0, "negative count");
    osh::launch_elem_qualities(ctx->ctx, dim, metric_ncomps == 1 ? 1 : dim, n, elems, ev2v,
        coords, metric, out);
  });
}

int osh_offset_scans(osh_ctx* ctx, int njobs, const int* const* in, const int* n,
    int* const* out) {
  return guarded([&] {
    OSH_CHECK(ctx && in && n && out, "null argument");
    OSH_CHECK(njobs >= 1 && njobs <= osh::kMaxScanJobs, "1 to 12 scans per call");
    osh::ScanJob jobs[osh::kMaxScanJobs];
    for (int j = 0; j < njobs; ++j) {
      OSH_CHECK(n[j] >= 0, "negative count");
      OSH_CHECK(out[j] && (in[j] || n[j] == 0), "null array");
      jobs[j] = {in[j], osh::kScanLO, n[j], out[j], nullptr};
    }
    osh::offset_scans(ctx->ctx, jobs, njobs);
  });
}

int osh_ctx_set_tuning(osh_ctx* ctx, const char* name, int value) {
  return guarded([&] {
    OSH_CHECK(ctx && name, "null argument");
    OSH_CHECK(value >= 1 && value <= 64 && (value & (value - 1)) == 0,
        "a lane count is a power of two from 1 to 64");
    *tuning_field(ctx->ctx.tuning, name) = value;
  });
}

int osh_ctx_get_tuning(osh_ctx* ctx, const char* name, int* out) {
  return guarded([&] {
    OSH_CHECK(ctx && name && out, "null argument");
    *out = *tuning_field(ctx->ctx.tuning, name);
  });
}

int osh_device_alloc(osh_ctx* ctx, long long bytes, void** out) {
  return guarded([&] {
    OSH_CHECK(ctx && out, "null argument");
    OSH_CHECK(bytes >= 0, "negative size");
    *out = bytes ? ctx->ctx.pool->get(std::size_t(bytes)) : nullptr;
  });
}

int osh_device_free(osh_ctx* ctx, void* p, long long bytes) {
  return guarded([&] {
    OSH_CHECK(ctx, "null argument");
    if (p) ctx->ctx.pool->put(p, std::size_t(bytes));
  });
}

int osh_copy(osh_ctx* ctx, void* dst, const void* src, long long bytes) {
  return guarded([&] {
    OSH_CHECK(ctx && ((dst && src) || bytes == 0), "null argument");
    OSH_CHECK(bytes >= 0, "negative size");
    if (bytes)
      OSH_HIP(hipMemcpyAsync(dst, src, std::size_t(bytes), hipMemcpyDefault, ctx->ctx.stream));
    ctx->ctx.sync();
  });
}

int osh_device_fill(osh_ctx* ctx, void* dst, long long n, int item_bytes, const void* value) {
  return guarded([&] {
    OSH_CHECK(ctx && value && (dst || n == 0), "null argument");
    OSH_CHECK(n >= 0, "negative count");
    OSH_CHECK(item_bytes == 1 || item_bytes == 4 || item_bytes == 8, "items of 1, 4 or 8 bytes");
    if (n == 0) return;
    if (item_bytes == 1)
      fill_device<unsigned char>(ctx->ctx, dst, n, value);
    else if (item_bytes == 4)
      fill_device<unsigned>(ctx->ctx, dst, n, value);
    else
      fill_device<unsigned long long>(ctx->ctx, dst, n, value);
    ctx->ctx.sync();
  });
}

int osh_pointer_is_device(const void* p, int* out) {
  return guarded([&] {
    OSH_CHECK(out, "null argument");
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      *out = 0;
      return;
    }
    *out = a.type == hipMemoryTypeDevice ? 1 : 0;
  });
}

int osh_mesh_min_quality(osh_mesh* mesh, double* out) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(out, "null argument");
    OSH_CHECK(mesh->m.metric.exists(), "mesh has no metric");
    auto& c = mesh->owner->ctx;
    *out = osh::min_real(c, osh::ask_qualities(c, mesh->m));
  });
}

int osh_mesh_max_length(osh_mesh* mesh, double* out) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(out, "null argument");
    OSH_CHECK(mesh->m.metric.exists(), "mesh has no metric");
    auto& c = mesh->owner->ctx;
    *out = osh::max_real(c, osh::ask_lengths(c, mesh->m));
  });
}

int osh_sort_by_keys(osh_ctx* ctx, int n, int width, const long long* keys, int* perm) {
  return guarded([&] {
    OSH_CHECK(ctx && perm && (keys || n == 0), "null argument");
    OSH_CHECK(n >= 0, "negative count");
    OSH_CHECK(width >= 1 && width <= 3, "tuples of 1, 2 or 3 keys");
    OSH_CHECK(std::int64_t(n) * width <= 0x7fffffff, "too many keys");
    auto& c = ctx->ctx;
    osh::DArray<osh::I64> k(c, n * width);
    osh::upload(c, k, reinterpret_cast<osh::I64 const*>(keys), n * width);
    download(c, osh::sort_by_keys(c, k, width), perm);
  });
}

int osh_ctx_set_profile(osh_ctx* ctx, int flags) {
  return guarded([&] {
    OSH_CHECK(ctx, "null argument");
    OSH_CHECK(flags >= 0 && flags <= (osh::kProfileTree | osh::kProfileMarkers | osh::kProfileSync),
        "profile flags are 1 (timer tree), 2 (roctx ranges) and 4 (wait at range ends)");
    OSH_CHECK(ctx->ctx.profile.current == osh::Profile::kNone,
        "profile flags change only while no range is open");
    ctx->ctx.profile.flags = flags;
  });
}

int osh_ctx_profile_clear(osh_ctx* ctx) {
  return guarded([&] {
    OSH_CHECK(ctx, "null argument");
    OSH_CHECK(ctx->ctx.profile.current == osh::Profile::kNone, "a range is open");
    ctx->ctx.profile.clear();
  });
}

int osh_ctx_profile_nframes(osh_ctx* ctx, int* n) {
  return guarded([&] {
    OSH_CHECK(ctx && n, "null argument");
    *n = int(ctx->ctx.profile.frames.size());
  });
}

int osh_ctx_profile_frame(osh_ctx* ctx, int i, int* parent, double* seconds, long long* calls,
    char* name, int name_cap) {
  return guarded([&] {
    OSH_CHECK(ctx && parent && seconds && calls && name && name_cap > 0, "null argument");
    auto const& fr = ctx->ctx.profile.frames;
    OSH_CHECK(i >= 0 && std::size_t(i) < fr.size(), "frame index out of range");
    auto const& f = fr[std::size_t(i)];
    *parent = (f.parent == osh::Profile::kNone) ? -1 : int(f.parent);
    *seconds = f.seconds;
    *calls = f.calls;
    std::snprintf(name, std::size_t(name_cap), "%s", f.name.c_str());
  });
}

int osh_ctx_profile_report(osh_ctx* ctx, double chop, int percent, char* buf, long long cap,
    long long* len) {
  return guarded([&] {
    OSH_CHECK(ctx && len, "null argument");
    OSH_CHECK(cap >= 0 && (buf || cap == 0), "null buffer");
    auto const& p = ctx->ctx.profile;
    std::string r = p.report(chop, percent != 0, p.root_seconds());
    *len = static_cast<long long>(r.size());
    if (cap > 0) std::snprintf(buf, std::size_t(cap), "%s", r.c_str());
  });
}

int osh_offset_scans_noalloc(osh_ctx* ctx, int njobs, const int* const* in, const int* n,
    int* const* out) {
  return guarded([&] {
    OSH_CHECK(ctx && in && n && out, "null argument");
    OSH_CHECK(njobs >= 1 && njobs <= osh::kMaxScanJobs, "1 to 12 scans per call");
    osh::ScanJob jobs[osh::kMaxScanJobs];
    for (int j = 0; j < njobs; ++j) {
      OSH_CHECK(n[j] >= 0, "negative count");
      OSH_CHECK(out[j] && (in[j] || n[j] == 0), "null array");
      jobs[j] = {in[j], osh::kScanLO, n[j], out[j], nullptr};
    }
    osh::offset_scans_noalloc(ctx->ctx, jobs, njobs);
  });
}

int osh_limit_metric_gradation(osh_mesh* mesh, double max_rate, double tol, int* nsteps) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(tol >= 0.0, "tolerance must not be negative");
    auto& c = mesh->owner->ctx;
    int steps = osh::limit_metric_gradation(c, mesh->m, max_rate, tol);
    c.sync();
    if (nsteps) *nsteps = steps;
  });
}

int osh_mesh_add_vertex_tag(osh_mesh* mesh, const char* name, int ncomps, int kind,
    const double* data) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(name && data, "null argument");
    OSH_CHECK(kind == 0 || kind == 1, "kind must be 0 (linear) or 1 (metric)");
    OSH_CHECK(ncomps >= 1 && ncomps <= 64, "1 to 64 components");
    if (kind == 1)
      OSH_CHECK(ncomps == 1 || ncomps == (mesh->m.dim == 3 ? 6 : 3), "metric tag size");
    auto& c = mesh->owner->ctx;
    osh::LO n = mesh->m.nents[0] * ncomps;
    osh::Reals d(c, n);
    osh::upload(c, d, data, n);
    c.sync();
    for (auto& t : mesh->m.vtags)
      if (t.name == name) {
        t = {name, ncomps, kind, d};
        return;
      }
    mesh->m.vtags.push_back({name, ncomps, kind, d});
  });
}

int osh_mesh_add_inherit_tag(osh_mesh* mesh, const char* name, int ncomps,
    const double* const* per_dim) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(name && per_dim, "null argument");
    OSH_CHECK(ncomps >= 1 && ncomps <= 64, "1 to 64 components");
    auto& c = mesh->owner->ctx;
    osh::DMesh::InheritTag t;
    t.name = name;
    t.ncomps = ncomps;
    for (int d = 0; d <= mesh->m.dim; ++d) {
      OSH_CHECK(per_dim[d] != nullptr, "null array for a dimension");
      osh::LO n = mesh->m.nents[d] * ncomps;
      t.data[d] = osh::Reals(c, n);
      osh::upload(c, t.data[d], per_dim[d], n);
    }
    c.sync();
    for (auto& x : mesh->m.itags)
      if (x.name == name) {
        x = t;
        return;
      }
    mesh->m.itags.push_back(t);
  });
}

int osh_mesh_inherit_tag_ncomps(osh_mesh* mesh, const char* name, int* out) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(name && out, "null argument");
    *out = -1;
    for (auto const& t : mesh->m.itags)
      if (t.name == name) *out = t.ncomps;
  });
}

int osh_mesh_get_inherit_tag(osh_mesh* mesh, const char* name, int ent_dim, double* out) {
  return guarded([&] {
    check_dim(mesh, ent_dim);
    OSH_CHECK(name && out, "null argument");
    for (auto const& t : mesh->m.itags)
      if (t.name == name) {
        download(mesh->owner->ctx, t.data[ent_dim], out);
        return;
      }
    throw osh::Error(std::string("no inheritance tag ") + name);
  });
}

int osh_mesh_vertex_tag_ncomps(osh_mesh* mesh, const char* name, int* out) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(name && out, "null argument");
    *out = -1;
    for (auto const& t : mesh->m.vtags)
      if (t.name == name) *out = t.ncomps;
  });
}

int osh_mesh_get_vertex_tag(osh_mesh* mesh, const char* name, double* out) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(name && out, "null argument");
    for (auto const& t : mesh->m.vtags)
      if (t.name == name) {
        download(mesh->owner->ctx, t.data, out);
        return;
      }
    throw osh::Error(std::string("no vertex tag ") + name);
  });
}

int osh_mesh_remove_vertex_tag(osh_mesh* mesh, const char* name) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(name != nullptr, "null argument");
    auto& v = mesh->m.vtags;
    for (auto it = v.begin(); it != v.end(); ++it)
      if (it->name == name) {
        v.erase(it);
        return;
      }
  });
}

int osh_mesh_add_elem_tag(osh_mesh* mesh, const char* name, int ncomps, const double* data) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(name && data, "null argument");
    OSH_CHECK(ncomps >= 1 && ncomps <= 64, "1 to 64 components");
    auto& c = mesh->owner->ctx;
    osh::LO n = mesh->m.nents[mesh->m.dim] * ncomps;
    osh::Reals d(c, n);
    osh::upload(c, d, data, n);
    c.sync();
    for (auto& t : mesh->m.etags)
      if (t.name == name) {  // new values; the transfer set for the tag stays
        t = t.with_data(d);
        t.ncomps = ncomps;
        return;
      }
    mesh->m.etags.push_back({name, ncomps, d});
  });
}

int osh_mesh_set_elem_transfer(osh_mesh* mesh, const char* name, int kind, int diffuse_type,
    double tol, double floor) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(name, "null argument");
    OSH_CHECK(kind >= 0 && kind <= 2, "kind must be 0 (pointwise), 1 (conserve) or 2 (density)");
    osh::Diffusion d = diffusion_of(diffuse_type, tol, floor);
    for (auto& t : mesh->m.etags)
      if (t.name == name) {
        t.kind = kind;
        t.diffuse = d;
        return;
      }
    throw osh::Error(std::string("no element tag ") + name);
  });
}

int osh_mesh_set_momentum_velocity(osh_mesh* mesh, const char* name, const char* density,
    int diffuse_type, double tol, double floor) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(name && density, "null argument");
    osh::Diffusion d = diffusion_of(diffuse_type, tol, floor);
    bool found = false;
    for (auto const& e : mesh->m.etags)
      if (e.name == density) {
        OSH_CHECK(e.ncomps == 1 && e.kind == 1,
            "a momentum velocity needs a conserved density of one component");
        found = true;
      }
    OSH_CHECK(found, std::string("no element tag ") + density);
    for (auto& t : mesh->m.vtags)
      if (t.name == name) {
        OSH_CHECK(t.ncomps == mesh->m.dim, "a momentum velocity has one component per dimension");
        t.kind = 2;
        t.density = density;
        t.diffuse = d;
        return;
      }
    throw osh::Error(std::string("no vertex tag ") + name);
  });
}

int osh_mesh_elem_tag_ncomps(osh_mesh* mesh, const char* name, int* out) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(name && out, "null argument");
    *out = -1;
    for (auto const& t : mesh->m.etags)
      if (t.name == name) *out = t.ncomps;
  });
}

int osh_mesh_get_elem_tag(osh_mesh* mesh, const char* name, double* out) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(name && out, "null argument");
    for (auto const& t : mesh->m.etags)
      if (t.name == name) {
        download(mesh->owner->ctx, t.data, out);
        return;
      }
    throw osh::Error(std::string("no element tag ") + name);
  });
}

int osh_mesh_remove_elem_tag(osh_mesh* mesh, const char* name) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(name != nullptr, "null argument");
    auto& v = mesh->m.etags;
    for (auto it = v.begin(); it != v.end(); ++it)
      if (it->name == name) {
        v.erase(it);
        return;
      }
  });
}

int osh_approach_metric(osh_mesh* mesh, const osh_adapt_opts* opts, double min_step, int* did) {
  return guarded([&] {
    check_metric(mesh);
    OSH_CHECK(opts != nullptr, "null options");
    osh::AdaptOptions o;
    to_options(opts, o);
    bool r = osh::approach_metric(mesh->owner->ctx, mesh->m, o, min_step);
    mesh->owner->ctx.sync();
    if (did) *did = r ? 1 : 0;
  });
}

int osh_warp_to_limit(osh_mesh* mesh, const osh_adapt_opts* opts, int exit_on_stall,
    int max_niters, int* did) {
  return guarded([&] {
    check_metric(mesh);
    OSH_CHECK(opts != nullptr, "null options");
    osh::AdaptOptions o;
    to_options(opts, o);
    bool r = osh::warp_to_limit(mesh->owner->ctx, mesh->m, o, exit_on_stall != 0, max_niters);
    mesh->owner->ctx.sync();
    if (did) *did = r ? 1 : 0;
  });
}

int osh_mesh_set_coords(osh_mesh* mesh, const double* data) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(data != nullptr, "null argument");
    auto& c = mesh->owner->ctx;
    osh::LO n = mesh->m.nents[0] * mesh->m.dim;
    osh::Reals d(c, n);
    osh::upload(c, d, data, n);
    c.sync();
    osh::set_coords(mesh->m, d);
  });
}

int osh_get_implied_isos(osh_mesh* mesh, int set_as_metric, double* out) {
  return guarded([&] {
    check_mesh(mesh);
    auto& c = mesh->owner->ctx;
    osh::Reals r = osh::implied_isos(c, mesh->m);
    if (out) download(c, r, out);
    if (set_as_metric) {
      mesh->m.metric = r;
      mesh->m.metric_ncomps = 1;
      mesh->m.length = osh::Reals();
      mesh->m.quality = osh::Reals();
    }
    c.sync();
  });
}

int osh_ill_metric(osh_mesh* mesh, const osh_adapt_opts* opts, int isotropy) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(opts != nullptr, "null options");
    osh::AdaptOptions o;
    to_options(opts, o);
    osh::ill_metric(mesh->owner->ctx, mesh->m, o, isotropy);
    mesh->owner->ctx.sync();
  });
}

int osh_fix(osh_mesh* mesh, const osh_adapt_opts* opts, int isotropy, int* ncalls) {
  return guarded([&] {
    check_classified(mesh);
    OSH_CHECK(opts != nullptr, "null options");
    osh::AdaptOptions o;
    to_options(opts, o);
    int r = osh::fix(mesh->owner->ctx, mesh->m, o, isotropy);
    mesh->owner->ctx.sync();
    if (ncalls) *ncalls = r;
  });
}

int osh_fix_for_given_metric(osh_mesh* mesh, const osh_adapt_opts* opts, int* ncalls) {
  return guarded([&] {
    check_adaptable(mesh);
    OSH_CHECK(opts != nullptr, "null options");
    osh::AdaptOptions o;
    to_options(opts, o);
    int r = osh::fix_for_given_metric(mesh->owner->ctx, mesh->m, o);
    mesh->owner->ctx.sync();
    if (ncalls) *ncalls = r;
  });
}

int osh_grade_fix_adapt(osh_mesh* mesh, const osh_adapt_opts* opts, int ncomps,
    const double* target, int* nsteps) {
  return guarded([&] {
    check_classified(mesh);
    OSH_CHECK(opts != nullptr && target != nullptr, "null argument");
    int dim = mesh->m.dim;
    OSH_CHECK(ncomps == 1 || ncomps == (dim * (dim + 1)) / 2, "target metric size");
    osh::AdaptOptions o;
    to_options(opts, o);
    auto& c = mesh->owner->ctx;
    osh::LO n = mesh->m.nents[0] * ncomps;
    osh::Reals t(c, n);
    osh::upload(c, t, target, n);
    int r = osh::grade_fix_adapt(c, mesh->m, o, ncomps, t);
    c.sync();
    if (nsteps) *nsteps = r;
  });
}

int osh_get_implied_metrics(osh_mesh* mesh, int set_as_metric, double* out) {
  return guarded([&] {
    check_mesh(mesh);
    auto& c = mesh->owner->ctx;
    osh::Reals r = osh::implied_metrics(c, mesh->m);
    if (out) download(c, r, out);
    if (set_as_metric) {
      mesh->m.metric = r;
      mesh->m.metric_ncomps = (mesh->m.dim * (mesh->m.dim + 1)) / 2;
      mesh->m.length = osh::Reals();
      mesh->m.quality = osh::Reals();
    }
    c.sync();
  });
}

int osh_interpolate_metrics(osh_ctx* ctx, int dim, int n, const double* a, const double* b,
    double t, double* out) {
  return guarded([&] {
    OSH_CHECK(ctx && (n == 0 || (a && b && out)), "null argument");
    OSH_CHECK(dim >= 1 && dim <= 3, "dimension must be 1, 2 or 3");
    OSH_CHECK(n >= 0, "negative count");
    auto& c = ctx->ctx;
    osh::LO nc = (dim * (dim + 1)) / 2;
    std::vector<double> ha = host_copy(c, a, std::size_t(n) * nc);
    std::vector<double> hb = host_copy(c, b, std::size_t(n) * nc);
    osh::Reals r;
    osh::interpolate_metric_arrays(c, dim, n, osh::to_device(c, ha), osh::to_device(c, hb), t, r);
    download(c, r, out);
  });
}

int osh_intersect_metrics(osh_ctx* ctx, int dim, int n, const double* m1, const double* m2,
    double* out) {
  return guarded([&] {
    OSH_CHECK(ctx && (n == 0 || (m1 && m2 && out)), "null argument");
    OSH_CHECK(dim >= 1 && dim <= 3, "dimension must be 1, 2 or 3");
    OSH_CHECK(n >= 0, "negative count");
    auto& c = ctx->ctx;
    osh::LO nc = (dim * (dim + 1)) / 2;
    std::vector<double> ha = host_copy(c, m1, std::size_t(n) * nc);
    std::vector<double> hb = host_copy(c, m2, std::size_t(n) * nc);
    osh::Reals r;
    osh::intersect_metric_arrays(c, dim, n, osh::to_device(c, ha), osh::to_device(c, hb), r);
    download(c, r, out);
  });
}

int osh_mident_metrics(osh_mesh* mesh, int n, const int* edges, double* out) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(mesh->m.metric.exists(), "mesh has no metric");
    OSH_CHECK(n >= 0 && (n == 0 || (edges && out)), "null argument");
    auto& c = mesh->owner->ctx;
    std::vector<int> h = host_copy(c, edges, std::size_t(n));
    for (int e : h) OSH_CHECK(e >= 0 && e < mesh->m.nents[1], "edge out of range");
    osh::LOs d = osh::to_device(c, h);
    osh::Reals r(c, n * mesh->m.metric_ncomps);
    if (n) osh::mident_metrics(c, mesh->m, n, d.data(), r.data());
    download(c, r, out);
  });
}

int osh_swap3d_qualities(osh_mesh* mesh, const osh_adapt_opts* opts, int n, const int* edges,
    double* quals, signed char* configs) {
  return guarded([&] {
    check_adaptable(mesh);
    OSH_CHECK(opts && n >= 0 && (n == 0 || (edges && quals && configs)), "null argument");
    OSH_CHECK(mesh->m.dim == 3, "swap3d qualities need a 3D mesh");
    osh::AdaptOptions o;
    to_options(opts, o);
    auto& c = mesh->owner->ctx;
    std::vector<int> h = host_copy(c, edges, std::size_t(n));
    std::vector<signed char> edim(std::size_t(mesh->m.nents[1]));
    download(c, mesh->m.class_dim[1], reinterpret_cast<osh::I8*>(edim.data()));
    std::vector<char> seen(edim.size(), 0);
    for (int e : h) {
      OSH_CHECK(e >= 0 && e < mesh->m.nents[1], "edge out of range");
      OSH_CHECK(edim[std::size_t(e)] == 3, "swap candidates are interior edges");
      OSH_CHECK(!seen[std::size_t(e)], "an edge listed twice");
      seen[std::size_t(e)] = 1;
    }
    osh::LOs cfg;
    osh::Reals q;
    osh::swap3d_choices(c, mesh->m, o, osh::to_device(c, h), cfg, q);
    std::vector<int> hc = osh::to_host(c, cfg);
    for (int i = 0; i < n; ++i) configs[i] = static_cast<signed char>(hc[std::size_t(i)]);
    download(c, q, quals);
  });
}

int osh_invert_adj(osh_ctx* ctx, int nhigh, int deg, int nlow, const int* hl2l,
    const signed char* down_codes, int* a2ab, int* ab2b, signed char* codes) {
  return guarded([&] {
    OSH_CHECK(ctx && a2ab && (nhigh == 0 || (hl2l && ab2b && codes)), "null argument");
    OSH_CHECK(nhigh >= 0 && nlow >= 0 && deg >= 1 && deg <= 6, "bad sizes");
    OSH_CHECK(std::int64_t(nhigh) * deg <= 0x7fffffff, "too many uses");
    auto& c = ctx->ctx;
    std::vector<int> h = host_copy(c, hl2l, std::size_t(nhigh) * std::size_t(deg));
    for (int l : h) OSH_CHECK(l >= 0 && l < nlow, "low entity out of range");
    osh::Bytes dc;
    if (down_codes)
      dc = osh::to_device(c, host_copy(c, reinterpret_cast<osh::I8 const*>(down_codes),
                                       std::size_t(nhigh) * std::size_t(deg)));
    osh::Adj a = osh::invert_adj_rows(c, nhigh, deg, nlow, osh::to_device(c, h),
        down_codes ? &dc : nullptr);
    download(c, a.a2ab, a2ab);
    osh::LO nuses = nhigh * deg;
    if (nuses) {
      OSH_HIP(hipMemcpyAsync(ab2b, a.ab2b.data(), sizeof(int) * std::size_t(nuses),
          hipMemcpyDefault, c.stream));
      OSH_HIP(hipMemcpyAsync(codes, a.codes.data(), std::size_t(nuses), hipMemcpyDefault,
          c.stream));
    }
    c.sync();
  });
}

int osh_reflect_down(osh_mesh* mesh, int high_dim, int nprods, const int* prod_verts,
    int* down, signed char* codes) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(high_dim >= 2 && high_dim <= mesh->m.dim, "high_dim between 2 and the mesh's");
    OSH_CHECK(nprods >= 0 && (nprods == 0 || (prod_verts && down && codes)), "null argument");
    auto& c = mesh->owner->ctx;
    std::vector<int> h = host_copy(c, prod_verts, std::size_t(nprods) * (high_dim + 1));
    for (int v : h) OSH_CHECK(v >= 0 && v < mesh->m.nents[0], "vertex out of range");
    osh::LOs d, found;
    osh::Bytes cd;
    osh::reflect_down(c, mesh->m, high_dim, nprods, osh::to_device(c, h), d, cd, &found);
    std::vector<int> hf = osh::to_host(c, found);
    for (int f : hf) OSH_CHECK(f == 1, "a side of a given simplex is not in the mesh");
    download(c, d, down);
    download(c, cd, reinterpret_cast<osh::I8*>(codes));
  });
}

int osh_collect_marked(osh_ctx* ctx, int n, const signed char* marks, int* out, int* count) {
  return guarded([&] {
    OSH_CHECK(ctx && count && n >= 0 && (n == 0 || (marks && out)), "null argument");
    auto& c = ctx->ctx;
    osh::Bytes m = osh::to_device(c, host_copy(c, reinterpret_cast<osh::I8 const*>(marks),
                                               std::size_t(n)));
    osh::LOs r = osh::collect_marked(c, m);
    *count = r.size();
    download(c, r, out);
  });
}

int osh_refine_qualities(osh_mesh* mesh, int n, const int* cands, double* out) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(mesh->m.metric.exists(), "mesh has no metric");
    auto& c = mesh->owner->ctx;
    std::vector<int> h = host_copy(c, cands, std::size_t(n));
    for (int e : h) OSH_CHECK(e >= 0 && e < mesh->m.nents[1], "edge out of range");
    auto q = osh::refine_qualities(c, mesh->m, osh::to_device(c, h));
    download(c, q, out);
  });
}

int osh_coarsen_qualities(osh_mesh* mesh, int ncands, const int* cands,
    const signed char* cand_codes, double* out, int precision, double min_quality) {
  return guarded([&] {
    check_metric(mesh);
    OSH_CHECK(ncands >= 0 && (ncands == 0 || (cands && cand_codes && out)), "null argument");
    OSH_CHECK(precision == 0 || precision == 1, "precision must be OSH_FP64 or OSH_MIXED");
    osh::coarsen_qualities(mesh->owner->ctx, mesh->m, ncands, cands,
        reinterpret_cast<osh::I8 const*>(cand_codes), out, precision, min_quality);
  });
}

int osh_find_loop(osh_mesh* mesh, int edge, int* eev2v, int* loop_verts, int* size) {
  return guarded([&] {
    check_mesh(mesh);
    OSH_CHECK(mesh->m.dim == 3, "edge loops exist in 3D");
    OSH_CHECK(edge >= 0 && edge < mesh->m.nents[1], "edge out of range");
    auto& c = mesh->owner->ctx;
    osh::LOs e2, lv, sz;
    osh::find_loops(c, mesh->m, osh::to_device(c, std::vector<int>{edge}), e2, lv, sz);
    download(c, e2, eev2v);
    download(c, lv, loop_verts);
    download(c, sz, size);
  });
}

int osh_find_indset(osh_mesh* mesh, int d, const double* quality, const signed char* cands,
    signed char* out) {
  return guarded([&] {
    check_dim(mesh, d);
    auto& c = mesh->owner->ctx;
    osh::LO n = mesh->m.nents[d];
    osh::Reals q(c, n);
    osh::Bytes cb(c, n);
    osh::upload(c, q, quality, n);
    osh::upload(c, cb, reinterpret_cast<osh::I8 const*>(cands), n);
    auto r = osh::find_indset(c, mesh->m, d, q, cb);
    download(c, r, reinterpret_cast<osh::I8*>(out));
  });
}

int osh_swap3d_tables(int loop_size, int* counts, int* tris, int* meshes, int* uedges,
    int* medges) {
  return guarded([&] {
    OSH_CHECK(loop_size >= 3 && loop_size <= osh::swap::kMaxLoop, "loop size must be 3 to 7");
    static osh::swap::Tables const t = osh::swap::make_tables();
    int n = loop_size;
    int ntris = t.ntris[n], nmeshes = t.mesh_counts[n], nue = t.nuedges[n];
    int per_mesh = t.mesh_sizes[n], edges_per_mesh = t.nedges[n];
    if (counts) {
      counts[0] = ntris;
      counts[1] = nmeshes;
      counts[2] = nue;
      counts[3] = per_mesh;
    }
    if (tris)
      for (int i = 0; i < ntris; ++i)
        for (int k = 0; k < 3; ++k) tris[i * 3 + k] = t.tris[n][i][k];
    if (meshes)
      for (int i = 0; i < nmeshes * per_mesh; ++i) meshes[i] = t.meshes[n][i];
    if (uedges)
      for (int i = 0; i < nue; ++i)
        for (int k = 0; k < 2; ++k) uedges[i * 2 + k] = t.uedges[n][i][k];
    if (medges)
      for (int i = 0; i < nmeshes * edges_per_mesh; ++i) medges[i] = t.medges[n][i];
  });
}

int osh_eval_math(osh_ctx* ctx, int which, int n, const double* in, double* out) {
  return guarded([&] {
    OSH_CHECK(ctx && in && out, "null argument");
    OSH_CHECK(which >= 0 && which <= 4, "unknown function");
    auto& c = ctx->ctx;
    osh::Reals din(c, n), dout(c, n);
    osh::upload(c, din, in, n);
    OSH_LAUNCH(eval_math_kernel, n, c, which, din.data(), dout.data());
    download(c, dout, out);
  });
}

}  // extern "C"

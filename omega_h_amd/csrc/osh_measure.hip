// Metric-space edge lengths and mean-ratio element qualities, and the
// log-Euclidean midpoint metrics of edges.
//
//   measure_edges_metric  src/Omega_h_shape.cpp:7-48
//   measure_qualities     src/Omega_h_quality.cpp:7-41
//   get_mident_metrics    src/Omega_h_metric.cpp:56-104
#include "osh_math.hpp"
#include "osh_mesh.hpp"

namespace osh {

namespace {

// one vertex's coordinates and metric; a 6-component metric row is 48 bytes
// at a 16-byte aligned offset and is read as three 16-byte loads
template <int S, int NC>
__device__ __forceinline__ void load_vertex(Real const* coords, Real const* metric, LO v,
    double* p, double* m) {
#pragma unroll
  for (int j = 0; j < S; ++j) p[j] = coords[v * S + j];
  if constexpr (NC == 6) {
    double2 const* q = reinterpret_cast<double2 const*>(metric + std::size_t(v) * 6);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double2 t = q[j];
      m[2 * j] = t.x;
      m[2 * j + 1] = t.y;
    }
  } else {
#pragma unroll
    for (int j = 0; j < NC; ++j) m[j] = metric[v * NC + j];
  }
}

template <int S, int M>
__device__ __forceinline__ Real length_of(LO e, LO const* ev2v, Real const* coords,
    Real const* metric) {
  constexpr int nc = (M == 1) ? 1 : (M == 2 ? 3 : 6);
  int2 ab = reinterpret_cast<int2 const*>(ev2v)[e];
  double p0[S], p1[S], m0[nc], m1[nc];
  load_vertex<S, nc>(coords, metric, ab.x, p0, m0);
  load_vertex<S, nc>(coords, metric, ab.y, p1, m1);
  return edge_length<S, M>(p0, p1, m0, m1);
}

template <int D, int M>
__device__ __forceinline__ Real quality_of(LO e, LO const* ev2v, Real const* coords,
    Real const* metric) {
  constexpr int nc = (M == 1) ? 1 : (M == 2 ? 3 : 6);
  LO vs[D + 1];
  if constexpr (D == 3) {
    int4 t = reinterpret_cast<int4 const*>(ev2v)[e];
    vs[0] = t.x;
    vs[1] = t.y;
    vs[2] = t.z;
    vs[3] = t.w;
  } else {
#pragma unroll
    for (int k = 0; k < D + 1; ++k) vs[k] = ev2v[e * (D + 1) + k];
  }
  double p[D + 1][D];
  double m[D + 1][nc];
  double const* pp[D + 1];
  double const* mm[D + 1];
#pragma unroll
  for (int k = 0; k < D + 1; ++k) {
    load_vertex<D, nc>(coords, metric, vs[k], p[k], m[k]);
    pp[k] = p[k];
    mm[k] = m[k];
  }
  return element_quality<D, M>(pp, mm);
}

template <int S, int M>
__global__ void __launch_bounds__(kBlock) edge_lengths_kernel(LO n, LO const* edges,
    LO const* ev2v, Real const* coords, Real const* metric, Real* out) {
  LO i = xcd_thread_index();
  if (i >= n) return;
  out[i] = length_of<S, M>(edges ? edges[i] : i, ev2v, coords, metric);
}

template <int D, int M>
__global__ void __launch_bounds__(kBlock) elem_qualities_kernel(LO n, LO const* elems,
    LO const* ev2v, Real const* coords, Real const* metric, Real* out) {
  LO i = xcd_thread_index();
  if (i >= n) return;
  out[i] = quality_of<D, M>(elems ? elems[i] : i, ev2v, coords, metric);
}

// the FP32 screening estimates of every edge length and element quality
// (the mixed-precision form's screens; reported against the FP64 values)
template <int S, int M>
__global__ void edge_lengths32_kernel(LO n, LO const* ev2v, Real const* coords,
    Real const* metric, Real* out) {
  LO e = thread_index();
  if (e >= n) return;
  constexpr int nc = (M == 1) ? 1 : (M == 2 ? 3 : 6);
  double p0[S], p1[S], m0[nc], m1[nc];
  load_vertex<S, nc>(coords, metric, ev2v[2 * e], p0, m0);
  load_vertex<S, nc>(coords, metric, ev2v[2 * e + 1], p1, m1);
  out[e] = double(f32::edge_length<S, M>(p0, p1, m0, m1));
}

template <int D, int M>
__global__ void elem_qualities32_kernel(LO n, LO const* ev2v, Real const* coords,
    Real const* metric, Real* out, I8* unsure) {
  LO e = thread_index();
  if (e >= n) return;
  constexpr int nc = (M == 1) ? 1 : (M == 2 ? 3 : 6);
  double p[D + 1][D];
  double m[D + 1][nc];
  double const* pp[D + 1];
  double const* mm[D + 1];
  for (int k = 0; k < D + 1; ++k) {
    load_vertex<D, nc>(coords, metric, ev2v[e * (D + 1) + k], p[k], m[k]);
    pp[k] = p[k];
    mm[k] = m[k];
  }
  bool u = false;
  out[e] = double(f32::element_quality<D, M>(pp, mm, &u));
  unsure[e] = u ? 1 : 0;
}

// the length (WHICH 1) or quality (WHICH 0) tag of a rebuilt mesh in one
// launch: threads [0, nsame) copy the old value of a same entity, the rest
// measure a product on the new mesh
template <int WHICH, int D, int M>
__global__ void __launch_bounds__(kBlock) measure_transfer_kernel(LO n, LO nsame,
    LO const* same2old, LO const* same2new, Real const* old_vals, LO const* prods2new,
    LO const* ev2v, Real const* coords, Real const* metric, Real* out) {
  LO i = thread_index();
  if (i >= n) return;
  if (i < nsame) {
    out[same2new[i]] = old_vals[same2old[i]];
    return;
  }
  LO e = prods2new[i - nsame];
  if constexpr (WHICH == 1)
    out[e] = length_of<D, M>(e, ev2v, coords, metric);
  else
    out[e] = quality_of<D, M>(e, ev2v, coords, metric);
}

template <int M>
__global__ void mident_kernel(LO n, LO const* edges, LO const* ev2v, Real const* metric,
    Real* out, int* error) {
  LO i = thread_index();
  if (i >= n) return;
  LO e = edges[i];
  constexpr int nc = (M == 1) ? 1 : (M == 2 ? 3 : 6);
  double a[nc], b[nc];
  for (int k = 0; k < nc; ++k) {
    a[k] = metric[ev2v[2 * e] * nc + k];
    b[k] = metric[ev2v[2 * e + 1] * nc + k];
  }
  int bad = 0;
  average_metric2<M>(a, b, out + i * nc, &bad);
  if (bad) *error = kDegenerateEigen;
}

// real_simplex_size (src/Omega_h_shape.hpp:77-100,166-170) of each element
template <int D>
__global__ void sizes_kernel(LO n, LO const* ev2v, Real const* coords, Real* out) {
  LO e = thread_index();
  if (e >= n) return;
  double b[3][3];
  LO v0 = ev2v[e * (D + 1)];
  for (int i = 0; i < D; ++i) {
    LO v = ev2v[e * (D + 1) + i + 1];
    for (int j = 0; j < D; ++j) b[i][j] = coords[v * D + j] - coords[v0 * D + j];
  }
  if (D == 3) {
    double c[3];
    cross3(b[0], b[1], c);
    out[e] = dot<3>(c, b[2]) / 6.0;
  } else {
    out[e] = (b[0][0] * b[1][1] - b[0][1] * b[1][0]) / 2.0;
  }
}

int metric_dim(DMesh const& m) { return (m.metric_ncomps == 1) ? 1 : m.dim; }

}  // namespace

// the kernels read a 6-component metric row with 16-byte loads, a tet's
// vertices with one 16-byte load and an edge's with one 8-byte load
static void check_alignment(int dim, int md, int nverts_per_ent, LO const* ev2v,
    Real const* metric) {
  auto off = [](void const* p, std::uintptr_t a) { return std::uintptr_t(p) % a; };
  OSH_CHECK(!(dim == 3 && md == 3) || off(metric, 16) == 0,
      "a 6-component metric array must be 16-byte aligned");
  OSH_CHECK(off(ev2v, nverts_per_ent == 4 ? 16 : (nverts_per_ent == 2 ? 8 : 4)) == 0,
      "the vertex list array is not aligned to its row size");
}

void launch_edge_lengths(Ctx& c, int dim, int md, LO n, LO const* edges, LO const* ev2v,
    Real const* coords, Real const* metric, Real* out) {
  check_alignment(dim, md, 2, ev2v, metric);
  if (dim == 3 && md == 3)
    OSH_LAUNCH((edge_lengths_kernel<3, 3>), n, c, edges, ev2v, coords, metric, out);
  else if (dim == 3)
    OSH_LAUNCH((edge_lengths_kernel<3, 1>), n, c, edges, ev2v, coords, metric, out);
  else if (md == 2)
    OSH_LAUNCH((edge_lengths_kernel<2, 2>), n, c, edges, ev2v, coords, metric, out);
  else
    OSH_LAUNCH((edge_lengths_kernel<2, 1>), n, c, edges, ev2v, coords, metric, out);
}

void launch_elem_qualities(Ctx& c, int dim, int md, LO n, LO const* elems, LO const* ev2v,
    Real const* coords, Real const* metric, Real* out) {
  check_alignment(dim, md, dim + 1, ev2v, metric);
  if (dim == 3 && md == 3)
    OSH_LAUNCH((elem_qualities_kernel<3, 3>), n, c, elems, ev2v, coords, metric, out);
  else if (dim == 3)
    OSH_LAUNCH((elem_qualities_kernel<3, 1>), n, c, elems, ev2v, coords, metric, out);
  else if (md == 2)
    OSH_LAUNCH((elem_qualities_kernel<2, 2>), n, c, elems, ev2v, coords, metric, out);
  else
    OSH_LAUNCH((elem_qualities_kernel<2, 1>), n, c, elems, ev2v, coords, metric, out);
}

void screen_estimates(Ctx& c, DMesh& m, Reals& lengths32, Reals& quals32, Bytes& unsure) {
  int dim = m.dim, md = metric_dim(m);
  LO ne = m.nents[1], nt = m.nents[dim];
  lengths32 = Reals(c, ne);
  quals32 = Reals(c, nt);
  unsure = Bytes(c, nt);
  LO const* ev2v = m.down[1].data();
  LO const* tv2v = verts_of(c, m, dim).data();
  Real const* x = m.coords.data();
  Real const* mt = m.metric.data();
#define OSH_EST(D, M)                                                                    \
  OSH_LAUNCH((edge_lengths32_kernel<D, M>), ne, c, ev2v, x, mt, lengths32.data());          \
  OSH_LAUNCH((elem_qualities32_kernel<D, M>), nt, c, tv2v, x, mt, quals32.data(),           \
      unsure.data())
  if (dim == 3 && md == 3) {
    OSH_EST(3, 3);
  } else if (dim == 3) {
    OSH_EST(3, 1);
  } else if (md == 2) {
    OSH_EST(2, 2);
  } else {
    OSH_EST(2, 1);
  }
#undef OSH_EST
}

void measure_edges(Ctx& c, DMesh& m, LO n, LO const* edges, Real* out) {
  launch_edge_lengths(c, m.dim, metric_dim(m), n, edges, m.down[1].data(), m.coords.data(),
      m.metric.data(), out);
}

void measure_elems(Ctx& c, DMesh& m, LO n, LO const* elems, Real* out) {
  LO const* ev2v = verts_of(c, m, m.dim).data();
  launch_elem_qualities(c, m.dim, metric_dim(m), n, elems, ev2v, m.coords.data(),
      m.metric.data(), out);
}

Reals const& ask_lengths(Ctx& c, DMesh& m) {
  if (!m.length.exists()) {
    Range range(c, "measure_edges_metric");
    m.length = Reals(c, m.nents[1]);
    measure_edges(c, m, m.nents[1], nullptr, m.length.data());
  }
  return m.length;
}

Reals const& ask_sizes(Ctx& c, DMesh& m) {
  if (!m.size.exists()) {
    LO n = m.nents[m.dim];
    m.size = Reals(c, n);
    LO const* ev2v = verts_of(c, m, m.dim).data();
    if (m.dim == 3)
      OSH_LAUNCH(sizes_kernel<3>, n, c, ev2v, m.coords.data(), m.size.data());
    else
      OSH_LAUNCH(sizes_kernel<2>, n, c, ev2v, m.coords.data(), m.size.data());
  }
  return m.size;
}

Reals const& ask_qualities(Ctx& c, DMesh& m) {
  if (!m.quality.exists()) {
    Range range(c, "measure_qualities");
    m.quality = Reals(c, m.nents[m.dim]);
    measure_elems(c, m, m.nents[m.dim], nullptr, m.quality.data());
  }
  return m.quality;
}

Reals measure_transfer(Ctx& c, DMesh& nw, int d, LO nsame, LO const* same2old,
    LO const* same2new, Real const* old_vals, LO nprods, LO const* prods2new) {
  OSH_CHECK(d == 1 || d == nw.dim, "lengths of edges or qualities of elements only");
  Reals out(c, nw.nents[d]);
  int md = metric_dim(nw);
  LO const* ev2v = (d == 1) ? nw.down[1].data() : verts_of(c, nw, d).data();
  check_alignment(nw.dim, md, d + 1, ev2v, nw.metric.data());
  LO n = nsame + nprods;
#define OSH_MT(W, D, M)                                                                    \
  OSH_LAUNCH((measure_transfer_kernel<W, D, M>), n, c, nsame, same2old, same2new, old_vals, \
      prods2new, ev2v, nw.coords.data(), nw.metric.data(), out.data())
  if (d == 1) {
    if (nw.dim == 3 && md == 3) OSH_MT(1, 3, 3);
    else if (nw.dim == 3) OSH_MT(1, 3, 1);
    else if (md == 2) OSH_MT(1, 2, 2);
    else OSH_MT(1, 2, 1);
  } else {
    if (nw.dim == 3 && md == 3) OSH_MT(0, 3, 3);
    else if (nw.dim == 3) OSH_MT(0, 3, 1);
    else if (md == 2) OSH_MT(0, 2, 2);
    else OSH_MT(0, 2, 1);
  }
#undef OSH_MT
  return out;
}

void mident_values(Ctx& c, DMesh& m, LO n, LO const* edges, Real const* data, int ncomps,
    Real* out) {
  int md = (ncomps == 1) ? 1 : (ncomps == 3 ? 2 : 3);
  OSH_CHECK(ncomps == 1 || ncomps == 3 || ncomps == 6, "a metric has 1, 3 or 6 components");
  if (md == 3)
    OSH_LAUNCH(mident_kernel<3>, n, c, edges, m.down[1].data(), data, out, c.error_word());
  else if (md == 2)
    OSH_LAUNCH(mident_kernel<2>, n, c, edges, m.down[1].data(), data, out, c.error_word());
  else
    OSH_LAUNCH(mident_kernel<1>, n, c, edges, m.down[1].data(), data, out, c.error_word());
}

void mident_metrics(Ctx& c, DMesh& m, LO n, LO const* edges, Real* out) {
  mident_values(c, m, n, edges, m.metric.data(), m.metric_ncomps, out);
}

}  // namespace osh

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

template <int S, int M>
__global__ void __launch_bounds__(kBlock) edge_lengths_kernel(LO n, LO const* edges,
    LO const* ev2v, Real const* coords, Real const* metric, Real* out) {
  LO i = thread_index();
  if (i >= n) return;
  LO e = edges ? edges[i] : i;
  constexpr int nc = (M == 1) ? 1 : (M == 2 ? 3 : 6);
  LO a = ev2v[2 * e], b = ev2v[2 * e + 1];
  double p0[S], p1[S], m0[nc], m1[nc];
  for (int k = 0; k < S; ++k) {
    p0[k] = coords[a * S + k];
    p1[k] = coords[b * S + k];
  }
  for (int k = 0; k < nc; ++k) {
    m0[k] = metric[a * nc + k];
    m1[k] = metric[b * nc + k];
  }
  out[i] = edge_length<S, M>(p0, p1, m0, m1);
}

template <int D, int M>
__global__ void __launch_bounds__(kBlock) elem_qualities_kernel(LO n, LO const* elems,
    LO const* ev2v, Real const* coords, Real const* metric, Real* out) {
  LO i = thread_index();
  if (i >= n) return;
  LO e = elems ? elems[i] : i;
  constexpr int nc = (M == 1) ? 1 : (M == 2 ? 3 : 6);
  double p[D + 1][D];
  double m[D + 1][nc];
  double const* pp[D + 1];
  double const* mm[D + 1];
  for (int k = 0; k < D + 1; ++k) {
    LO v = ev2v[e * (D + 1) + k];
    for (int j = 0; j < D; ++j) p[k][j] = coords[v * D + j];
    for (int j = 0; j < nc; ++j) m[k][j] = metric[v * nc + j];
    pp[k] = p[k];
    mm[k] = m[k];
  }
  out[i] = element_quality<D, M>(pp, mm);
}

template <int M>
__global__ void mident_kernel(LO n, LO const* edges, LO const* ev2v, Real const* metric,
    Real* out) {
  LO i = thread_index();
  if (i >= n) return;
  LO e = edges[i];
  constexpr int nc = (M == 1) ? 1 : (M == 2 ? 3 : 6);
  double a[nc], b[nc];
  for (int k = 0; k < nc; ++k) {
    a[k] = metric[ev2v[2 * e] * nc + k];
    b[k] = metric[ev2v[2 * e + 1] * nc + k];
  }
  average_metric2<M>(a, b, out + i * nc);
}

int metric_dim(DMesh const& m) { return (m.metric_ncomps == 1) ? 1 : m.dim; }

}  // namespace

void launch_edge_lengths(Ctx& c, int dim, int md, LO n, LO const* edges, LO const* ev2v,
    Real const* coords, Real const* metric, Real* out) {
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
  if (dim == 3 && md == 3)
    OSH_LAUNCH((elem_qualities_kernel<3, 3>), n, c, elems, ev2v, coords, metric, out);
  else if (dim == 3)
    OSH_LAUNCH((elem_qualities_kernel<3, 1>), n, c, elems, ev2v, coords, metric, out);
  else if (md == 2)
    OSH_LAUNCH((elem_qualities_kernel<2, 2>), n, c, elems, ev2v, coords, metric, out);
  else
    OSH_LAUNCH((elem_qualities_kernel<2, 1>), n, c, elems, ev2v, coords, metric, out);
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
    m.length = Reals(c, m.nents[1]);
    measure_edges(c, m, m.nents[1], nullptr, m.length.data());
  }
  return m.length;
}

Reals const& ask_qualities(Ctx& c, DMesh& m) {
  if (!m.quality.exists()) {
    m.quality = Reals(c, m.nents[m.dim]);
    measure_elems(c, m, m.nents[m.dim], nullptr, m.quality.data());
  }
  return m.quality;
}

void mident_metrics(Ctx& c, DMesh& m, LO n, LO const* edges, Real* out) {
  int md = metric_dim(m);
  if (md == 3)
    OSH_LAUNCH(mident_kernel<3>, n, c, edges, m.down[1].data(), m.metric.data(), out);
  else if (md == 2)
    OSH_LAUNCH(mident_kernel<2>, n, c, edges, m.down[1].data(), m.metric.data(), out);
  else
    OSH_LAUNCH(mident_kernel<1>, n, c, edges, m.down[1].data(), m.metric.data(), out);
}

}  // namespace osh

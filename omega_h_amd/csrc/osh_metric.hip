// Metric preparation on the device: gradation limiting by metric
// intersection, and intersection of metric arrays.
//
//   limit_gradation_once_tmpl  src/Omega_h_metric.cpp:156-180
//   limit_metric_gradation     src/Omega_h_metric.cpp:200-220
//   are_close (arrays)         src/Omega_h_array_ops.cpp:75-87,
//                              src/Omega_h_scalar.hpp:266-277
#include "osh_metric.hpp"
#include "osh_passes.hpp"

namespace osh {

namespace {

template <int S, int M>
__global__ void limit_gradation_kernel(LO nv, LO const* a2ab, LO const* ab2b,
    Real const* coords, Real const* in, Real max_rate, Real* out) {
  LO v = xcd_thread_index();
  if (v >= nv) return;
  constexpr int nc = (M * (M + 1)) / 2;
  Mat m = get_symm<M>(in + std::size_t(v) * nc);
  double x[S];
  for (int i = 0; i < S; ++i) x[i] = coords[std::size_t(v) * S + i];
  for (LO j = a2ab[v]; j < a2ab[v + 1]; ++j) {
    LO av = ab2b[j];
    Mat am = get_symm<M>(in + std::size_t(av) * nc);
    double vec[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < S; ++i) vec[i] = coords[std::size_t(av) * S + i] - x[i];
    double dist = math::sqrt_(metric_product<S, M>(am, vec));
    double h = 1.0 + dist * max_rate;
    double factor = 1.0 / (h * h);
    for (int c = 0; c < M; ++c)
      for (int r = 0; r < M; ++r) am.m[c][r] = am.m[c][r] * factor;
    m = intersect_metrics<M>(m, am);
  }
  set_symm<M>(m, out + std::size_t(v) * nc);
}

__global__ void not_close_kernel(LO n, Real const* a, Real const* b, Real tol, Real floor,
    I8* out) {
  LO i = thread_index();
  if (i >= n) return;
  double am = math::fabs_(a[i]), bm = math::fabs_(b[i]);
  double rd = (am <= floor && bm <= floor) ? 0.0
                                           : math::fabs_(b[i] - a[i]) / math::max2(am, bm);
  out[i] = !(rd <= tol);
}

template <int M>
__global__ void intersect_kernel(LO n, Real const* a, Real const* b, Real* out) {
  LO i = thread_index();
  if (i >= n) return;
  constexpr int nc = (M * (M + 1)) / 2;
  Mat r = intersect_metrics<M>(get_symm<M>(a + std::size_t(i) * nc),
      get_symm<M>(b + std::size_t(i) * nc));
  set_symm<M>(r, out + std::size_t(i) * nc);
}

// interpolate_between_metrics, src/Omega_h_metric.cpp:106-111: linearize
// (log), a (1 - t) + b t per stored component
// (src/Omega_h_array_ops.cpp:567-574), delinearize (exp)
template <int M>
__global__ void interpolate_kernel(LO n, Real const* a, Real const* b, Real t, Real* out) {
  LO i = thread_index();
  if (i >= n) return;
  constexpr int nc = (M * (M + 1)) / 2;
  double va[6], vb[6], vc[6];
  set_symm<M>(log_spd<M>(get_symm<M>(a + std::size_t(i) * nc)), va);
  set_symm<M>(log_spd<M>(get_symm<M>(b + std::size_t(i) * nc)), vb);
  for (int k = 0; k < nc; ++k) vc[k] = va[k] * (1.0 - t) + vb[k] * t;
  set_symm<M>(exp_spd<M>(get_symm<M>(vc)), out + std::size_t(i) * nc);
}

Reals limit_gradation_once(Ctx& c, DMesh& m, Reals const& values, Real max_rate) {
  LO nv = m.nents[0];
  Adj const& star = ask_star(c, m, 0);
  Reals out(c, values.size());
  int md = (m.metric_ncomps == 1) ? 1 : m.dim;
#define OSH_LG(S, M)                                                                       \
  OSH_LAUNCH((limit_gradation_kernel<S, M>), nv, c, star.a2ab.data(), star.ab2b.data(),    \
      m.coords.data(), values.data(), max_rate, out.data())
  if (m.dim == 3 && md == 3)
    OSH_LG(3, 3);
  else if (m.dim == 3)
    OSH_LG(3, 1);
  else if (md == 2)
    OSH_LG(2, 2);
  else
    OSH_LG(2, 1);
#undef OSH_LG
  return out;
}

}  // namespace

int limit_metric_gradation(Ctx& c, DMesh& m, Real max_rate, Real tol) {
  OSH_CHECK(max_rate > 0.0, "max_rate must be positive");
  OSH_CHECK(m.metric.exists(), "mesh has no metric");
  Reals values2 = m.metric;
  Reals values;
  LO n = values2.size();
  Bytes flags(c, n);
  int steps = 0;
  bool close;
  do {
    values = values2;
    values2 = limit_gradation_once(c, m, values, max_rate);
    ++steps;
    OSH_LAUNCH(not_close_kernel, n, c, values.data(), values2.data(), tol, kMetricEpsilon,
        flags.data());
    close = max_i8(c, flags) <= 0;
    OSH_CHECK(close || steps < 10000, "gradation limiting does not converge in 10000 sweeps");
  } while (!close);
  // a new metric invalidates cached lengths and qualities
  // (src/Omega_h_mesh.cpp:167-177)
  m.metric = values2;
  m.length = Reals();
  m.quality = Reals();
  return steps;
}

void interpolate_metric_arrays(Ctx& c, int dim, LO n, Reals const& a, Reals const& b, Real t,
    Reals& out) {
  out = Reals(c, a.size());
  if (dim == 3)
    OSH_LAUNCH(interpolate_kernel<3>, n, c, a.data(), b.data(), t, out.data());
  else if (dim == 2)
    OSH_LAUNCH(interpolate_kernel<2>, n, c, a.data(), b.data(), t, out.data());
  else
    OSH_LAUNCH(interpolate_kernel<1>, n, c, a.data(), b.data(), t, out.data());
}

namespace {

void replace_metric(DMesh& m, Reals const& v) {
  m.metric = v;
  m.length = Reals();
  m.quality = Reals();
}

bool acceptable(Ctx& c, DMesh& m, AdaptOptions const& o) {
  if (min_real(c, ask_qualities(c, m)) < o.min_quality_allowed) return false;
  return !(max_real(c, ask_lengths(c, m)) > o.max_length_allowed);
}

}  // namespace

bool approach_metric(Ctx& c, DMesh& m, AdaptOptions const& o, Real min_step) {
  auto it = m.vtags.begin();
  for (; it != m.vtags.end(); ++it)
    if (it->name == "target_metric") break;
  if (it == m.vtags.end()) return false;
  OSH_CHECK(it->ncomps == m.metric_ncomps, "target_metric and metric differ in size");
  Real minq = min_real(c, ask_qualities(c, m));
  OSH_CHECK(minq >= o.min_quality_allowed,
      "approach_metric: minimum element quality " + std::to_string(minq) +
          " < minimum allowed quality " + std::to_string(o.min_quality_allowed));
  Real maxl = max_real(c, ask_lengths(c, m));
  OSH_CHECK(!(maxl > o.max_length_allowed),
      "approach_metric: maximum edge length " + std::to_string(maxl) +
          " > maximum allowed length " + std::to_string(o.max_length_allowed));
  Reals orig = m.metric;
  Reals target = it->data;
  replace_metric(m, target);
  if (acceptable(c, m, o)) {
    m.vtags.erase(it);
    return true;
  }
  int mdim = (m.metric_ncomps == 1) ? 1 : m.dim;
  Real factor = 1.0;
  do {
    factor /= 2.0;
    OSH_CHECK(factor >= min_step, "approach_metric has stalled at step size " +
                                      std::to_string(factor) + " < " + std::to_string(min_step));
    Reals cur;
    interpolate_metric_arrays(c, mdim, m.nents[0], orig, target, factor, cur);
    replace_metric(m, cur);
  } while (!acceptable(c, m, o));
  return true;
}

void intersect_metric_arrays(Ctx& c, int dim, LO n, Reals const& a, Reals const& b,
    Reals& out) {
  out = Reals(c, a.size());
  if (dim == 3)
    OSH_LAUNCH(intersect_kernel<3>, n, c, a.data(), b.data(), out.data());
  else if (dim == 2)
    OSH_LAUNCH(intersect_kernel<2>, n, c, a.data(), b.data(), out.data());
  else
    OSH_LAUNCH(intersect_kernel<1>, n, c, a.data(), b.data(), out.data());
}

}  // namespace osh

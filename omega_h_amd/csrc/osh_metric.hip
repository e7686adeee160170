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

// size-scaled implied metric of each element (src/Omega_h_metric.cpp:257-301)
template <int D>
__global__ void element_implied_kernel(LO ne, LO const* ev2v, Real const* coords, Real scalar,
    Real* e2m) {
  LO e = xcd_thread_index();
  if (e >= ne) return;
  constexpr int nc = (D * (D + 1)) / 2;
  double p[D + 1][3];
  for (int k = 0; k <= D; ++k) {
    LO v = ev2v[e * (D + 1) + k];
    for (int j = 0; j < D; ++j) p[k][j] = coords[std::size_t(v) * D + j];
  }
  double x[6];
  element_implied_metric<D>(p, x);
  for (int k = 0; k < nc; ++k) e2m[std::size_t(e) * nc + k] = x[k] * scalar;
}

// project_metrics, src/Omega_h_metric.cpp:221-245: log-Euclidean average of
// the metrics of a vertex's elements (average_metric_contrib/_finish with
// degenerate metrics skipped, src/Omega_h_metric.hpp:125-139)
template <int D>
__global__ void project_metrics_kernel(LO nv, LO const* a2ab, LO const* ab2b, Real const* e2m,
    Real* out) {
  LO v = thread_index();
  if (v >= nv) return;
  constexpr int nc = (D * (D + 1)) / 2;
  Mat am;
  mat_zero(am);
  int n = 0;
  for (LO j = a2ab[v]; j < a2ab[v + 1]; ++j) {
    Mat em = get_symm<D>(e2m + std::size_t(ab2b[j]) * nc);
    double mx = 0.0;
    for (int c = 0; c < D; ++c)
      for (int r = 0; r < D; ++r) mx = math::max2(mx, math::fabs_(em.m[c][r]));
    if (mx < kMetricEpsilon) continue;
    Mat lm = log_spd<D>(em);
    for (int c = 0; c < D; ++c)
      for (int r = 0; r < D; ++r) am.m[c][r] = am.m[c][r] + lm.m[c][r];
    ++n;
  }
  Mat r = am;
  if (n > 0) {
    for (int c = 0; c < D; ++c)
      for (int rr = 0; rr < D; ++rr) am.m[c][rr] = am.m[c][rr] / n;
    r = exp_spd<D>(am);
  }
  set_symm<D>(r, out + std::size_t(v) * nc);
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

Reals implied_metrics(Ctx& c, DMesh& m) {
  int dim = m.dim;
  int nc = (dim * (dim + 1)) / 2;
  LO ne = m.nents[dim];
  // power(typical_unit_simplex_size / equilateral_simplex_size, 2, dim)
  double typical = dim == 3 ? 0.0838934100219 : 0.377645136635;
  double perfect = dim == 3 ? 0.1178511301977579 : 0.4330127018922193;
  double s = typical / perfect;
  double scalar = dim == 3 ? math::cbrt_(s * s) : s;
  Reals e2m(c, ne * nc);
  LOs const& ev2v = verts_of(c, m, dim);
  Adj const& v2e = ask_up(c, m, 0, dim);
  Reals out(c, m.nents[0] * nc);
  if (dim == 3) {
    OSH_LAUNCH(element_implied_kernel<3>, ne, c, ev2v.data(), m.coords.data(), scalar, e2m.data());
    OSH_LAUNCH(project_metrics_kernel<3>, m.nents[0], c, v2e.a2ab.data(), v2e.ab2b.data(),
        e2m.data(), out.data());
  } else {
    OSH_LAUNCH(element_implied_kernel<2>, ne, c, ev2v.data(), m.coords.data(), scalar, e2m.data());
    OSH_LAUNCH(project_metrics_kernel<2>, m.nents[0], c, v2e.a2ab.data(), v2e.ab2b.data(),
        e2m.data(), out.data());
  }
  return out;
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

// the first move of warp_to_limit: coordinates + warp
__global__ void warp_add_kernel(LO n, Real const* coords, Real const* warp, Real* out) {
  LO i = thread_index();
  if (i >= n) return;
  out[i] = coords[i] + warp[i];
}

// one halving of warp_to_limit (src/Omega_h_approach.cpp:66-70): the warp
// halves, the remainder (0 before the first halving) gains the half, and
// the vertices move by the half from where they started
__global__ void warp_halve_kernel(LO n, Real const* coords, Real const* warp_in,
    Real* warp_out, Real* rem, bool first, Real* out) {
  LO i = thread_index();
  if (i >= n) return;
  Real h = warp_in[i] / 2.0;
  warp_out[i] = h;
  Real r = first ? 0.0 : rem[i];
  rem[i] = r + h;
  out[i] = coords[i] + h;
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

void set_coords(DMesh& m, Reals const& v) {
  m.coords = v;
  m.length = Reals();
  m.quality = Reals();
  m.size = Reals();
}

bool warp_to_limit(Ctx& c, DMesh& m, AdaptOptions const& o, bool exit_on_stall, int max_niters) {
  auto it = m.vtags.begin();
  for (; it != m.vtags.end(); ++it)
    if (it->name == "warp") break;
  if (it == m.vtags.end()) return false;
  OSH_CHECK(it->ncomps == m.dim, "the warp tag needs one vector per vertex");
  Real minq = min_real(c, ask_qualities(c, m));
  OSH_CHECK(minq >= o.min_quality_allowed,
      "warp_to_limit: minimum element quality " + std::to_string(minq) +
          " < minimum allowed quality " + std::to_string(o.min_quality_allowed));
  Real maxl = max_real(c, ask_lengths(c, m));
  OSH_CHECK(!(maxl > o.max_length_allowed),
      "warp_to_limit: maximum edge length " + std::to_string(maxl) +
          " > maximum allowed length " + std::to_string(o.max_length_allowed));
  LO n = m.nents[0] * m.dim;
  Reals coords = m.coords;
  Reals warp = it->data;
  {
    Reals moved(c, n);
    OSH_LAUNCH(warp_add_kernel, n, c, coords.data(), warp.data(), moved.data());
    set_coords(m, moved);
  }
  if (acceptable(c, m, o)) {
    m.vtags.erase(it);
    return true;
  }
  // the halved warp goes to an array of its own: the tag's array may be
  // shared with copies of this mesh
  Reals half(c, n), rem(c, n);
  int i = 0;
  do {
    ++i;
    if (i > max_niters) {
      OSH_CHECK(exit_on_stall, "warp step " + std::to_string(i) +
                                   ": unable to satisfy this warp under this size field");
      for (it = m.vtags.begin(); it != m.vtags.end(); ++it)
        if (it->name == "warp") break;
      m.vtags.erase(it);
      return true;
    }
    Reals moved(c, n);
    OSH_LAUNCH(warp_halve_kernel, n, c, coords.data(), (i == 1) ? warp.data() : half.data(),
        half.data(), rem.data(), i == 1, moved.data());
    set_coords(m, moved);
  } while (!acceptable(c, m, o));
  for (auto& t : m.vtags)
    if (t.name == "warp") t.data = rem;
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

// Metric preparation on the device: gradation limiting by metric
// intersection, and intersection of metric arrays.
//
//   limit_gradation_once_tmpl  src/Omega_h_metric.cpp:156-180
//   limit_metric_gradation     src/Omega_h_metric.cpp:200-220
//   are_close (arrays)         src/Omega_h_array_ops.cpp:75-87,
//                              src/Omega_h_scalar.hpp:266-277
// Repairing an initial mesh's quality and the grade-fix-adapt driver:
//   apply_isotropy             src/Omega_h_metric.cpp:576-612,
//                              src/Omega_h_eigen.cpp:7-40
//   compute_ill_metric_dim     src/Omega_h_fix.cpp:16-68
//   fix_for_given_metric, fix  src/Omega_h_fix.cpp:90-134
//   grade_fix_adapt            src/Omega_h_fix.cpp:136-176
#include <cmath>
#include <cstdio>

#include "osh_metric.hpp"
#include "osh_passes.hpp"

namespace osh {

namespace {

template <int S, int M>
__global__ void limit_gradation_kernel(LO nv, LO const* a2ab, LO const* ab2b,
    Real const* coords, Real const* in, Real max_rate, Real* out, int* error) {
  LO v = xcd_thread_index();
  if (v >= nv) return;
  int bad = 0;
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
    m = intersect_metrics<M>(m, am, &bad);
  }
  set_symm<M>(m, out + std::size_t(v) * nc);
  if (bad) *error = kDegenerateEigen;
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
__global__ void intersect_kernel(LO n, Real const* a, Real const* b, Real* out, int* error) {
  LO i = thread_index();
  if (i >= n) return;
  constexpr int nc = (M * (M + 1)) / 2;
  int bad = 0;
  Mat r = intersect_metrics<M>(get_symm<M>(a + std::size_t(i) * nc),
      get_symm<M>(b + std::size_t(i) * nc), &bad);
  set_symm<M>(r, out + std::size_t(i) * nc);
  if (bad) *error = kDegenerateEigen;
}

// interpolate_between_metrics, src/Omega_h_metric.cpp:106-111: linearize
// (log), a (1 - t) + b t per stored component
// (src/Omega_h_array_ops.cpp:567-574), delinearize (exp)
template <int M>
__global__ void interpolate_kernel(LO n, Real const* a, Real const* b, Real t, Real* out,
    int* error) {
  LO i = thread_index();
  if (i >= n) return;
  constexpr int nc = (M * (M + 1)) / 2;
  double va[6], vb[6], vc[6];
  int bad = 0;
  set_symm<M>(log_spd<M>(get_symm<M>(a + std::size_t(i) * nc), &bad), va);
  set_symm<M>(log_spd<M>(get_symm<M>(b + std::size_t(i) * nc), &bad), vb);
  for (int k = 0; k < nc; ++k) vc[k] = va[k] * (1.0 - t) + vb[k] * t;
  set_symm<M>(exp_spd<M>(get_symm<M>(vc), &bad), out + std::size_t(i) * nc);
  if (bad) *error = kDegenerateEigen;
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
    Real* out, int* error) {
  LO v = thread_index();
  if (v >= nv) return;
  int bad = 0;
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
    Mat lm = log_spd<D>(em, &bad);
    for (int c = 0; c < D; ++c)
      for (int r = 0; r < D; ++r) am.m[c][r] = am.m[c][r] + lm.m[c][r];
    ++n;
  }
  Mat r = am;
  if (n > 0) {
    for (int c = 0; c < D; ++c)
      for (int rr = 0; rr < D; ++rr) am.m[c][rr] = am.m[c][rr] / n;
    r = exp_spd<D>(am, &bad);
  }
  set_symm<D>(r, out + std::size_t(v) * nc);
  if (bad) *error = kDegenerateEigen;
}

Reals limit_gradation_once(Ctx& c, DMesh& m, Reals const& values, Real max_rate) {
  LO nv = m.nents[0];
  Adj const& star = ask_star(c, m, 0);
  Reals out(c, values.size());
  int md = (m.metric_ncomps == 1) ? 1 : m.dim;
#define OSH_LG(S, M)                                                                       \
  OSH_LAUNCH((limit_gradation_kernel<S, M>), nv, c, star.a2ab.data(), star.ab2b.data(),    \
      m.coords.data(), values.data(), max_rate, out.data(), c.error_word())
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

Reals element_implied_array(Ctx& c, DMesh& m, Real scalar) {
  int dim = m.dim;
  int nc = (dim * (dim + 1)) / 2;
  LO ne = m.nents[dim];
  Reals e2m(c, ne * nc);
  LOs const& ev2v = verts_of(c, m, dim);
  if (dim == 3)
    OSH_LAUNCH(element_implied_kernel<3>, ne, c, ev2v.data(), m.coords.data(), scalar, e2m.data());
  else
    OSH_LAUNCH(element_implied_kernel<2>, ne, c, ev2v.data(), m.coords.data(), scalar, e2m.data());
  return e2m;
}

Reals project_metric_array(Ctx& c, DMesh& m, int mdim, Reals const& e2m) {
  Adj const& v2e = ask_up(c, m, 0, m.dim);
  int nc = (mdim * (mdim + 1)) / 2;
  Reals out(c, m.nents[0] * nc);
  if (mdim == 3)
    OSH_LAUNCH(project_metrics_kernel<3>, m.nents[0], c, v2e.a2ab.data(), v2e.ab2b.data(),
        e2m.data(), out.data(), c.error_word());
  else if (mdim == 2)
    OSH_LAUNCH(project_metrics_kernel<2>, m.nents[0], c, v2e.a2ab.data(), v2e.ab2b.data(),
        e2m.data(), out.data(), c.error_word());
  else
    OSH_LAUNCH(project_metrics_kernel<1>, m.nents[0], c, v2e.a2ab.data(), v2e.ab2b.data(),
        e2m.data(), out.data(), c.error_word());
  return out;
}

Reals implied_metrics(Ctx& c, DMesh& m) {
  int dim = m.dim;
  // power(typical_unit_simplex_size / equilateral_simplex_size, 2, dim)
  double typical = dim == 3 ? 0.0838934100219 : 0.377645136635;
  double perfect = dim == 3 ? 0.1178511301977579 : 0.4330127018922193;
  double s = typical / perfect;
  double scalar = dim == 3 ? math::cbrt_(s * s) : s;
  return project_metric_array(c, m, dim, element_implied_array(c, m, scalar));
}

void interpolate_metric_arrays(Ctx& c, int dim, LO n, Reals const& a, Reals const& b, Real t,
    Reals& out) {
  out = Reals(c, a.size());
  if (dim == 3)
    OSH_LAUNCH(interpolate_kernel<3>, n, c, a.data(), b.data(), t, out.data(), c.error_word());
  else if (dim == 2)
    OSH_LAUNCH(interpolate_kernel<2>, n, c, a.data(), b.data(), t, out.data(), c.error_word());
  else
    OSH_LAUNCH(interpolate_kernel<1>, n, c, a.data(), b.data(), t, out.data(), c.error_word());
}

namespace {

// apply_isotropy of n metrics of dimension M: 1 (OMEGA_H_ISO_LENGTH) the
// largest eigenvalue, 2 (OMEGA_H_ISO_SIZE) root<M>(det)
// (src/Omega_h_metric.cpp:576-612)
template <int M>
__global__ void isotropy_kernel(LO n, Real const* in, int isotropy, Real* out) {
  LO i = thread_index();
  if (i >= n) return;
  constexpr int nc = (M * (M + 1)) / 2;
  Mat m = get_symm<M>(in + std::size_t(i) * nc);
  if (isotropy == 1) {
    out[i] = max_eigenvalue<M>(m);
  } else {
    double d = det<M>(m);
    out[i] = (M == 3) ? math::cbrt_(d) : math::sqrt_(d);
  }
}

// get_mident_metrics(mesh, dim, v2m, has_degen = true)
// (src/Omega_h_metric.cpp:56-104): each element's vertex metrics averaged,
// degenerate ones skipped (src/Omega_h_metric.hpp:125-149)
template <int D, int M>
__global__ void element_mident_kernel(LO ne, LO const* ev2v, Real const* v2m, Real* out,
    int* error) {
  LO e = thread_index();
  if (e >= ne) return;
  int bad = 0;
  constexpr int nc = (M * (M + 1)) / 2;
  Mat am;
  mat_zero(am);
  int n = 0;
  for (int k = 0; k <= D; ++k) {
    Mat vm = get_symm<M>(v2m + std::size_t(ev2v[e * (D + 1) + k]) * nc);
    double mx = 0.0;
    for (int cc = 0; cc < M; ++cc)
      for (int r = 0; r < M; ++r) mx = math::max2(mx, math::fabs_(vm.m[cc][r]));
    if (mx < kMetricEpsilon) continue;
    Mat lm = log_spd<M>(vm, &bad);
    for (int cc = 0; cc < M; ++cc)
      for (int r = 0; r < M; ++r) am.m[cc][r] = am.m[cc][r] + lm.m[cc][r];
    ++n;
  }
  Mat r = am;
  if (n > 0) {
    for (int cc = 0; cc < M; ++cc)
      for (int rr = 0; rr < M; ++rr) am.m[cc][rr] = am.m[cc][rr] / n;
    r = exp_spd<M>(am, &bad);
  }
  set_symm<M>(r, out + std::size_t(e) * nc);
  if (bad) *error = kDegenerateEigen;
}

// compute_ill_metric_dim (src/Omega_h_fix.cpp:40-48): the element metrics of
// elements below the allowed quality stop contributing
__global__ void zero_low_quality_kernel(LO ne, int nc, Real const* q, Real floor,
    Real const* in, Real* out) {
  LO e = thread_index();
  if (e >= ne) return;
  bool low = q[e] < floor;
  for (int k = 0; k < nc; ++k)
    out[std::size_t(e) * nc + k] = low ? 0. : in[std::size_t(e) * nc + k];
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

// n metrics of dimension mdim made isotropic (a copy for isotropy 0 or
// scalar metrics); *out_mdim is the dimension of the result
Reals apply_isotropy_array(Ctx& c, int mdim, LO n, Reals const& in, int isotropy,
    int* out_mdim) {
  if (isotropy == 0 || mdim == 1) {
    *out_mdim = mdim;
    return in;
  }
  *out_mdim = 1;
  Reals out(c, n);
  if (mdim == 3)
    OSH_LAUNCH(isotropy_kernel<3>, n, c, in.data(), isotropy, out.data());
  else
    OSH_LAUNCH(isotropy_kernel<2>, n, c, in.data(), isotropy, out.data());
  return out;
}

void set_mesh_metric(DMesh& m, Reals const& v, int ncomps) {
  m.metric = v;
  m.metric_ncomps = ncomps;
  m.length = Reals();
  m.quality = Reals();
}

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
  if (o.verbosity >= 2) std::printf("approach_metric moved by factor %g\n", factor);
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
    if (o.verbosity >= 2) std::printf("warp_to_limit completed in one step\n");
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
      std::printf("warp_to_limit stalled, dropping warp field and continuing anyway\n");
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
  if (o.verbosity >= 2) std::printf("warp_to_limit moved by factor %g\n", std::ldexp(1.0, -i));
  for (auto& t : m.vtags)
    if (t.name == "warp") t.data = rem;
  return true;
}

Reals implied_isos(Ctx& c, DMesh& m) {
  int mdim;
  return apply_isotropy_array(c, m.dim, m.nents[0], implied_metrics(c, m), 2, &mdim);
}

void ill_metric(Ctx& c, DMesh& m, AdaptOptions const& o, int isotropy) {
  OSH_CHECK(isotropy >= 0 && isotropy <= 2, "isotropy must be 0, 1 or 2");
  int dim = m.dim;
  LO ne = m.nents[dim];
  int mdim;
  Reals elem = apply_isotropy_array(c, dim, ne, element_implied_array(c, m, 1.0), isotropy,
      &mdim);
  int nc = (mdim * (mdim + 1)) / 2;
  Reals vm = project_metric_array(c, m, mdim, elem);
  if (isotropy == 0) {
    LOs const& ev2v = verts_of(c, m, dim);
    for (int i = 0; i < 5; ++i) {
      Reals mid(c, ne * nc);
      if (dim == 3)
        OSH_LAUNCH((element_mident_kernel<3, 3>), ne, c, ev2v.data(), vm.data(), mid.data(),
            c.error_word());
      else
        OSH_LAUNCH((element_mident_kernel<2, 2>), ne, c, ev2v.data(), vm.data(), mid.data(),
            c.error_word());
      vm = project_metric_array(c, m, mdim, mid);
    }
  }
  set_mesh_metric(m, vm, nc);
  Reals const& q = ask_qualities(c, m);
  Reals kept(c, ne * nc);
  OSH_LAUNCH(zero_low_quality_kernel, ne, c, nc, q.data(), o.min_quality_allowed, elem.data(),
      kept.data());
  set_mesh_metric(m, project_metric_array(c, m, mdim, kept), nc);
  limit_metric_gradation(c, m, 1.0, 1e-2);
}

int fix_for_given_metric(Ctx& c, DMesh& m, AdaptOptions const& o) {
  Real minqual = min_real(c, ask_qualities(c, m));
  Real maxlen = max_real(c, ask_lengths(c, m));
  AdaptOptions opts = o;
  int calls = 0;
  while (true) {
    Real minqual_old = minqual;
    opts.min_quality_allowed = minqual;
    Real twice = opts.max_length_desired * 2.0;
    opts.max_length_allowed = (twice > maxlen) ? twice : maxlen;
    opts.nsliver_layers = 10;
    Real want = minqual + 0.1;
    opts.min_quality_desired = (1.0 < want) ? 1.0 : want;
    adapt(c, m, opts, nullptr);
    ++calls;
    minqual = min_real(c, ask_qualities(c, m));
    maxlen = max_real(c, ask_lengths(c, m));
    if (minqual < minqual_old + 1e-3) break;
  }
  return calls;
}

int fix(Ctx& c, DMesh& m, AdaptOptions const& o, int isotropy) {
  ill_metric(c, m, o, isotropy);
  return fix_for_given_metric(c, m, o);
}

int grade_fix_adapt(Ctx& c, DMesh& m, AdaptOptions const& o, int ncomps, Reals const& target) {
  int dim = m.dim;
  OSH_CHECK(ncomps == 1 || ncomps == (dim * (dim + 1)) / 2, "target metric size");
  OSH_CHECK(target.size() == m.nents[0] * ncomps, "target metric size");
  int metric_dim = (ncomps == 1) ? 1 : dim;
  // limit_metric_gradation(mesh, target, 0.5) over the mesh's own star
  Reals had = m.metric;
  int had_ncomps = m.metric_ncomps;
  set_mesh_metric(m, target, ncomps);
  Reals graded;
  try {
    limit_metric_gradation(c, m, 0.5, 1e-2);
    graded = m.metric;
  } catch (...) {
    set_mesh_metric(m, had, had_ncomps);  // the caller's mesh as it was
    throw;
  }
  bool found = false;
  for (auto& t : m.vtags)
    if (t.name == "target_metric") {
      t = {"target_metric", ncomps, 1, graded};
      found = true;
    }
  if (!found) m.vtags.push_back({"target_metric", ncomps, 1, graded});
  if (had.exists())
    set_mesh_metric(m, had, had_ncomps);
  else if (metric_dim == 1)
    set_mesh_metric(m, implied_isos(c, m), 1);
  else
    set_mesh_metric(m, implied_metrics(c, m), (dim * (dim + 1)) / 2);
  if (min_real(c, ask_qualities(c, m)) < o.min_quality_allowed)
    fix(c, m, o, (metric_dim == 1) ? 1 : 0);
  int steps = 0;
  while (approach_metric(c, m, o, 1e-4)) {
    adapt(c, m, o, nullptr);
    ++steps;
  }
  return steps;
}

void intersect_metric_arrays(Ctx& c, int dim, LO n, Reals const& a, Reals const& b,
    Reals& out) {
  out = Reals(c, a.size());
  if (dim == 3)
    OSH_LAUNCH(intersect_kernel<3>, n, c, a.data(), b.data(), out.data(), c.error_word());
  else if (dim == 2)
    OSH_LAUNCH(intersect_kernel<2>, n, c, a.data(), b.data(), out.data(), c.error_word());
  else
    OSH_LAUNCH(intersect_kernel<1>, n, c, a.data(), b.data(), out.data(), c.error_word());
}

}  // namespace osh

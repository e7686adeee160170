/* Oracle (test infrastructure only): repairing an initial mesh's quality and
 * the grade-fix-adapt driver.
 *
 *   apply_isotropy            src/Omega_h_metric.cpp:576-612
 *     get_max_eigenvalues     src/Omega_h_eigen.cpp:7-40
 *   compute_ill_metric_dim    src/Omega_h_fix.cpp:16-68
 *     smooth_metric_once      src/Omega_h_metric.cpp:254-257
 *   fix_for_given_metric      src/Omega_h_fix.cpp:90-118
 *   fix                       src/Omega_h_fix.cpp:120-134
 *   grade_fix_adapt           src/Omega_h_fix.cpp:136-176
 *   add_implied_isos_tag      src/Omega_h_metric_input.cpp:271-280
 *
 * Isotropy codes follow Omega_h_Isotropy: 0 anisotropic, 1 OMEGA_H_ISO_LENGTH,
 * 2 OMEGA_H_ISO_SIZE. */
#include <stdlib.h>
#include <string.h>

#include "om_internal.h"

/* n metrics of dimension mdim -> n metrics (isotropy 0) or n scalars */
static double* apply_isotropy(int mdim, int n, double const* in, int isotropy, int* out_mdim) {
  int nc = om_symm_ncomps(mdim);
  if (isotropy == 0 || mdim == 1) {
    double* out = (double*)om_malloc(sizeof(double) * ((size_t)n * nc + 1));
    memcpy(out, in, sizeof(double) * (size_t)n * nc);
    *out_mdim = mdim;
    return out;
  }
  double* out = (double*)om_malloc(sizeof(double) * ((size_t)n + 1));
  for (int i = 0; i < n; ++i) {
    om_mat m;
    om_get_symm(mdim, in + (size_t)i * nc, &m);
    if (isotropy == 1) {
      out[i] = om_max_eigenvalue(mdim, &m);
    } else {
      double d = om_det(mdim, &m);
      out[i] = (mdim == 3) ? om_cbrt(d) : om_sqrt(d);
    }
  }
  *out_mdim = 1;
  return out;
}

/* get_implied_isos (src/Omega_h_metric.cpp:316-321): the implied metrics
 * made isotropic with the same volume */
void om_get_implied_isos(om_mesh* m, double* out) {
  int dim = m->dim;
  int nv = m->nents[0];
  double* full = (double*)om_malloc(sizeof(double) * ((size_t)nv * om_symm_ncomps(dim) + 1));
  om_get_implied_metrics(m, full);
  int mdim;
  double* iso = apply_isotropy(dim, nv, full, 2, &mdim);
  memcpy(out, iso, sizeof(double) * (size_t)nv);
  free(iso);
  free(full);
}

/* compute_ill_metric_dim: a metric implied by the elements, smoothed (when
 * anisotropic), with the contributions of elements of low quality in it
 * removed, and graded at rate 1; it becomes the mesh's metric */
void om_ill_metric(om_mesh* m, om_adapt_opts const* o, int isotropy) {
  int dim = m->dim;
  int ne = m->nents[dim];
  int nv = m->nents[0];
  double* full = (double*)om_malloc(sizeof(double) * ((size_t)ne * om_symm_ncomps(dim) + 1));
  om_element_implied_metrics(m, 1.0, full);
  int mdim;
  double* elem = apply_isotropy(dim, ne, full, isotropy, &mdim);
  free(full);
  int nc = om_symm_ncomps(mdim);
  double* vm = (double*)om_malloc(sizeof(double) * ((size_t)nv * nc + 1));
  om_project_metrics(m, mdim, elem, vm);
  if (isotropy == 0) {
    double* mid = (double*)om_malloc(sizeof(double) * ((size_t)ne * nc + 1));
    for (int i = 0; i < 5; ++i) {
      om_element_mident_metrics(m, mdim, vm, mid);
      om_project_metrics(m, mdim, mid, vm);
    }
    free(mid);
  }
  om_set_metric(m, nc, vm);
  double* q = (double*)om_malloc(sizeof(double) * ((size_t)ne + 1));
  om_get_qualities(m, q);
  for (int e = 0; e < ne; ++e)
    if (q[e] < o->min_quality_allowed)
      for (int k = 0; k < nc; ++k) elem[(size_t)e * nc + k] = 0.;
  om_project_metrics(m, mdim, elem, vm);
  om_set_metric(m, nc, vm);
  om_limit_metric_gradation(m, 1.0, 1e-2);
  free(q);
  free(vm);
  free(elem);
}

/* fix_for_given_metric: adapt with the allowed quality raised to what the
 * mesh has, until the minimum quality stops improving by 1e-3; returns the
 * number of adapt() calls */
int om_fix_for_given_metric(om_mesh* m, om_adapt_opts const* o) {
  double minqual = om_min_quality(m);
  double maxlen = om_max_length(m);
  om_adapt_opts opts = *o;
  int calls = 0;
  while (1) {
    double minqual_old = minqual;
    opts.min_quality_allowed = minqual;
    double twice = opts.max_length_desired * 2.0;
    opts.max_length_allowed = (twice > maxlen) ? twice : maxlen;
    opts.nsliver_layers = 10;
    double want = minqual + 0.1;
    opts.min_quality_desired = (1.0 < want) ? 1.0 : want;
    om_adapt(m, &opts, NULL);
    ++calls;
    minqual = om_min_quality(m);
    maxlen = om_max_length(m);
    if (minqual < minqual_old + 1e-3) break;
  }
  return calls;
}

int om_fix(om_mesh* m, om_adapt_opts const* o, int isotropy) {
  om_ill_metric(m, o, isotropy);
  return om_fix_for_given_metric(m, o);
}

/* grade_fix_adapt: grade the target (rate 0.5), keep it as "target_metric",
 * derive the implied metric when the mesh has none, repair the mesh when
 * its quality is below the allowed minimum, then
 * while (approach_metric) adapt.  Returns the number of adapt() calls after
 * the repair, or the negative code of a failed approach_metric. */
int om_grade_fix_adapt(om_mesh* m, om_adapt_opts const* o, int ncomps, double const* target) {
  int dim = m->dim;
  int nv = m->nents[0];
  int metric_dim = (ncomps == 1) ? 1 : dim;
  size_t n = (size_t)nv * ncomps;
  /* limit_metric_gradation(mesh, target, 0.5) on the mesh's own star */
  int had_ncomps = m->metric_ncomps;
  double* had = NULL;
  if (m->metric) {
    had = (double*)om_malloc(sizeof(double) * ((size_t)nv * had_ncomps + 1));
    memcpy(had, m->metric, sizeof(double) * (size_t)nv * had_ncomps);
  }
  om_set_metric(m, ncomps, target);
  om_limit_metric_gradation(m, 0.5, 1e-2);
  double* graded = (double*)om_malloc(sizeof(double) * (n + 1));
  memcpy(graded, m->metric, sizeof(double) * n);
  om_remove_vtag(m, "target_metric");
  om_add_vtag(m, "target_metric", ncomps, OM_METRIC, graded);
  free(graded);
  if (had) {
    om_set_metric(m, had_ncomps, had);
    free(had);
  } else {
    double* implied;
    int nc = (metric_dim == 1) ? 1 : om_symm_ncomps(dim);
    implied = (double*)om_malloc(sizeof(double) * ((size_t)nv * nc + 1));
    if (metric_dim == 1)
      om_get_implied_isos(m, implied);
    else
      om_get_implied_metrics(m, implied);
    om_set_metric(m, nc, implied);
    free(implied);
  }
  if (om_min_quality(m) < o->min_quality_allowed) om_fix(m, o, (metric_dim == 1) ? 1 : 0);
  int steps = 0;
  int r;
  while ((r = om_approach_metric(m, o, 1e-4)) == 1) {
    om_adapt(m, o, NULL);
    ++steps;
  }
  return (r < 0) ? r : steps;
}

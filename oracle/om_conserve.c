/* Oracle (test infrastructure only): conservative and momentum transfer
 * through adapt(), restated serially from the reference:
 *
 *   setup_conservation_tags              src/Omega_h_conserve.cpp:578-606
 *   separate_cavities / separate_by_color src/Omega_h_conserve.cpp:65-181
 *   carry_class_bdry_integ_error         src/Omega_h_conserve.cpp:183-207
 *   introduce_class_integ_error          src/Omega_h_conserve.cpp:209-243
 *   transfer_integ_error and the density
 *   and momentum errors                  src/Omega_h_conserve.cpp:245-336
 *   OpConservation flags per operation   src/Omega_h_conserve.cpp:294-305,
 *                                        372-389,460-537
 *   transfer_by_intersection             src/Omega_h_conserve.cpp:391-455
 *   diffusion and correct_integral_errors src/Omega_h_conserve.cpp:608-829
 *
 * A cavity is a key's old elements (the key's upward row) and its new
 * elements (its products, in product order).  Cavities are sorted into
 * NOT_BDRY (no old element has a vertex off the interior), KEY_BDRY (the key
 * is classified below the mesh dimension) and TOUCH_BDRY (the rest); KEY_BDRY
 * cavities are split into sub-cavities of one element class_id each, in the
 * order in which a class first appears among the cavity's old elements (new
 * elements of a class no old element has form a last sub-cavity with no old
 * elements).  Element errors ("<integral>_error") are introduced where an
 * operation does not conserve (new minus old integral of the sub-cavity) and
 * carried where they may already exist (the old sub-cavity's error), spread
 * over the new elements in proportion to their sizes.  Every sum starts from
 * 0.0 and runs in list order. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "om_internal.h"

enum { NOT_BDRY = 0, TOUCH_BDRY = 1, KEY_BDRY = 2 };

static double min2(double a, double b) { return (b < a) ? b : a; }
static double max2(double a, double b) { return (a < b) ? b : a; }

static double* sizes_of(om_mesh* m) {
  int dim = m->dim;
  double* s = (double*)om_malloc(sizeof(double) * ((size_t)m->nents[dim] + 1));
  om_get_sizes(m, s);
  return s;
}

/* an element touches the boundary when one of its vertices is classified
 * below the mesh dimension (get_elems_are_bdry, :125-131) */
static signed char* elems_are_bdry(om_mesh* m) {
  int dim = m->dim;
  int const* ev2v = om_verts_of(m, dim);
  signed char* b = (signed char*)om_calloc((size_t)m->nents[dim] + 1, 1);
  for (int e = 0; e < m->nents[dim]; ++e)
    for (int k = 0; k <= dim; ++k)
      if (m->class_dim[0][ev2v[e * (dim + 1) + k]] != dim) b[e] = 1;
  return b;
}

/* element averages of a vertex field (average_field, src/Omega_h_mesh.cpp:822-844) */
static double* elem_average(om_mesh* m, int nc, double const* vdata) {
  int dim = m->dim;
  int ne = m->nents[dim];
  int const* ev2v = om_verts_of(m, dim);
  double* out = (double*)om_malloc(sizeof(double) * ((size_t)ne * nc + 1));
  for (int e = 0; e < ne; ++e)
    for (int j = 0; j < nc; ++j) {
      double comp = 0;
      for (int k = 0; k <= dim; ++k) comp += vdata[(size_t)ev2v[e * (dim + 1) + k] * nc + j];
      comp /= (dim + 1);
      out[(size_t)e * nc + j] = comp;
    }
  return out;
}

static om_etag* find_etag(om_mesh* m, char const* name) {
  for (int t = 0; t < m->netags; ++t)
    if (strcmp(m->etags[t].name, name) == 0) return &m->etags[t];
  return NULL;
}

int om_conserve_any(om_mesh* m) {
  for (int t = 0; t < m->netags; ++t)
    if (m->etags[t].kind == OM_CONSERVE) return 1;
  return 0;
}

static int densities_or_conserved(om_mesh* m) {
  for (int t = 0; t < m->netags; ++t)
    if (m->etags[t].kind == OM_CONSERVE || m->etags[t].kind == OM_DENSITY) return 1;
  return 0;
}

/* setup_conservation_tags: zero errors for every conserved density and
 * momentum velocity */
void om_conserve_setup(om_mesh* m) {
  int ne = m->nents[m->dim];
  for (int t = 0; t < m->netags; ++t) {
    om_etag* g = &m->etags[t];
    if (g->kind != OM_CONSERVE) continue;
    free(g->err);
    g->err = (double*)om_calloc((size_t)ne * g->ncomps + 1, sizeof(double));
  }
  for (int t = 0; t < m->nvtags; ++t) {
    om_vtag* g = &m->vtags[t];
    if (g->kind != OM_MOMENTUM) continue;
    free(g->err);
    g->err = (double*)om_calloc((size_t)ne * g->ncomps + 1, sizeof(double));
  }
}

/* ---- sub-cavities: (key, old list, new list) */
typedef struct {
  int n;
  int* key;  /* n */
  int* o0;   /* n + 1 offsets into olist */
  int* olist;
  int* n0;   /* n + 1 offsets into nlist */
  int* nlist;
} subcavs;

static void subcavs_free(subcavs* s) {
  free(s->key);
  free(s->o0);
  free(s->olist);
  free(s->n0);
  free(s->nlist);
  memset(s, 0, sizeof(*s));
}

typedef struct {
  int nkeys;
  int const* o_a2ab; /* old elements per key */
  int* o_ab2b;
  int const* k2p;    /* new elements per key */
  int* n_ab2b;
  int* category;     /* per key */
} cavities;

/* the sub-cavities of one category; `colored` splits each cavity by
 * element class_id (separate_by_color), otherwise one sub-cavity per key */
static void make_subcavs(om_mesh* old, om_mesh* nw, cavities const* c, int cat, int colored,
    subcavs* s) {
  int dim = old->dim;
  int cap_s = 1, cap_o = 1, cap_n = 1;
  for (int k = 0; k < c->nkeys; ++k) {
    if (c->category[k] != cat) continue;
    int no = c->o_a2ab[k + 1] - c->o_a2ab[k];
    int nn = c->k2p[k + 1] - c->k2p[k];
    cap_s += no + nn + 1;
    cap_o += no;
    cap_n += nn;
  }
  s->n = 0;
  s->key = (int*)om_malloc(sizeof(int) * (size_t)cap_s);
  s->o0 = (int*)om_malloc(sizeof(int) * (size_t)(cap_s + 1));
  s->n0 = (int*)om_malloc(sizeof(int) * (size_t)(cap_s + 1));
  s->olist = (int*)om_malloc(sizeof(int) * (size_t)cap_o);
  s->nlist = (int*)om_malloc(sizeof(int) * (size_t)cap_n);
  int no_tot = 0, nn_tot = 0;
  s->o0[0] = 0;
  s->n0[0] = 0;
  int const* ocls = old->class_id[dim];
  int const* ncls = nw->class_id[dim];
  int* otaken = NULL;
  int* ntaken = NULL;
  for (int k = 0; k < c->nkeys; ++k) {
    if (c->category[k] != cat) continue;
    int ob = c->o_a2ab[k], oe = c->o_a2ab[k + 1];
    int nb = c->k2p[k], ne = c->k2p[k + 1];
    if (!colored) {
      for (int i = ob; i < oe; ++i) s->olist[no_tot++] = c->o_ab2b[i];
      for (int i = nb; i < ne; ++i) s->nlist[nn_tot++] = c->n_ab2b[i];
      s->key[s->n] = k;
      ++s->n;
      s->o0[s->n] = no_tot;
      s->n0[s->n] = nn_tot;
      continue;
    }
    otaken = (int*)realloc(otaken, sizeof(int) * (size_t)(oe - ob + 1));
    ntaken = (int*)realloc(ntaken, sizeof(int) * (size_t)(ne - nb + 1));
    memset(otaken, 0, sizeof(int) * (size_t)(oe - ob + 1));
    memset(ntaken, 0, sizeof(int) * (size_t)(ne - nb + 1));
    int oleft = oe - ob, nleft = ne - nb;
    while (oleft > 0 || nleft > 0) {
      int any_old = 0, color = 0;
      for (int i = ob; i < oe; ++i)
        if (!otaken[i - ob]) {
          any_old = 1;
          color = ocls[c->o_ab2b[i]];
          break;
        }
      for (int i = ob; i < oe; ++i) {
        if (otaken[i - ob]) continue;
        if (ocls[c->o_ab2b[i]] == color) {
          otaken[i - ob] = 1;
          --oleft;
          s->olist[no_tot++] = c->o_ab2b[i];
        }
      }
      for (int i = nb; i < ne; ++i) {
        if (ntaken[i - nb]) continue;
        if (!any_old || ncls[c->n_ab2b[i]] == color) {
          ntaken[i - nb] = 1;
          --nleft;
          s->nlist[nn_tot++] = c->n_ab2b[i];
        }
      }
      s->key[s->n] = k;
      ++s->n;
      s->o0[s->n] = no_tot;
      s->n0[s->n] = nn_tot;
    }
  }
  free(otaken);
  free(ntaken);
}

/* introduce_class_integ_error: (new - old) integral of each sub-cavity,
 * spread over its new elements by size, added into err */
static void introduce(subcavs const* s, int nc, double const* odens, double const* osize,
    double const* ndens, double const* nsize, double* err) {
  for (int i = 0; i < s->n; ++i) {
    double nsum = 0.0;
    for (int j = s->n0[i]; j < s->n0[i + 1]; ++j) nsum = nsum + nsize[s->nlist[j]];
    for (int q = 0; q < nc; ++q) {
      double oint = 0.0, nint = 0.0;
      for (int j = s->o0[i]; j < s->o0[i + 1]; ++j) {
        int e = s->olist[j];
        oint = oint + odens[(size_t)e * nc + q] * osize[e];
      }
      for (int j = s->n0[i]; j < s->n0[i + 1]; ++j) {
        int e = s->nlist[j];
        nint = nint + ndens[(size_t)e * nc + q] * nsize[e];
      }
      double cerr = nint - oint;
      double d;
      if (nsum != 0.0) {
        d = cerr / nsum;
      } else {
        if (cerr != 0.0) {
          fprintf(stderr, "oracle: integral error in a cavity of zero size\n");
          abort();
        }
        d = 0.0;
      }
      for (int j = s->n0[i]; j < s->n0[i + 1]; ++j) {
        int e = s->nlist[j];
        err[(size_t)e * nc + q] += d * nsize[e];
      }
    }
  }
}

/* carry_class_bdry_integ_error: the old sub-cavity's error, spread over its
 * new elements by size, added into err */
static void carry(subcavs const* s, int nc, double const* oerr, double const* nsize,
    double* err) {
  for (int i = 0; i < s->n; ++i) {
    double nsum = 0.0;
    for (int j = s->n0[i]; j < s->n0[i + 1]; ++j) nsum = nsum + nsize[s->nlist[j]];
    for (int q = 0; q < nc; ++q) {
      double oe = 0.0;
      for (int j = s->o0[i]; j < s->o0[i + 1]; ++j) oe = oe + oerr[(size_t)s->olist[j] * nc + q];
      double d;
      if (nsum != 0.0) {
        d = oe / nsum;
      } else {
        if (oe != 0.0) {
          fprintf(stderr, "oracle: integral error in a cavity of zero size\n");
          abort();
        }
        d = 0.0;
      }
      for (int j = s->n0[i]; j < s->n0[i + 1]; ++j) {
        int e = s->nlist[j];
        err[(size_t)e * nc + q] += d * nsize[e];
      }
    }
  }
}

/* transfer_by_intersection_dim: each new element of a sub-cavity takes the
 * intersection-weighted mean of the old elements' values */
static void by_intersection(om_mesh* old, om_mesh* nw, subcavs const* s, int nc,
    double const* odata, double* ndata) {
  int dim = old->dim;
  int const* oev2v = om_verts_of(old, dim);
  int const* nev2v = om_verts_of(nw, dim);
  for (int i = 0; i < s->n; ++i) {
    for (int j = s->n0[i]; j < s->n0[i + 1]; ++j) {
      int ne = s->nlist[j];
      double npts[12], opts[12];
      for (int k = 0; k <= dim; ++k)
        for (int x = 0; x < dim; ++x)
          npts[k * dim + x] = nw->coords[(size_t)nev2v[ne * (dim + 1) + k] * dim + x];
      for (int q = 0; q < nc; ++q) ndata[(size_t)ne * nc + q] = 0;
      double total = 0.0;
      for (int jo = s->o0[i]; jo < s->o0[i + 1]; ++jo) {
        int oe = s->olist[jo];
        for (int k = 0; k <= dim; ++k)
          for (int x = 0; x < dim; ++x)
            opts[k * dim + x] = old->coords[(size_t)oev2v[oe * (dim + 1) + k] * dim + x];
        int bad = 0;
        double isz = om_intersect_size(dim, npts, opts, &bad);
        if (bad) {
          fprintf(stderr, "oracle: polytope buffer overflow\n");
          abort();
        }
        for (int q = 0; q < nc; ++q)
          ndata[(size_t)ne * nc + q] += isz * odata[(size_t)oe * nc + q];
        total += isz;
      }
      for (int q = 0; q < nc; ++q) ndata[(size_t)ne * nc + q] /= total;
    }
  }
}

/* the new mesh's error array of one conserved quantity: same elements keep
 * their errors, products start at zero, then introduce / carry per
 * category (transfer_integ_error) */
static double* transfer_errors(om_mesh* old, om_mesh* nw, int const* same2old,
    int const* same2new, int nsame, subcavs const* cls, int const* this_time,
    int const* always, int nc, double const* oerr, double const* odens, double const* osize,
    double const* ndens, double const* nsize) {
  int nne = nw->nents[nw->dim];
  double* err = (double*)om_calloc((size_t)nne * nc + 1, sizeof(double));
  for (int i = 0; i < nsame; ++i)
    for (int q = 0; q < nc; ++q)
      err[(size_t)same2new[i] * nc + q] = oerr[(size_t)same2old[i] * nc + q];
  for (int cat = 0; cat < 3; ++cat) {
    if (!this_time[cat]) introduce(&cls[cat], nc, odens, osize, ndens, nsize, err);
    if (!always[cat]) carry(&cls[cat], nc, oerr, nsize, err);
  }
  return err;
}

static double* momenta_of(om_mesh* m, om_vtag const* vel, double const* vdata) {
  int dim = m->dim;
  int ne = m->nents[dim];
  om_etag* dens = find_etag(m, vel->density);
  double* ev = elem_average(m, vel->ncomps, vdata);
  for (int e = 0; e < ne; ++e)
    for (int q = 0; q < vel->ncomps; ++q) ev[(size_t)e * vel->ncomps + q] *= dens->data[e];
  return ev;
}

/* densities by intersection (coarsening and swap) and the errors of every
 * conserved quantity, for one pass.  op: 0 refine, 1 coarsen, 2 swap.  The
 * new mesh's elements, coordinates, classification and vertex tags are
 * set; its element tags hold the products' values of the other transfers,
 * which this overwrites for densities. */
void om_conserve_pass(om_mesh* old, om_mesh* nw, int op, int key_dim, int nkeys,
    int const* keys2kds, int const* k2p, int const* prods2new, int nsame,
    int const* same2old, int const* same2new) {
  if (!densities_or_conserved(old)) return;
  int dim = old->dim;
  int nprods = k2p[nkeys];
  om_graph const* kd2e = om_ask_up(old, key_dim, dim);
  cavities c;
  c.nkeys = nkeys;
  int* o_a2ab = (int*)om_malloc(sizeof(int) * (size_t)(nkeys + 1));
  o_a2ab[0] = 0;
  for (int k = 0; k < nkeys; ++k)
    o_a2ab[k + 1] = o_a2ab[k] + kd2e->a2ab[keys2kds[k] + 1] - kd2e->a2ab[keys2kds[k]];
  c.o_a2ab = o_a2ab;
  c.o_ab2b = (int*)om_malloc(sizeof(int) * (size_t)(o_a2ab[nkeys] + 1));
  for (int k = 0; k < nkeys; ++k)
    for (int j = kd2e->a2ab[keys2kds[k]], i = o_a2ab[k]; j < kd2e->a2ab[keys2kds[k] + 1]; ++j, ++i)
      c.o_ab2b[i] = kd2e->ab2b[j];
  c.k2p = k2p;
  c.n_ab2b = (int*)om_malloc(sizeof(int) * (size_t)(nprods + 1));
  for (int p = 0; p < nprods; ++p) c.n_ab2b[p] = prods2new[p];
  c.category = (int*)om_malloc(sizeof(int) * (size_t)(nkeys + 1));
  signed char* obdry = elems_are_bdry(old);
  for (int k = 0; k < nkeys; ++k) {
    int nb = 0;
    for (int i = o_a2ab[k]; i < o_a2ab[k + 1]; ++i) nb += obdry[c.o_ab2b[i]];
    int key_bdry = old->class_dim[key_dim][keys2kds[k]] < dim;
    c.category[k] = key_bdry ? KEY_BDRY : (nb > 0 ? TOUCH_BDRY : NOT_BDRY);
  }
  free(obdry);
  subcavs whole[3], cls[3];
  for (int cat = 0; cat < 3; ++cat) {
    make_subcavs(old, nw, &c, cat, 0, &whole[cat]);
    make_subcavs(old, nw, &c, cat, cat == KEY_BDRY, &cls[cat]);
  }
  /* densities: intersection (coarsen over NOT, TOUCH and KEY by class;
   * swap over whole cavities); refinement products already inherited */
  if (op != 0) {
    subcavs all;
    if (op == 2) {
      for (int k = 0; k < nkeys; ++k) c.category[k] = 0;
      make_subcavs(old, nw, &c, 0, 0, &all);
    }
    for (int t = 0; t < old->netags; ++t) {
      om_etag const* a = &old->etags[t];
      if (a->kind != OM_CONSERVE && a->kind != OM_DENSITY) continue;
      om_etag* b = &nw->etags[t];
      if (op == 2) {
        by_intersection(old, nw, &all, a->ncomps, a->data, b->data);
      } else {
        by_intersection(old, nw, &whole[NOT_BDRY], a->ncomps, a->data, b->data);
        by_intersection(old, nw, &whole[TOUCH_BDRY], a->ncomps, a->data, b->data);
        by_intersection(old, nw, &cls[KEY_BDRY], a->ncomps, a->data, b->data);
      }
    }
    if (op == 2) subcavs_free(&all);
  }
  if (om_conserve_any(old)) {
    /* OpConservation: density always conserved away from the boundary;
     * this operation conserves density except in KEY_BDRY cavities of
     * coarsening and swap, and momentum only in refinement */
    int d_this[3] = {1, 1, op == 0 ? 1 : 0};
    int d_always[3] = {1, 0, 0};
    int m_this[3] = {op == 0, op == 0, op == 0};
    int m_always[3] = {0, 0, 0};
    double* osize = sizes_of(old);
    double* nsize = sizes_of(nw);
    for (int t = 0; t < old->netags; ++t) {
      om_etag const* a = &old->etags[t];
      if (a->kind != OM_CONSERVE || !a->err) continue;
      om_etag* b = &nw->etags[t];
      free(b->err);
      b->err = transfer_errors(old, nw, same2old, same2new, nsame, cls, d_this, d_always,
          a->ncomps, a->err, a->data, osize, b->data, nsize);
    }
    for (int t = 0; t < old->nvtags; ++t) {
      om_vtag const* a = &old->vtags[t];
      if (a->kind != OM_MOMENTUM || !a->err) continue;
      om_vtag* b = &nw->vtags[t];
      double* om = momenta_of(old, a, a->data);
      double* nm = momenta_of(nw, b, b->data);
      free(b->err);
      b->err = transfer_errors(old, nw, same2old, same2new, nsame, cls, m_this, m_always,
          a->ncomps, a->err, om, osize, nm, nsize);
      free(om);
      free(nm);
    }
    free(osize);
    free(nsize);
  }
  for (int cat = 0; cat < 3; ++cat) {
    subcavs_free(&whole[cat]);
    subcavs_free(&cls[cat]);
  }
  free(o_a2ab);
  free(c.o_ab2b);
  free(c.n_ab2b);
  free(c.category);
}

/* ---- correct_integral_errors */

/* the element graph across sides classified in the interior
 * (get_elem_diffusion_graph, elements_across_sides src/Omega_h_adj.cpp:622-657) */
static om_graph* diffusion_graph(om_mesh* m) {
  int dim = m->dim;
  int ne = m->nents[dim];
  int const* es2s = m->down[dim];
  om_graph const* s2e = om_ask_up(m, dim - 1, dim);
  om_graph* g = (om_graph*)om_calloc(1, sizeof(om_graph));
  g->n = ne;
  g->a2ab = (int*)om_malloc(sizeof(int) * (size_t)(ne + 1));
  g->a2ab[0] = 0;
  for (int e = 0; e < ne; ++e) {
    int deg = 0;
    for (int k = 0; k <= dim; ++k)
      if (m->class_dim[dim - 1][es2s[e * (dim + 1) + k]] == dim) ++deg;
    g->a2ab[e + 1] = g->a2ab[e] + deg;
  }
  g->ab2b = (int*)om_malloc(sizeof(int) * ((size_t)g->a2ab[ne] + 1));
  for (int e = 0; e < ne; ++e) {
    int i = g->a2ab[e];
    for (int k = 0; k <= dim; ++k) {
      int s = es2s[e * (dim + 1) + k];
      if (m->class_dim[dim - 1][s] != dim) continue;
      int se = s2e->a2ab[s];
      if (s2e->ab2b[se] == e) ++se;
      g->ab2b[i++] = s2e->ab2b[se];
    }
  }
  return g;
}

/* diffuse_densities: repeat diffuse_densities_once until every |density|
 * is at most tol (at most max_iters sweeps; returns -1 beyond) */
static int diffuse(om_mesh* m, om_graph const* g, double* dens, double const* csize, double tol,
    int max_iters) {
  int ne = g->n;
  int max_deg = m->dim + 1;
  double* out = (double*)om_malloc(sizeof(double) * ((size_t)ne + 1));
  int it = 0;
  for (;;) {
    int bounded = 1;
    for (int e = 0; e < ne; ++e)
      if (!(fabs(dens[e]) <= tol)) bounded = 0;
    if (bounded) break;
    if (it >= max_iters) {
      free(out);
      return -1;
    }
    for (int e = 0; e < ne; ++e) {
      double s = csize[e];
      double o = dens[e];
      for (int j = g->a2ab[e]; j < g->a2ab[e + 1]; ++j) {
        int oe = g->ab2b[j];
        double os = csize[oe];
        double mins = min2(s, os);
        double factor = mins / (s * max_deg);
        double x = dens[e];
        double ox = dens[oe];
        o += (ox - x) * factor;
      }
      out[e] = o;
    }
    memcpy(dens, out, sizeof(double) * (size_t)ne);
    ++it;
  }
  free(out);
  return it;
}

/* diffuse_integrals_weighted, one component at a time */
static int diffuse_weighted(om_mesh* m, om_graph const* g, om_diffuse const* opt, int nc,
    double* err, double const* quantity, int max_iters) {
  if (opt->type == OM_DIFFUSE_NONE) return 0;
  int ne = g->n;
  double* ws = (double*)om_malloc(sizeof(double) * ((size_t)ne + 1));
  double* wd = (double*)om_malloc(sizeof(double) * ((size_t)ne + 1));
  int rc = 0;
  for (int q = 0; q < nc && rc == 0; ++q) {
    for (int e = 0; e < ne; ++e) {
      ws[e] = max2(fabs(quantity[(size_t)e * nc + q]), opt->floor);
      double x = err[(size_t)e * nc + q];
      if (ws[e] != 0.0) {
        wd[e] = x / ws[e];
      } else {
        if (x != 0.0) rc = -2;
        wd[e] = 0.0;
      }
    }
    if (rc) break;
    if (diffuse(m, g, wd, ws, opt->tol, max_iters) < 0) rc = -1;
    for (int e = 0; e < ne; ++e) err[(size_t)e * nc + q] = wd[e] * ws[e];
  }
  free(ws);
  free(wd);
  return rc;
}

/* correct_integral_errors (with correct_density_error and
 * correct_momentum_error): returns 0, or nonzero where the reference would
 * not finish (diffusion beyond max_iters sweeps) or would fail a check */
int om_conserve_correct(om_mesh* m, int max_iters) {
  if (!om_conserve_any(m)) return 0;
  int dim = m->dim;
  int ne = m->nents[dim];
  int nv = m->nents[0];
  om_graph* g = diffusion_graph(m);
  double* sizes = sizes_of(m);
  int rc = 0;
  for (int t = 0; t < m->netags && rc == 0; ++t) {
    om_etag* a = &m->etags[t];
    if (a->kind != OM_CONSERVE || !a->err) continue;
    int nc = a->ncomps;
    free(a->old);
    a->old = (double*)om_malloc(sizeof(double) * ((size_t)ne * nc + 1));
    memcpy(a->old, a->data, sizeof(double) * (size_t)ne * nc);
    double* oint = (double*)om_malloc(sizeof(double) * ((size_t)ne * nc + 1));
    for (int e = 0; e < ne; ++e)
      for (int q = 0; q < nc; ++q) oint[(size_t)e * nc + q] = a->data[(size_t)e * nc + q] * sizes[e];
    rc = diffuse_weighted(m, g, &a->diffuse, nc, a->err, oint, max_iters);
    for (int e = 0; e < ne; ++e)
      for (int q = 0; q < nc; ++q) {
        size_t i = (size_t)e * nc + q;
        double nint = oint[i] - a->err[i];
        a->data[i] = nint / sizes[e];
      }
    free(oint);
    free(a->err);
    a->err = NULL;
  }
  om_graph const* v2e = om_ask_up(m, 0, dim);
  int const* ev2v = om_verts_of(m, dim);
  for (int t = 0; t < m->nvtags && rc == 0; ++t) {
    om_vtag* a = &m->vtags[t];
    if (a->kind != OM_MOMENTUM || !a->err) continue;
    int nc = a->ncomps;
    om_etag* dens = find_etag(m, a->density);
    double* mass = (double*)om_malloc(sizeof(double) * ((size_t)ne + 1));
    double* omass = (double*)om_malloc(sizeof(double) * ((size_t)ne + 1));
    for (int e = 0; e < ne; ++e) {
      mass[e] = dens->data[e] * sizes[e];
      omass[e] = dens->old[e] * sizes[e];
    }
    double* ev = elem_average(m, nc, a->data);
    double* newm = (double*)om_malloc(sizeof(double) * ((size_t)ne * nc + 1));
    for (int e = 0; e < ne; ++e)
      for (int q = 0; q < nc; ++q) {
        size_t i = (size_t)e * nc + q;
        double om = ev[i] * omass[e];
        double nm = ev[i] * mass[e];
        newm[i] = nm;
        a->err[i] = a->err[i] + (nm - om);
      }
    double* vmass = (double*)om_malloc(sizeof(double) * ((size_t)nv + 1));
    for (int v = 0; v < nv; ++v) {
      double s = 0.0;
      for (int j = v2e->a2ab[v]; j < v2e->a2ab[v + 1]; ++j) s = s + mass[v2e->ab2b[j]];
      vmass[v] = s / (double)(dim + 1);
    }
    rc = diffuse_weighted(m, g, &a->diffuse, nc, a->err, newm, max_iters);
    double const* fixed = NULL;
    for (int i = 0; i < m->nitags; ++i)
      if (strcmp(m->itags[i].name, "momentum_velocity_fixed") == 0) fixed = m->itags[i].data[0];
    double* out = (double*)om_malloc(sizeof(double) * ((size_t)nv * nc + 1));
    memcpy(out, a->data, sizeof(double) * (size_t)nv * nc);
    for (int v = 0; v < nv; ++v) {
      int vf = fixed ? (int)fixed[v] : 0;
      double vm = vmass[v];
      for (int j = v2e->a2ab[v]; j < v2e->a2ab[v + 1]; ++j) {
        int e = v2e->ab2b[j];
        double nfree[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k <= dim; ++k) {
          int v2 = ev2v[e * (dim + 1) + k];
          int f2 = fixed ? (int)fixed[v2] : 0;
          for (int q = 0; q < nc; ++q)
            if (!(f2 & (1 << q))) nfree[q] += 1.0;
        }
        for (int q = 0; q < nc; ++q)
          if (!(vf & (1 << q)))
            out[(size_t)v * nc + q] -= a->err[(size_t)e * nc + q] / (nfree[q] * vm);
      }
    }
    memcpy(a->data, out, sizeof(double) * (size_t)nv * nc);
    free(out);
    free(vmass);
    free(newm);
    free(ev);
    free(mass);
    free(omass);
    free(a->err);
    a->err = NULL;
  }
  for (int t = 0; t < m->netags; ++t) {
    free(m->etags[t].old);
    m->etags[t].old = NULL;
    if (rc) {
      free(m->etags[t].err);
      m->etags[t].err = NULL;
    }
  }
  if (rc)
    for (int t = 0; t < m->nvtags; ++t) {
      free(m->vtags[t].err);
      m->vtags[t].err = NULL;
    }
  om_graph_free(g);
  free(sizes);
  return rc;
}

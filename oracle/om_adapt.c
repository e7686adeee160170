/* Oracle (test infrastructure only): the passes of Omega_h::adapt(),
 * restated serially in C for one rank (globals equal local indices).
 *
 *   adapt / satisfy_lengths / satisfy_quality src/Omega_h_adapt.cpp:52-288
 *   refine_by_size / refine_qualities       src/Omega_h_refine.cpp:17-100,
 *                                           src/Omega_h_refine_qualities.cpp:13-109
 *   refine_products                         src/Omega_h_refine_topology.cpp:13-203
 *   coarsen_by_size / coarsen_slivers       src/Omega_h_coarsen.cpp:29-220
 *   check_collapse_class / _exposure        src/Omega_h_check_collapse.cpp:10-117
 *   prevent_coarsen_overshoot               src/Omega_h_coarsen_overshoot.cpp:12-49
 *   prevent_coarsen_flip / most normal      src/Omega_h_coarsen_flip.cpp:14-150,
 *                                           src/Omega_h_most_normal.hpp:15-88
 *   coarsen_qualities / filters             src/Omega_h_coarsen_qualities.cpp:16-122
 *   choose_rails / find_rails               src/Omega_h_collapse_rail.cpp:15-101
 *   mark_dead_ents / find_coarsen_domains / coarsen_topology
 *                                           src/Omega_h_coarsen_topology.cpp:15-122
 *   swap_edges / swap3d / swap2d            src/Omega_h_swap.cpp, src/Omega_h_swap3d*.cpp,
 *                                           src/Omega_h_swap2d*.cpp
 *   find_loop / choose                      src/Omega_h_swap3d_loop.hpp:24-78,
 *                                           src/Omega_h_swap3d_choice.hpp:16-100
 *   find_indset / QualityCompare            src/Omega_h_indset.cpp:5-35,
 *                                           src/Omega_h_indset_inline.hpp:12-72
 *   mark_sliver_layers / mark_down / mark_up src/Omega_h_mark.cpp:12-63,190-212
 *   modify_ents_adapt and numbering         src/Omega_h_modify.cpp:20-70,141-517
 *   transfer_refine / _coarsen / _swap      src/Omega_h_transfer.cpp:160-730
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "om_internal.h"

enum { NOT_IN = 0, IN = 1, UNKNOWN = 2 };
enum { DONT_COLLAPSE = 0, COLLAPSE_LEFT = 1, COLLAPSE_RIGHT = 2 };

static double min2(double a, double b) { return (b < a) ? b : a; }

static int* collect_marked(int n, signed char const* marks, int* count) {
  int c = 0;
  for (int i = 0; i < n; ++i)
    if (marks[i]) ++c;
  int* out = (int*)om_malloc(sizeof(int) * (size_t)c);
  int k = 0;
  for (int i = 0; i < n; ++i)
    if (marks[i]) out[k++] = i;
  *count = c;
  return out;
}

static int any_marked(int n, signed char const* marks) {
  for (int i = 0; i < n; ++i)
    if (marks[i] == 1) return 1;
  return 0;
}

void om_default_opts(int dim, om_adapt_opts* o) {
  o->min_length_desired = 1.0 / om_sqrt(2.0);
  o->max_length_desired = om_sqrt(2.0);
  o->max_length_allowed = o->max_length_desired * 2.0;
  if (dim == 3) {
    o->min_quality_allowed = 0.20;
    o->min_quality_desired = 0.30;
  } else {
    o->min_quality_allowed = 0.30;
    o->min_quality_desired = 0.40;
  }
  o->nsliver_layers = 4;
  o->verbosity = 2; /* EACH_REBUILD, as the reference; the oracle prints nothing */
  o->should_refine = 1;
  o->should_coarsen = 1;
  o->should_swap = 1;
  o->should_coarsen_slivers = 1;
  o->should_prevent_coarsen_flip = 0;
}

/* ------------------------------------------------------------ indset */
void om_find_indset(om_mesh* m, int d, double const* q, signed char const* cands,
    signed char* out) {
  int n = m->nents[d];
  if (d == m->dim) {
    memcpy(out, cands, (size_t)n);
    return;
  }
  om_graph const* g = om_ask_star(m, d);
  signed char* old = (signed char*)om_malloc((size_t)n);
  for (int i = 0; i < n; ++i) old[i] = cands[i] ? UNKNOWN : NOT_IN;
  for (int round = 0;; ++round) {
    if (round > n + 1) {
      fprintf(stderr, "oracle: independent set does not converge\n");
      abort();
    }
    int any = 0;
    for (int i = 0; i < n; ++i)
      if (old[i] == UNKNOWN) any = 1;
    if (!any) break;
    memcpy(out, old, (size_t)n);
    for (int v = 0; v < n; ++v) {
      if (old[v] != UNKNOWN) continue;
      int b = g->a2ab[v], e = g->a2ab[v + 1];
      int rejected = 0;
      for (int j = b; j < e; ++j)
        if (old[g->ab2b[j]] == IN) {
          out[v] = NOT_IN;
          rejected = 1;
          break;
        }
      if (rejected) continue;
      int local_max = 1;
      for (int j = b; j < e; ++j) {
        int u = g->ab2b[j];
        if (old[u] == NOT_IN) continue;
        int beats;
        if (q[u] != q[v])
          beats = q[u] < q[v];
        else
          beats = u < v;
        if (!beats) {
          local_max = 0;
          break;
        }
      }
      if (local_max) out[v] = IN;
    }
    memcpy(old, out, (size_t)n);
  }
  memcpy(out, old, (size_t)n);
  free(old);
}

void om_kat_indset(om_mesh* m, int d, double const* q, signed char const* c,
    signed char* out) {
  om_find_indset(m, d, q, c, out);
}

/* ------------------------------------------------------------ marking */
static signed char* mark_down(om_mesh* m, int high, int low,
    signed char const* marks) {
  int nl = m->nents[low];
  signed char* out = (signed char*)om_calloc((size_t)nl, 1);
  if (high == low) {
    memcpy(out, marks, (size_t)nl);
    return out;
  }
  om_graph const* l2h = om_ask_up(m, low, high);
  for (int l = 0; l < nl; ++l)
    for (int j = l2h->a2ab[l]; j < l2h->a2ab[l + 1]; ++j)
      if (marks[l2h->ab2b[j]]) out[l] = 1;
  return out;
}

static signed char* mark_up(om_mesh* m, int low, int high,
    signed char const* marks) {
  int nh = m->nents[high];
  int deg = om_simplex_degree(high, low);
  int const* hl2l = om_ask_down(m, high, low, NULL);
  signed char* out = (signed char*)om_calloc((size_t)nh, 1);
  for (int h = 0; h < nh; ++h)
    for (int j = 0; j < deg; ++j)
      if (marks[hl2l[h * deg + j]]) out[h] = 1;
  return out;
}

void om_kat_mark_down(om_mesh* m, int high, int low, signed char const* marks,
    signed char* out) {
  signed char* r = mark_down(m, high, low, marks);
  memcpy(out, r, (size_t)m->nents[low]);
  free(r);
}

void om_kat_mark_up(om_mesh* m, int low, int high, signed char const* marks,
    signed char* out) {
  signed char* r = mark_up(m, low, high, marks);
  memcpy(out, r, (size_t)m->nents[high]);
  free(r);
}

static signed char* mark_sliver_layers(om_mesh* m, double qual_ceil, int nlayers) {
  int dim = m->dim;
  int ne = m->nents[dim];
  double const* q = om_ask_qualities(m);
  signed char* marks = (signed char*)om_malloc((size_t)ne);
  for (int i = 0; i < ne; ++i) marks[i] = q[i] < qual_ceil;
  om_graph const* dual = om_ask_dual(m);
  signed char* tmp = (signed char*)om_malloc((size_t)ne);
  for (int layer = 0; layer < nlayers; ++layer) {
    for (int e = 0; e < ne; ++e) {
      signed char r = -128; /* ArithTraits<I8>::min() */
      for (int j = dual->a2ab[e]; j < dual->a2ab[e + 1]; ++j) {
        signed char x = marks[dual->ab2b[j]];
        r = (r < x) ? x : r;
      }
      tmp[e] = r;
    }
    memcpy(marks, tmp, (size_t)ne);
  }
  free(tmp);
  return marks;
}

/* ------------------------------------------------------------ modify */
typedef struct {
  int nnew;
  int nsame;
  int* same2old;
  int* same2new;
  int nprods;
  int* prods2new;
  int* old2new;
} mod_out;

static void mod_out_free(mod_out* o) {
  free(o->same2old);
  free(o->same2new);
  free(o->prods2new);
  free(o->old2new);
  memset(o, 0, sizeof(*o));
}

static void modify_ents_adapt(om_mesh* old, om_mesh* nw, int ent_dim,
    int key_dim, int nkeys, int const* keys2kds, int const* keys2prods,
    int const* prod_verts, int const* old_lows2new_lows, mod_out* out) {
  int dim = old->dim;
  int nold = old->nents[ent_dim];
  signed char* mods[4] = {NULL, NULL, NULL, NULL};
  mods[key_dim] = (signed char*)om_calloc((size_t)old->nents[key_dim], 1);
  for (int k = 0; k < nkeys; ++k) mods[key_dim][keys2kds[k]] = 1;
  for (int md = key_dim + 1; md <= dim; ++md)
    mods[md] = mark_up(old, key_dim, md, mods[key_dim]);
  /* collect_same */
  if (mods[ent_dim]) {
    int c = 0;
    for (int i = 0; i < nold; ++i)
      if (!mods[ent_dim][i]) ++c;
    out->same2old = (int*)om_malloc(sizeof(int) * (size_t)c);
    out->nsame = c;
    c = 0;
    for (int i = 0; i < nold; ++i)
      if (!mods[ent_dim][i]) out->same2old[c++] = i;
  } else {
    out->nsame = nold;
    out->same2old = (int*)om_malloc(sizeof(int) * (size_t)nold);
    for (int i = 0; i < nold; ++i) out->same2old[i] = i;
  }
  /* get_mods2reps */
  int* reps = (int*)om_malloc(sizeof(int) * (size_t)nkeys);
  if (ent_dim == key_dim) {
    for (int k = 0; k < nkeys; ++k) reps[k] = keys2kds[k];
  } else if (ent_dim > key_dim) {
    om_graph const* up = om_ask_up(old, key_dim, ent_dim);
    for (int k = 0; k < nkeys; ++k) reps[k] = up->ab2b[up->a2ab[keys2kds[k]]];
  } else {
    int deg = om_simplex_degree(key_dim, ent_dim);
    int const* down = om_ask_down(old, key_dim, ent_dim, NULL);
    for (int k = 0; k < nkeys; ++k) reps[k] = down[keys2kds[k] * deg];
  }
  /* get_rep_counts + offset_scan */
  int* counts = (int*)om_calloc((size_t)nold, sizeof(int));
  for (int s = 0; s < out->nsame; ++s) counts[out->same2old[s]] = 1;
  for (int k = 0; k < nkeys; ++k) counts[reps[k]] += keys2prods[k + 1] - keys2prods[k];
  int* offsets = om_offset_scan(nold, counts);
  out->nnew = offsets[nold];
  /* get_rep2md_order (only when the keys have higher dimension) */
  int* order = NULL;
  if (key_dim > ent_dim) {
    int nkd = old->nents[key_dim];
    order = (int*)om_malloc(sizeof(int) * (size_t)nkd);
    int* md2mod = (int*)om_malloc(sizeof(int) * (size_t)nkd);
    for (int i = 0; i < nkd; ++i) order[i] = md2mod[i] = -1;
    for (int k = 0; k < nkeys; ++k) md2mod[keys2kds[k]] = k;
    om_graph const* up = om_ask_up(old, ent_dim, key_dim);
    for (int rep = 0; rep < nold; ++rep) {
      int offset = 0;
      for (int j = up->a2ab[rep]; j < up->a2ab[rep + 1]; ++j) {
        int md = up->ab2b[j];
        int mod = md2mod[md];
        if (mod >= 0 && om_code_which_down(up->codes[j]) == 0) {
          order[md] = offset;
          offset += keys2prods[mod + 1] - keys2prods[mod];
        }
      }
    }
    free(md2mod);
  }
  /* assign_new_numbering */
  out->same2new = (int*)om_malloc(sizeof(int) * (size_t)out->nsame);
  for (int s = 0; s < out->nsame; ++s) out->same2new[s] = offsets[out->same2old[s]];
  out->nprods = keys2prods[nkeys];
  out->prods2new = (int*)om_malloc(sizeof(int) * (size_t)out->nprods);
  for (int k = 0; k < nkeys; ++k) {
    int offset = offsets[reps[k]];
    if (order) offset += order[keys2kds[k]] + 1;
    for (int p = keys2prods[k]; p < keys2prods[k + 1]; ++p) out->prods2new[p] = offset++;
  }
  out->old2new = (int*)om_malloc(sizeof(int) * (size_t)nold);
  for (int i = 0; i < nold; ++i) out->old2new[i] = -1;
  for (int s = 0; s < out->nsame; ++s) out->old2new[out->same2old[s]] = out->same2new[s];
  /* modify_conn */
  if (ent_dim == 0) {
    nw->nents[0] = out->nnew;
  } else {
    int low_dim = ent_dim - 1;
    int deg = ent_dim + 1;
    int nnew = out->nnew;
    int* down = (int*)om_malloc(sizeof(int) * (size_t)nnew * deg);
    signed char* codes = NULL;
    if (low_dim > 0) codes = (signed char*)om_malloc((size_t)nnew * deg);
    int const* odown = old->down[ent_dim];
    signed char const* ocodes = old->down_codes[ent_dim];
    for (int s = 0; s < out->nsame; ++s)
      for (int j = 0; j < deg; ++j) {
        down[out->same2new[s] * deg + j] =
            old_lows2new_lows[odown[out->same2old[s] * deg + j]];
        if (codes) codes[out->same2new[s] * deg + j] = ocodes[out->same2old[s] * deg + j];
      }
    if (low_dim > 0) {
      int* pd = (int*)om_malloc(sizeof(int) * (size_t)out->nprods * deg);
      signed char* pc = (signed char*)om_malloc((size_t)out->nprods * deg);
      om_reflect_down(nw, ent_dim, out->nprods, prod_verts, pd, pc);
      for (int p = 0; p < out->nprods; ++p)
        for (int j = 0; j < deg; ++j) {
          down[out->prods2new[p] * deg + j] = pd[p * deg + j];
          codes[out->prods2new[p] * deg + j] = pc[p * deg + j];
        }
      free(pd);
      free(pc);
    } else {
      for (int p = 0; p < out->nprods; ++p)
        for (int j = 0; j < deg; ++j)
          down[out->prods2new[p] * deg + j] = prod_verts[p * deg + j];
    }
    nw->down[ent_dim] = down;
    nw->down_codes[ent_dim] = codes;
    nw->nents[ent_dim] = nnew;
  }
  for (int d = 0; d < 4; ++d) free(mods[d]);
  free(reps);
  free(counts);
  free(offsets);
  free(order);
}

/* transfer_common: same entities copy, products take prod_data */
static void transfer_i(int nnew, mod_out const* mo, int const* old_data,
    int const* prod_data, int** dst) {
  int* d = (int*)om_malloc(sizeof(int) * (size_t)nnew);
  for (int p = 0; p < mo->nprods; ++p) d[mo->prods2new[p]] = prod_data[p];
  for (int s = 0; s < mo->nsame; ++s) d[mo->same2new[s]] = old_data[mo->same2old[s]];
  *dst = d;
}
static void transfer_b(int nnew, mod_out const* mo, signed char const* old_data,
    signed char const* prod_data, signed char** dst) {
  signed char* d = (signed char*)om_malloc((size_t)nnew);
  for (int p = 0; p < mo->nprods; ++p) d[mo->prods2new[p]] = prod_data[p];
  for (int s = 0; s < mo->nsame; ++s) d[mo->same2new[s]] = old_data[mo->same2old[s]];
  *dst = d;
}
static void transfer_r(int nnew, int nc, mod_out const* mo, double const* old_data,
    double const* prod_data, double** dst) {
  double* d = (double*)om_malloc(sizeof(double) * (size_t)nnew * nc);
  for (int p = 0; p < mo->nprods; ++p)
    for (int c = 0; c < nc; ++c) d[mo->prods2new[p] * nc + c] = prod_data[p * nc + c];
  for (int s = 0; s < mo->nsame; ++s)
    for (int c = 0; c < nc; ++c) d[mo->same2new[s] * nc + c] = old_data[mo->same2old[s] * nc + c];
  *dst = d;
}

static int metric_dim_of(int ncomps) { return ncomps == 1 ? 1 : (ncomps == 3 ? 2 : 3); }

/* vertex tags of surviving vertices; with mo == NULL every vertex survives
 * in place (swap) */
static void copy_vtags(om_mesh const* m, om_mesh* nw, mod_out const* mo) {
  nw->nvtags = m->nvtags;
  for (int t = 0; t < m->nvtags; ++t) {
    om_vtag const* a = &m->vtags[t];
    om_vtag* b = &nw->vtags[t];
    *b = *a;
    b->data = NULL;
    b->err = NULL;
    if (mo) {
      transfer_r(mo->nnew, a->ncomps, mo, a->data, NULL, &b->data);
    } else {
      size_t n = (size_t)m->nents[0] * a->ncomps;
      b->data = (double*)om_malloc(sizeof(double) * (n + 1));
      memcpy(b->data, a->data, sizeof(double) * n);
    }
  }
}

/* vertex tags at refinement midpoints */
static void refine_vtags(om_mesh const* m, om_mesh* nw, int nkeys, int const* keys2edges,
    mod_out const* mo) {
  int const* ev2v = m->down[1];
  nw->nvtags = m->nvtags;
  for (int t = 0; t < m->nvtags; ++t) {
    om_vtag const* a = &m->vtags[t];
    om_vtag* b = &nw->vtags[t];
    *b = *a;
    b->data = NULL;
    b->err = NULL;
    int nc = a->ncomps;
    double* pd = (double*)om_malloc(sizeof(double) * (size_t)(nkeys + 1) * nc);
    for (int k = 0; k < nkeys; ++k) {
      int e = keys2edges[k];
      double const* x = a->data + (size_t)ev2v[e * 2] * nc;
      double const* y = a->data + (size_t)ev2v[e * 2 + 1] * nc;
      if (a->kind == OM_METRIC) {
        double const* ms[2] = {x, y};
        om_average_metric(metric_dim_of(nc), 2, ms, pd + (size_t)k * nc);
      } else {
        for (int j = 0; j < nc; ++j) {
          double comp = 0;
          comp += x[j];
          comp += y[j];
          comp /= 2;
          pd[(size_t)k * nc + j] = comp;
        }
      }
    }
    transfer_r(mo->nnew, nc, mo, a->data, pd, &b->data);
    free(pd);
  }
}

/* inheritance tags of dimension d: products take the value of entity
 * prod_src[p] of dimension src_dims[p] (or src_dim when src_dims is NULL);
 * prod_src == NULL means there are no products */
static void inherit_itags(om_mesh const* old, om_mesh* nw, int d, mod_out const* mo,
    int src_dim, int const* src_dims, int const* prod_src) {
  nw->nitags = old->nitags;
  for (int t = 0; t < old->nitags; ++t) {
    om_itag const* a = &old->itags[t];
    om_itag* b = &nw->itags[t];
    memcpy(b->name, a->name, sizeof(b->name));
    b->ncomps = a->ncomps;
    int nc = a->ncomps;
    double* pd = NULL;
    if (prod_src) {
      pd = (double*)om_malloc(sizeof(double) * (size_t)(mo->nprods + 1) * nc);
      for (int p = 0; p < mo->nprods; ++p) {
        int sd = src_dims ? src_dims[p] : src_dim;
        for (int k = 0; k < nc; ++k)
          pd[(size_t)p * nc + k] = a->data[sd][(size_t)prod_src[p] * nc + k];
      }
    }
    transfer_r(mo->nnew, nc, mo, a->data[d], pd, &b->data[d]);
    free(pd);
  }
}

/* element tags: products take the value of old element prod_src[p];
 * prod_src == NULL means there are no products */
static void etags_from(om_mesh const* old, om_mesh* nw, mod_out const* mo, int const* prod_src) {
  nw->netags = old->netags;
  for (int t = 0; t < old->netags; ++t) {
    om_etag const* a = &old->etags[t];
    om_etag* b = &nw->etags[t];
    *b = *a;
    b->data = NULL;
    b->err = NULL;
    b->old = NULL;
    int nc = a->ncomps;
    double* pd = NULL;
    if (prod_src) {
      pd = (double*)om_malloc(sizeof(double) * (size_t)(mo->nprods + 1) * nc);
      for (int p = 0; p < mo->nprods; ++p)
        for (int k = 0; k < nc; ++k) pd[(size_t)p * nc + k] = a->data[(size_t)prod_src[p] * nc + k];
    }
    transfer_r(mo->nnew, nc, mo, a->data, pd, &b->data);
    free(pd);
  }
}

/* transfer_pointwise_tmpl (src/Omega_h_transfer.cpp:515-563): each product
 * element takes the value of the old element around its key (a vertex or an
 * edge, rows in the order of the upward adjacency) in which the product's
 * centroid has the largest smallest barycentric coordinate, the first one on
 * ties; pv holds the products' vertices in the new mesh, whose coordinates
 * are already set */
static void etags_pointwise(om_mesh* old, om_mesh* nw, mod_out const* mo, int nkeys,
    int const* keys2kds, int key_dim, int const* k2p, int const* pv) {
  if (old->netags == 0) {
    nw->netags = 0;
    return;
  }
  int dim = old->dim;
  int nprods = k2p[nkeys];
  int* best = (int*)om_malloc(sizeof(int) * (size_t)(nprods + 1));
  om_graph const* kd2e = om_ask_up(old, key_dim, dim);
  int const* ev2v = om_verts_of(old, dim);
  for (int k = 0; k < nkeys; ++k) {
    int kd = keys2kds[k];
    for (int p = k2p[k]; p < k2p[k + 1]; ++p) {
      double x[3];
      double const* q0 = nw->coords + (size_t)pv[p * (dim + 1)] * dim;
      for (int j = 0; j < dim; ++j) x[j] = q0[j];
      for (int i = 1; i <= dim; ++i) {
        double const* qi = nw->coords + (size_t)pv[p * (dim + 1) + i] * dim;
        for (int j = 0; j < dim; ++j) x[j] = x[j] + qi[j];
      }
      for (int j = 0; j < dim; ++j) x[j] = x[j] / (double)(dim + 1);
      int b = -1;
      double bd = 0.0;
      for (int ke = kd2e->a2ab[kd]; ke < kd2e->a2ab[kd + 1]; ++ke) {
        int e = kd2e->ab2b[ke];
        double const* pp[4];
        for (int i = 0; i <= dim; ++i) pp[i] = old->coords + (size_t)ev2v[e * (dim + 1) + i] * dim;
        double xi[4];
        om_barycentric(dim, x, pp, xi);
        double mn = xi[0];
        for (int i = 1; i <= dim; ++i) mn = (mn < xi[i]) ? mn : xi[i];
        double dist = -mn;
        if (b == -1 || dist < bd) {
          b = e;
          bd = dist;
        }
      }
      best[p] = b;
    }
  }
  etags_from(old, nw, mo, best);
  free(best);
}

/* class_id/class_dim of products from a list of old entities of dimension
 * src_dim (one per product) */
static void transfer_class_from(om_mesh* old, om_mesh* nw, int d, int src_dim,
    int const* prod_src, mod_out const* mo) {
  int* pid = (int*)om_malloc(sizeof(int) * (size_t)(mo->nprods + 1));
  signed char* pdim = (signed char*)om_malloc((size_t)(mo->nprods + 1));
  for (int p = 0; p < mo->nprods; ++p) {
    pid[p] = old->class_id[src_dim][prod_src[p]];
    pdim[p] = old->class_dim[src_dim][prod_src[p]];
  }
  transfer_i(mo->nnew, mo, old->class_id[d], pid, &nw->class_id[d]);
  transfer_b(mo->nnew, mo, old->class_dim[d], pdim, &nw->class_dim[d]);
  free(pid);
  free(pdim);
}

static void transfer_lengths_qualities(om_mesh* old, om_mesh* nw, int d,
    mod_out const* mo) {
  if (d == 1) {
    double* pd = (double*)om_malloc(sizeof(double) * (size_t)(mo->nprods + 1));
    om_measure_edges(nw, mo->nprods, mo->prods2new, pd);
    transfer_r(mo->nnew, 1, mo, om_ask_lengths(old), pd, &nw->length);
    free(pd);
  }
  if (d == old->dim) {
    double* pd = (double*)om_malloc(sizeof(double) * (size_t)(mo->nprods + 1));
    om_measure_elems(nw, mo->nprods, mo->prods2new, pd);
    transfer_r(mo->nnew, 1, mo, om_ask_qualities(old), pd, &nw->quality);
    free(pd);
  }
}

static void replace_mesh(om_mesh* m, om_mesh* nw) {
  om_mesh tmp = *m;
  *m = *nw;
  *nw = tmp;
  om_mesh_free(nw);
}

/* ------------------------------------------------------------ mident */
static void mident_metrics(om_mesh* m, int n, int const* edges, double* out) {
  int nc = m->metric_ncomps;
  int mdim = (nc == 1) ? 1 : m->dim;
  int const* ev2v = m->down[1];
  for (int i = 0; i < n; ++i) {
    int e = edges[i];
    double const* ms[2] = {m->metric + ev2v[e * 2] * nc, m->metric + ev2v[e * 2 + 1] * nc};
    om_average_metric(mdim, 2, ms, out + i * nc);
  }
}

void om_kat_mident_metrics(om_mesh* m, int n, int const* edges, double* out) {
  mident_metrics(m, n, edges, out);
}

/* ------------------------------------------------------------ refine */
static void refine_qualities(om_mesh* m, int ncands, int const* cands,
    double const* midpt_metrics, double* out) {
  int dim = m->dim;
  int nc = m->metric_ncomps;
  int mdim = (nc == 1) ? 1 : dim;
  int const* ev2v = m->down[1];
  int const* cv2v = om_verts_of(m, dim);
  om_graph const* e2c = om_ask_up(m, 1, dim);
  for (int cand = 0; cand < ncands; ++cand) {
    int e = cands[cand];
    double midp[3];
    for (int j = 0; j < dim; ++j)
      midp[j] = (m->coords[ev2v[e * 2] * dim + j] + m->coords[ev2v[e * 2 + 1] * dim + j]) / 2.;
    double minqual = 1.0;
    for (int ec = e2c->a2ab[e]; ec < e2c->a2ab[e + 1]; ++ec) {
      int c = e2c->ab2b[ec];
      signed char code = e2c->codes[ec];
      int cce = om_code_which_down(code);
      int rot = om_code_rotation(code);
      int const* ccv2v = cv2v + c * (dim + 1);
      for (int eev = 0; eev < 2; ++eev) {
        int cev = eev ^ rot;
        int ccv = om_down_template(dim, 1, cce, cev);
        int ccs = om_opposite_template(dim, 0, ccv);
        int csv2v[3];
        double const* p[4];
        double const* ms[4];
        for (int csv = 0; csv < dim; ++csv) {
          int ccv2 = om_down_template(dim, dim - 1, ccs, csv);
          csv2v[csv] = ccv2v[ccv2];
        }
        for (int csv = 0; csv < dim; ++csv) {
          p[csv] = m->coords + csv2v[csv] * dim;
          ms[csv] = m->metric + csv2v[csv] * nc;
        }
        p[dim] = midp;
        ms[dim] = midpt_metrics + cand * nc;
        if (dim == 3) {
          double const* t = p[1];
          p[1] = p[2];
          p[2] = t;
          t = ms[1];
          ms[1] = ms[2];
          ms[2] = t;
        }
        double q = om_element_quality(dim, mdim, p, ms);
        minqual = min2(minqual, q);
      }
    }
    out[cand] = minqual;
  }
}

void om_kat_refine_qualities(om_mesh* m, int n, int const* cands, double* out) {
  double* mm = (double*)om_malloc(sizeof(double) * (size_t)n * m->metric_ncomps);
  mident_metrics(m, n, cands, mm);
  refine_qualities(m, n, cands, mm, out);
  free(mm);
}

/* refine_products for one dimension, src/Omega_h_refine_topology.cpp:48-203 */
static void refine_products(om_mesh* m, int ent_dim, int nkeys,
    int const* keys2edges, int const* keys2midverts, int const* o2n,
    int** keys2prods_out, int** prod_verts_out, int** prod_src_out,
    int* prod_src_dim_pairs) {
  int dim = m->dim;
  int const* ev2v = m->down[1];
  int nv = ent_dim + 1;
  int* npairs = (int*)om_malloc(sizeof(int) * (size_t)nkeys);
  int* ncuts = (int*)om_calloc((size_t)nkeys, sizeof(int));
  om_graph const* pairs_up = (ent_dim > 1) ? om_ask_up(m, 1, ent_dim) : NULL;
  om_graph const* cuts_up = (ent_dim < dim) ? om_ask_up(m, 1, ent_dim + 1) : NULL;
  for (int k = 0; k < nkeys; ++k) {
    int e = keys2edges[k];
    npairs[k] = pairs_up ? 2 * (pairs_up->a2ab[e + 1] - pairs_up->a2ab[e]) : 2;
    if (cuts_up) ncuts[k] = cuts_up->a2ab[e + 1] - cuts_up->a2ab[e];
  }
  int* tot = (int*)om_malloc(sizeof(int) * (size_t)nkeys);
  for (int k = 0; k < nkeys; ++k) tot[k] = npairs[k] + ncuts[k];
  int* k2p = om_offset_scan(nkeys, tot);
  int nprods = k2p[nkeys];
  int* pv = (int*)om_malloc(sizeof(int) * (size_t)(nprods + 1) * nv);
  int* psrc = (int*)om_malloc(sizeof(int) * (size_t)(nprods + 1));
  for (int k = 0; k < nkeys; ++k) {
    int e = keys2edges[k];
    int mid = keys2midverts[k];
    int prod = k2p[k];
    if (ent_dim == 1) {
      pv[prod * 2 + 0] = o2n[ev2v[e * 2 + 0]];
      pv[prod * 2 + 1] = mid;
      psrc[prod] = e;
      ++prod;
      pv[prod * 2 + 0] = mid;
      pv[prod * 2 + 1] = o2n[ev2v[e * 2 + 1]];
      psrc[prod] = e;
      ++prod;
    } else {
      int const* dv2v = om_verts_of(m, ent_dim);
      for (int ed = pairs_up->a2ab[e]; ed < pairs_up->a2ab[e + 1]; ++ed) {
        int dom = pairs_up->ab2b[ed];
        signed char code = pairs_up->codes[ed];
        int dde = om_code_which_down(code);
        int rot = om_code_rotation(code);
        for (int eev = 0; eev < 2; ++eev) {
          int dev = eev ^ rot;
          int ddv = om_down_template(ent_dim, 1, dde, dev);
          int dds = om_opposite_template(ent_dim, 0, ddv);
          int* ppv = pv + prod * nv;
          for (int dsv = 0; dsv < ent_dim; ++dsv) {
            int ddv2 = om_down_template(ent_dim, ent_dim - 1, dds, dsv);
            ppv[dsv] = o2n[dv2v[dom * (ent_dim + 1) + ddv2]];
          }
          ppv[ent_dim] = mid;
          if (ent_dim == 3) {
            int t = ppv[1];
            ppv[1] = ppv[2];
            ppv[2] = t;
          }
          psrc[prod] = dom;
          ++prod;
        }
      }
    }
    if (cuts_up) {
      int cdim = ent_dim + 1;
      int const* dv2v = om_verts_of(m, cdim);
      for (int ed = cuts_up->a2ab[e]; ed < cuts_up->a2ab[e + 1]; ++ed) {
        int dom = cuts_up->ab2b[ed];
        int dde = om_code_which_down(cuts_up->codes[ed]);
        int ddt = om_opposite_template(cdim, 1, dde);
        int* ccv = pv + prod * nv;
        for (int dtv = 0; dtv < cdim - 1; ++dtv) {
          int ddv2 = om_down_template(cdim, cdim - 2, ddt, dtv);
          ccv[dtv] = o2n[dv2v[dom * (cdim + 1) + ddv2]];
        }
        ccv[cdim - 1] = mid;
        psrc[prod] = dom;
        ++prod;
      }
    }
  }
  *prod_src_dim_pairs = ent_dim;
  *keys2prods_out = k2p;
  *prod_verts_out = pv;
  *prod_src_out = psrc;
  free(npairs);
  free(ncuts);
  free(tot);
}

int om_refine_by_size(om_mesh* m, om_adapt_opts const* o) {
  int dim = m->dim;
  int ne = m->nents[1];
  double const* lengths = om_ask_lengths(m);
  om_ask_qualities(m);
  signed char* cand = (signed char*)om_malloc((size_t)ne);
  for (int e = 0; e < ne; ++e) cand[e] = lengths[e] > o->max_length_desired;
  if (!any_marked(ne, cand)) {
    free(cand);
    return 0;
  }
  int ncands;
  int* cands = collect_marked(ne, cand, &ncands);
  free(cand);
  int nc = m->metric_ncomps;
  double* mm = (double*)om_malloc(sizeof(double) * (size_t)ncands * nc);
  mident_metrics(m, ncands, cands, mm);
  double* quals = (double*)om_malloc(sizeof(double) * (size_t)ncands);
  refine_qualities(m, ncands, cands, mm, quals);
  free(mm);
  signed char* initial = (signed char*)om_calloc((size_t)ne, 1);
  double* edge_quals = (double*)om_calloc((size_t)ne, sizeof(double));
  int anygood = 0;
  for (int c = 0; c < ncands; ++c) {
    int good = quals[c] >= o->min_quality_allowed;
    initial[cands[c]] = (signed char)good;
    edge_quals[cands[c]] = quals[c];
    anygood |= good;
  }
  free(cands);
  free(quals);
  if (!anygood) {
    free(initial);
    free(edge_quals);
    return 0;
  }
  signed char* keys_mark = (signed char*)om_malloc((size_t)ne);
  om_find_indset(m, 1, edge_quals, initial, keys_mark);
  free(initial);
  free(edge_quals);
  int nkeys;
  int* keys2edges = collect_marked(ne, keys_mark, &nkeys);
  free(keys_mark);
  /* refine_element_based */
  om_mesh* nw = om_mesh_new(dim);
  nw->metric_ncomps = nc;
  int* keys2midverts = NULL;
  int* o2n_verts = NULL;
  int* old_lows2new_lows = NULL;
  for (int d = 0; d <= dim; ++d) {
    int* k2p = NULL;
    int* pv = NULL;
    int* psrc = NULL;
    int dummy;
    if (d == 0) {
      k2p = (int*)om_malloc(sizeof(int) * (size_t)(nkeys + 1));
      for (int k = 0; k <= nkeys; ++k) k2p[k] = k;
    } else {
      refine_products(m, d, nkeys, keys2edges, keys2midverts, o2n_verts, &k2p,
          &pv, &psrc, &dummy);
    }
    mod_out mo;
    memset(&mo, 0, sizeof(mo));
    modify_ents_adapt(m, nw, d, 1, nkeys, keys2edges, k2p, pv, old_lows2new_lows, &mo);
    if (d == 0) {
      keys2midverts = (int*)om_malloc(sizeof(int) * (size_t)(nkeys + 1));
      memcpy(keys2midverts, mo.prods2new, sizeof(int) * (size_t)nkeys);
      o2n_verts = (int*)om_malloc(sizeof(int) * (size_t)m->nents[0]);
      memcpy(o2n_verts, mo.old2new, sizeof(int) * (size_t)m->nents[0]);
      /* coordinates (average_field) and metric (get_mident_metrics) */
      double* pc = (double*)om_malloc(sizeof(double) * (size_t)(nkeys + 1) * dim);
      int const* ev2v = m->down[1];
      for (int k = 0; k < nkeys; ++k) {
        int e = keys2edges[k];
        for (int j = 0; j < dim; ++j) {
          double comp = 0;
          comp += m->coords[ev2v[e * 2] * dim + j];
          comp += m->coords[ev2v[e * 2 + 1] * dim + j];
          comp /= 2;
          pc[k * dim + j] = comp;
        }
      }
      transfer_r(mo.nnew, dim, &mo, m->coords, pc, &nw->coords);
      free(pc);
      double* pm = (double*)om_malloc(sizeof(double) * (size_t)(nkeys + 1) * nc);
      mident_metrics(m, nkeys, keys2edges, pm);
      transfer_r(mo.nnew, nc, &mo, m->metric, pm, &nw->metric);
      free(pm);
      refine_vtags(m, nw, nkeys, keys2edges, &mo);
      /* the midpoint inherits the key edge's classification */
      transfer_class_from(m, nw, 0, 1, keys2edges, &mo);
      inherit_itags(m, nw, 0, &mo, 1, NULL, keys2edges);
    } else {
      /* pairs inherit from their domain (same dim), cuts from dim+1 */
      int nprods = k2p[nkeys];
      int* pid = (int*)om_malloc(sizeof(int) * (size_t)(nprods + 1));
      signed char* pdim = (signed char*)om_malloc((size_t)(nprods + 1));
      int* psd = (int*)om_malloc(sizeof(int) * (size_t)(nprods + 1));
      om_graph const* cuts_up = (d < dim) ? om_ask_up(m, 1, d + 1) : NULL;
      for (int k = 0; k < nkeys; ++k) {
        int ncut = cuts_up ? (cuts_up->a2ab[keys2edges[k] + 1] - cuts_up->a2ab[keys2edges[k]]) : 0;
        for (int p = k2p[k]; p < k2p[k + 1]; ++p) {
          int src_dim = (p >= k2p[k + 1] - ncut) ? d + 1 : d;
          psd[p] = src_dim;
          pid[p] = m->class_id[src_dim][psrc[p]];
          pdim[p] = m->class_dim[src_dim][psrc[p]];
        }
      }
      transfer_i(mo.nnew, &mo, m->class_id[d], pid, &nw->class_id[d]);
      transfer_b(mo.nnew, &mo, m->class_dim[d], pdim, &nw->class_dim[d]);
      inherit_itags(m, nw, d, &mo, d, psd, psrc);
      if (d == dim) {
        etags_from(m, nw, &mo, psrc);
        om_conserve_pass(m, nw, 0, 1, nkeys, keys2edges, k2p, mo.prods2new, mo.nsame,
            mo.same2old, mo.same2new);
      }
      free(pid);
      free(pdim);
      free(psd);
      transfer_lengths_qualities(m, nw, d, &mo);
    }
    free(old_lows2new_lows);
    old_lows2new_lows = (int*)om_malloc(sizeof(int) * (size_t)(m->nents[d] + 1));
    memcpy(old_lows2new_lows, mo.old2new, sizeof(int) * (size_t)m->nents[d]);
    mod_out_free(&mo);
    free(k2p);
    free(pv);
    free(psrc);
  }
  free(old_lows2new_lows);
  free(keys2midverts);
  free(o2n_verts);
  free(keys2edges);
  replace_mesh(m, nw);
  return 1;
}

/* ------------------------------------------------------------ coarsen */
static int collapses(signed char code, int col_v) { return (code & (1 << col_v)) != 0; }
static signed char dont_collapse(signed char code, int col_v) {
  return (signed char)(code & ~(1 << col_v));
}

/* coarsen_qualities for one edge, src/Omega_h_coarsen_qualities.cpp:16-63:
 * per collapse direction the minimum quality of the collapsing vertex's
 * elements that survive, with that vertex moved onto the other end; -1 for
 * a direction that does not collapse */
static void coarsen_edge_qualities(om_mesh* m, int e, signed char code, double* out) {
  int dim = m->dim;
  int nc = m->metric_ncomps;
  int mdim = (nc == 1) ? 1 : dim;
  int const* ev2v = m->down[1];
  int const* cv2v = om_verts_of(m, dim);
  om_graph const* v2c = om_ask_up(m, 0, dim);
  out[0] = out[1] = -1.0;
  if (code == DONT_COLLAPSE) return;
  for (int col = 0; col < 2; ++col) {
    if (!collapses(code, col)) continue;
    int v_col = ev2v[e * 2 + col];
    int v_onto = ev2v[e * 2 + (1 - col)];
    double minqual = 1.0;
    for (int vc = v2c->a2ab[v_col]; vc < v2c->a2ab[v_col + 1]; ++vc) {
      int c = v2c->ab2b[vc];
      int ccv_col = om_code_which_down(v2c->codes[vc]);
      int cv[4];
      int will_die = 0;
      for (int k = 0; k < dim + 1; ++k) {
        cv[k] = cv2v[c * (dim + 1) + k];
        if (k != ccv_col && cv[k] == v_onto) will_die = 1;
      }
      if (will_die) continue;
      cv[ccv_col] = v_onto;
      double const* p[4];
      double const* ms[4];
      for (int k = 0; k < dim + 1; ++k) {
        p[k] = m->coords + cv[k] * dim;
        ms[k] = m->metric + cv[k] * nc;
      }
      double q = om_element_quality(dim, mdim, p, ms);
      minqual = min2(minqual, q);
    }
    out[col] = minqual;
  }
}

void om_kat_coarsen_qualities(om_mesh* m, int n, int const* cands, signed char const* codes,
    double* out) {
  for (int i = 0; i < n; ++i) coarsen_edge_qualities(m, cands[i], codes[i], out + 2 * i);
}

/* prevent_coarsen_flip, src/Omega_h_coarsen_flip.cpp:68-150.  Exposed sides
 * (fewer than two elements, mark_exposed_sides src/Omega_h_mark.cpp:12-19)
 * that touch a vertex of a candidate edge "matter", and so do their
 * vertices.  Each vertex that matters gets a normal
 * (compute_flip_normals_dim, :14-64): the normalized sum of its exposed
 * sides' area vectors, replaced by the most normal normal of their unit
 * normals when one of those is more than acos(0.95) away from it.  A
 * collapse direction is vetoed when a side that matters around the
 * collapsing vertex, and does not hold the vertex it collapses onto, would
 * after the collapse have a normal whose dot product with the normal of one
 * of its vertices falls below 1e-10 (:70-121).  Returns 1 where the
 * reference's checks would fail. */
static int prevent_coarsen_flip(om_mesh* m, signed char* edge_codes) {
  int dim = m->dim;
  int sd = dim - 1;
  int nv = m->nents[0], ne = m->nents[1], ns = m->nents[sd];
  int const* ev2v = m->down[1];
  int const* sv2v = om_verts_of(m, sd);
  om_graph const* v2s = om_ask_up(m, 0, sd);
  om_graph const* s2e = om_ask_up(m, sd, dim);
  signed char* vcand = (signed char*)om_calloc((size_t)nv, 1);
  for (int e = 0; e < ne; ++e) {
    if (edge_codes[e] == DONT_COLLAPSE) continue;
    vcand[ev2v[e * 2]] = 1;
    vcand[ev2v[e * 2 + 1]] = 1;
  }
  signed char* exposed = (signed char*)om_calloc((size_t)ns, 1);
  signed char* smatter = (signed char*)om_calloc((size_t)ns, 1);
  signed char* vmatter = (signed char*)om_calloc((size_t)nv, 1);
  for (int s = 0; s < ns; ++s) {
    exposed[s] = (s2e->a2ab[s + 1] - s2e->a2ab[s]) < 2;
    int adj = 0;
    for (int k = 0; k < sd + 1; ++k) adj |= vcand[sv2v[s * (sd + 1) + k]];
    smatter[s] = (adj && exposed[s]) ? 1 : 0;
    if (smatter[s])
      for (int k = 0; k < sd + 1; ++k) vmatter[sv2v[s * (sd + 1) + k]] = 1;
  }
  int const max_adj = (dim == 3) ? 24 : 8;
  double* normals = (double*)om_calloc((size_t)nv * dim, sizeof(double));
  double N[24 * 3];
  int bad = 0;
  for (int v = 0; v < nv && !bad; ++v) {
    if (!vmatter[v]) continue;
    double Nc[3] = {0.0, 0.0, 0.0};
    int n = 0;
    for (int vs = v2s->a2ab[v]; vs < v2s->a2ab[v + 1]; ++vs) {
      int s = v2s->ab2b[vs];
      if (!exposed[s]) continue;
      if (n >= max_adj) {
        bad = 1;
        break;
      }
      double const* p[3];
      for (int k = 0; k < sd + 1; ++k) p[k] = m->coords + (size_t)sv2v[s * (sd + 1) + k] * dim;
      double sv[3];
      om_side_vector(dim, p, sv);
      double l = 0.0;
      for (int k = 0; k < dim; ++k) Nc[k] = Nc[k] + sv[k];
      l = sv[0] * sv[0];
      for (int k = 1; k < dim; ++k) l = l + sv[k] * sv[k];
      l = om_sqrt(l);
      for (int k = 0; k < dim; ++k) N[n * dim + k] = sv[k] / l;
      ++n;
    }
    if (bad) break;
    double l = Nc[0] * Nc[0];
    for (int k = 1; k < dim; ++k) l = l + Nc[k] * Nc[k];
    l = om_sqrt(l);
    for (int k = 0; k < dim; ++k) Nc[k] = Nc[k] / l;
    int i;
    for (i = 0; i < n; ++i) {
      double d = Nc[0] * N[i * dim];
      for (int k = 1; k < dim; ++k) d = d + Nc[k] * N[i * dim + k];
      if (d < 0.95) break;
    }
    if (i < n && om_most_normal(dim, n, N, Nc)) {
      bad = 1;
      break;
    }
    for (int k = 0; k < dim; ++k) normals[(size_t)v * dim + k] = Nc[k];
  }
  for (int e = 0; e < ne && !bad; ++e) {
    signed char code = edge_codes[e];
    if (code == DONT_COLLAPSE) continue;
    for (int col = 0; col < 2; ++col) {
      if (!collapses(code, col)) continue;
      int v_col = ev2v[e * 2 + col];
      if (!vmatter[v_col]) continue;
      int v_onto = ev2v[e * 2 + (1 - col)];
      for (int vs = v2s->a2ab[v_col]; vs < v2s->a2ab[v_col + 1]; ++vs) {
        int s = v2s->ab2b[vs];
        if (!smatter[s]) continue;
        int ssv_col = om_code_which_down(v2s->codes[vs]);
        int sv[3];
        int k;
        for (k = 0; k < sd + 1; ++k) sv[k] = sv2v[s * (sd + 1) + k];
        for (k = 0; k < sd + 1; ++k)
          if (sv[k] == v_onto) break;
        if (k != sd + 1) continue;
        sv[ssv_col] = v_onto;
        double const* p[3];
        for (k = 0; k < sd + 1; ++k) p[k] = m->coords + (size_t)sv[k] * dim;
        double sn[3];
        om_side_vector(dim, p, sn);
        for (k = 0; k < sd + 1; ++k) {
          double const* vn = normals + (size_t)sv[k] * dim;
          double d = vn[0] * sn[0];
          for (int q = 1; q < dim; ++q) d = d + vn[q] * sn[q];
          if (d < 1e-10) break;
        }
        if (k < sd + 1) code = dont_collapse(code, col);
      }
    }
    edge_codes[e] = code;
  }
  free(vcand);
  free(exposed);
  free(smatter);
  free(vmatter);
  free(normals);
  return bad;
}

static int coarsen(om_mesh* m, om_adapt_opts const* o, signed char* edge_codes,
    int overshoot_allowed, int improve) {
  int dim = m->dim;
  int ne = m->nents[1];
  int nc = m->metric_ncomps;
  int mdim = (nc == 1) ? 1 : dim;
  int const* ev2v = m->down[1];
  /* coarsen_element_based1: check_collapse_class */
  int any = 0;
  for (int e = 0; e < ne; ++e) {
    signed char code = edge_codes[e];
    if (code == DONT_COLLAPSE) continue;
    int cd0 = m->class_dim[0][ev2v[e * 2]], cd1 = m->class_dim[0][ev2v[e * 2 + 1]];
    int ecd = m->class_dim[1][e];
    signed char orig = code;
    if (cd0 == cd1) {
      if (cd0 != ecd) code = DONT_COLLAPSE;
    } else {
      int cds[2] = {cd0, cd1};
      for (int col = 0; col < 2; ++col)
        if (collapses(orig, col) && cds[col] != ecd) code = dont_collapse(code, col);
    }
    edge_codes[e] = code;
    if (code > any) any = code;
  }
  if (any <= DONT_COLLAPSE) return 0;
  /* coarsen_ghosted: check_collapse_exposure */
  for (int cell_dim = 2; cell_dim <= dim; ++cell_dim) {
    om_graph const* e2c = om_ask_up(m, 1, cell_dim);
    int const* cs2s = m->down[cell_dim];
    int nccs = cell_dim + 1;
    for (int e = 0; e < ne; ++e) {
      signed char code = edge_codes[e];
      if (code == DONT_COLLAPSE) continue;
      for (int ec = e2c->a2ab[e]; ec < e2c->a2ab[e + 1]; ++ec) {
        int c = e2c->ab2b[ec];
        int cce = om_code_which_down(e2c->codes[ec]);
        int rot = om_code_rotation(e2c->codes[ec]);
        int c_dim = m->class_dim[cell_dim][c];
        for (int col = 0; col < 2; ++col) {
          if (!collapses(code, col)) continue;
          int onto = 1 - col;
          int cev_onto = rot ^ onto;
          int ccv_onto = om_down_template(cell_dim, 1, cce, cev_onto);
          int ccs_opp = om_opposite_template(cell_dim, 0, ccv_onto);
          int s_opp = cs2s[c * nccs + ccs_opp];
          if (m->class_dim[cell_dim - 1][s_opp] != c_dim) code = dont_collapse(code, col);
        }
      }
      edge_codes[e] = code;
    }
  }
  /* prevent_coarsen_overshoot */
  double max_length = overshoot_allowed ? o->max_length_allowed : o->max_length_desired;
  om_graph const* v2e = om_ask_up(m, 0, 1);
  int anyc = 0;
  for (int e = 0; e < ne; ++e) {
    signed char code = edge_codes[e];
    if (code == DONT_COLLAPSE) continue;
    for (int col = 0; col < 2; ++col) {
      if (!collapses(code, col)) continue;
      int v_col = ev2v[e * 2 + col];
      int v_onto = ev2v[e * 2 + (1 - col)];
      for (int ve = v2e->a2ab[v_col]; ve < v2e->a2ab[v_col + 1]; ++ve) {
        int e2 = v2e->ab2b[ve];
        if (e2 == e) continue;
        int eev_in = om_code_which_down(v2e->codes[ve]);
        int eev_out = 1 - eev_in;
        int nv[2];
        nv[eev_in] = v_onto;
        nv[eev_out] = ev2v[e2 * 2 + eev_out];
        double len = om_metric_edge_length(dim, mdim, m->coords + nv[0] * dim,
            m->coords + nv[1] * dim, m->metric + nv[0] * nc, m->metric + nv[1] * nc);
        if (len >= max_length) {
          code = dont_collapse(code, col);
          break;
        }
      }
    }
    edge_codes[e] = code;
    if (code != DONT_COLLAPSE) anyc = 1;
  }
  if (!anyc) return 0;
  if (o->should_prevent_coarsen_flip) {
    if (prevent_coarsen_flip(m, edge_codes)) {
      fprintf(stderr, "oracle: prevent_coarsen_flip: a check of the reference failed\n");
      abort();
    }
    anyc = 0;
    for (int e = 0; e < ne; ++e)
      if (edge_codes[e] != DONT_COLLAPSE) anyc = 1;
    if (!anyc) return 0;
  }
  /* coarsen_qualities */
  double* cq = (double*)om_malloc(sizeof(double) * (size_t)ne * 2);
  om_graph const* v2c = om_ask_up(m, 0, dim);
  for (int e = 0; e < ne; ++e) coarsen_edge_qualities(m, e, edge_codes[e], cq + e * 2);
  /* filter_coarsen_min_qual and filter_coarsen_improve */
  double const* eq = om_ask_qualities(m);
  double* vold = NULL;
  if (improve) {
    int nv = m->nents[0];
    vold = (double*)om_malloc(sizeof(double) * (size_t)nv);
    for (int v = 0; v < nv; ++v) {
      double r = 1.79769313486231570815e+308;
      for (int vc = v2c->a2ab[v]; vc < v2c->a2ab[v + 1]; ++vc) r = min2(r, eq[v2c->ab2b[vc]]);
      vold[v] = r;
    }
  }
  anyc = 0;
  for (int e = 0; e < ne; ++e) {
    signed char code = edge_codes[e];
    if (code == DONT_COLLAPSE) continue;
    for (int col = 0; col < 2; ++col) {
      if (!(cq[e * 2 + col] >= o->min_quality_allowed)) code = dont_collapse(code, col);
    }
    if (improve) {
      for (int col = 0; col < 2; ++col)
        if (!(cq[e * 2 + col] > vold[ev2v[e * 2 + col]])) code = dont_collapse(code, col);
    }
    edge_codes[e] = code;
    if (code != DONT_COLLAPSE) anyc = 1;
  }
  free(vold);
  if (!anyc) {
    free(cq);
    return 0;
  }
  /* choose_rails */
  int nv = m->nents[0];
  double const* lengths = om_ask_lengths(m);
  signed char* vcand = (signed char*)om_calloc((size_t)nv, 1);
  double* vqual = (double*)om_malloc(sizeof(double) * (size_t)nv);
  int* vrail = (int*)om_malloc(sizeof(int) * (size_t)nv);
  for (int v = 0; v < nv; ++v) {
    int best = -1;
    double best_len = -1, best_q = -1;
    for (int ve = v2e->a2ab[v]; ve < v2e->a2ab[v + 1]; ++ve) {
      int e = v2e->ab2b[ve];
      signed char code = edge_codes[e];
      if (code == DONT_COLLAPSE) continue;
      int eev = om_code_which_down(v2e->codes[ve]);
      if (!collapses(code, eev)) continue;
      double len = lengths[e];
      double q = cq[e * 2 + eev];
      if (best == -1 || len < best_len || (len == best_len && q > best_q)) {
        vcand[v] = 1;
        best = e;
        best_len = len;
        best_q = q;
      }
    }
    vqual[v] = best_q;
    vrail[v] = best;
  }
  free(cq);
  signed char* keys_mark = (signed char*)om_malloc((size_t)nv);
  om_find_indset(m, 0, vqual, vcand, keys_mark);
  free(vcand);
  free(vqual);
  int nkeys;
  int* keys2verts = collect_marked(nv, keys_mark, &nkeys);
  free(keys_mark);
  /* coarsen_element_based2: find_rails */
  int* rails2edges = (int*)om_malloc(sizeof(int) * (size_t)(nkeys + 1));
  int* col_dirs = (int*)om_malloc(sizeof(int) * (size_t)(nkeys + 1));
  for (int k = 0; k < nkeys; ++k) {
    int v = keys2verts[k];
    for (int ve = v2e->a2ab[v]; ve < v2e->a2ab[v + 1]; ++ve) {
      if (v2e->ab2b[ve] == vrail[v]) {
        rails2edges[k] = v2e->ab2b[ve];
        col_dirs[k] = om_code_which_down(v2e->codes[ve]);
        break;
      }
    }
  }
  free(vrail);
  /* mark_dead_ents */
  signed char* dead[4] = {NULL, NULL, NULL, NULL};
  dead[1] = (signed char*)om_calloc((size_t)ne, 1);
  for (int k = 0; k < nkeys; ++k) dead[1][rails2edges[k]] = 1;
  for (int d = 2; d <= dim; ++d) dead[d] = (signed char*)om_calloc((size_t)m->nents[d], 1);
  for (int cell_dim = dim; cell_dim > 1; --cell_dim) {
    om_graph const* e2c = om_ask_up(m, 1, cell_dim);
    int const* cs2s = m->down[cell_dim];
    int nccs = cell_dim + 1;
    for (int k = 0; k < nkeys; ++k) {
      int e = rails2edges[k];
      int onto = 1 - col_dirs[k];
      for (int ec = e2c->a2ab[e]; ec < e2c->a2ab[e + 1]; ++ec) {
        int c = e2c->ab2b[ec];
        int cce = om_code_which_down(e2c->codes[ec]);
        int rot = om_code_rotation(e2c->codes[ec]);
        int ccv_onto = om_down_template(cell_dim, 1, cce, rot ^ onto);
        int ccs_opp = om_opposite_template(cell_dim, 0, ccv_onto);
        dead[cell_dim][c] = 1;
        dead[cell_dim - 1][cs2s[c * nccs + ccs_opp]] = 1;
      }
    }
  }
  int* keys2onto = (int*)om_malloc(sizeof(int) * (size_t)(nkeys + 1));
  for (int k = 0; k < nkeys; ++k) keys2onto[k] = ev2v[rails2edges[k] * 2 + (1 - col_dirs[k])];
  /* rebuild */
  om_mesh* nw = om_mesh_new(dim);
  nw->metric_ncomps = nc;
  int* o2n_verts = NULL;
  int* old_lows2new_lows = NULL;
  for (int d = 0; d <= dim; ++d) {
    int* k2p = (int*)om_calloc((size_t)(nkeys + 1), sizeof(int));
    int* pv = NULL;
    int* psrc = NULL;
    if (d > 0) {
      om_graph const* v2d = om_ask_up(m, 0, d);
      for (int k = 0; k < nkeys; ++k) {
        int v = keys2verts[k];
        int c = 0;
        for (int j = v2d->a2ab[v]; j < v2d->a2ab[v + 1]; ++j)
          if (!dead[d][v2d->ab2b[j]]) ++c;
        k2p[k + 1] = k2p[k] + c;
      }
      int nprods = k2p[nkeys];
      int nvd = d + 1;
      int const* dv2v = om_verts_of(m, d);
      pv = (int*)om_malloc(sizeof(int) * (size_t)(nprods + 1) * nvd);
      psrc = (int*)om_malloc(sizeof(int) * (size_t)(nprods + 1));
      for (int k = 0; k < nkeys; ++k) {
        int v = keys2verts[k];
        int prod = k2p[k];
        for (int j = v2d->a2ab[v]; j < v2d->a2ab[v + 1]; ++j) {
          int c = v2d->ab2b[j];
          if (dead[d][c]) continue;
          int ccv_col = om_code_which_down(v2d->codes[j]);
          for (int ccv = 0; ccv < nvd; ++ccv) {
            int ov = (ccv == ccv_col) ? keys2onto[k] : dv2v[c * nvd + ccv];
            pv[prod * nvd + ccv] = o2n_verts[ov];
          }
          psrc[prod] = c;
          ++prod;
        }
      }
    }
    mod_out mo;
    memset(&mo, 0, sizeof(mo));
    modify_ents_adapt(m, nw, d, 0, nkeys, keys2verts, k2p, pv, old_lows2new_lows, &mo);
    if (d == 0) {
      o2n_verts = (int*)om_malloc(sizeof(int) * (size_t)nv);
      memcpy(o2n_verts, mo.old2new, sizeof(int) * (size_t)nv);
      transfer_r(mo.nnew, dim, &mo, m->coords, NULL, &nw->coords);
      transfer_r(mo.nnew, nc, &mo, m->metric, NULL, &nw->metric);
      copy_vtags(m, nw, &mo);
      transfer_i(mo.nnew, &mo, m->class_id[0], NULL, &nw->class_id[0]);
      transfer_b(mo.nnew, &mo, m->class_dim[0], NULL, &nw->class_dim[0]);
      inherit_itags(m, nw, 0, &mo, 0, NULL, NULL);
    } else {
      transfer_class_from(m, nw, d, d, psrc, &mo);
      inherit_itags(m, nw, d, &mo, d, NULL, psrc);
      if (d == dim) {
        etags_pointwise(m, nw, &mo, nkeys, keys2verts, 0, k2p, pv);
        om_conserve_pass(m, nw, 1, 0, nkeys, keys2verts, k2p, mo.prods2new, mo.nsame,
            mo.same2old, mo.same2new);
      }
      transfer_lengths_qualities(m, nw, d, &mo);
    }
    free(old_lows2new_lows);
    old_lows2new_lows = (int*)om_malloc(sizeof(int) * (size_t)(m->nents[d] + 1));
    memcpy(old_lows2new_lows, mo.old2new, sizeof(int) * (size_t)m->nents[d]);
    mod_out_free(&mo);
    free(k2p);
    free(pv);
    free(psrc);
  }
  free(old_lows2new_lows);
  free(o2n_verts);
  for (int d = 0; d < 4; ++d) free(dead[d]);
  free(keys2onto);
  free(rails2edges);
  free(col_dirs);
  free(keys2verts);
  replace_mesh(m, nw);
  return 1;
}

static int coarsen_verts(om_mesh* m, om_adapt_opts const* o,
    signed char const* vert_marks, int overshoot_allowed, int improve) {
  int ne = m->nents[1];
  int const* ev2v = m->down[1];
  signed char* codes = (signed char*)om_malloc((size_t)ne);
  for (int e = 0; e < ne; ++e) {
    signed char code = DONT_COLLAPSE;
    for (int eev = 0; eev < 2; ++eev)
      if (vert_marks[ev2v[e * 2 + eev]]) code = (signed char)(code | (1 << eev));
    codes[e] = code;
  }
  int r = coarsen(m, o, codes, overshoot_allowed, improve);
  free(codes);
  return r;
}

int om_coarsen_by_size(om_mesh* m, om_adapt_opts const* o) {
  int ne = m->nents[1];
  double const* lengths = om_ask_lengths(m);
  om_ask_qualities(m);
  signed char* cand = (signed char*)om_malloc((size_t)ne);
  for (int e = 0; e < ne; ++e) cand[e] = lengths[e] < o->min_length_desired;
  int r = 0;
  if (any_marked(ne, cand)) {
    signed char* vm = mark_down(m, 1, 0, cand);
    r = coarsen_verts(m, o, vm, 0, 0);
    free(vm);
  }
  free(cand);
  return r;
}

int om_coarsen_slivers(om_mesh* m, om_adapt_opts const* o) {
  om_ask_lengths(m);
  signed char* marks = mark_sliver_layers(m, o->min_quality_desired, o->nsliver_layers);
  signed char* vm = mark_down(m, m->dim, 0, marks);
  free(marks);
  int r = coarsen_verts(m, o, vm, 1, 1);
  free(vm);
  return r;
}

/* ------------------------------------------------------------ swap 3D */
typedef struct {
  int size;
  int eev2v[2];
  int loop_verts[OM_MAX_EDGE_SWAP];
} swap_loop;

static void find_loop(om_mesh* m, int edge, swap_loop* loop) {
  om_graph const* e2t = om_ask_up(m, 1, 3);
  int const* ev2v = m->down[1];
  int const* tv2v = om_verts_of(m, 3);
  int begin = e2t->a2ab[edge], end = e2t->a2ab[edge + 1];
  loop->size = end - begin;
  if (loop->size > OM_MAX_EDGE_SWAP) return;
  loop->eev2v[0] = ev2v[edge * 2];
  loop->eev2v[1] = ev2v[edge * 2 + 1];
  int tmp[OM_MAX_EDGE_SWAP][2];
  for (int i = 0; i < OM_MAX_EDGE_SWAP; ++i) tmp[i][0] = tmp[i][1] = -1;
  for (int le = 0; le < loop->size; ++le) {
    int et = begin + le;
    int tet = e2t->ab2b[et];
    int rre = om_code_which_down(e2t->codes[et]);
    int rot = om_code_rotation(e2t->codes[et]);
    int rre_opp = om_opposite_template(3, 1, rre);
    for (int eev = 0; eev < 2; ++eev) {
      int rev = rot ^ eev;
      int rrv = om_down_template(3, 1, rre_opp, rev);
      tmp[le][eev] = tv2v[tet * 4 + rrv];
    }
  }
  for (int i = 0; i < loop->size - 1; ++i) {
    int j;
    for (j = i + 1; j < loop->size; ++j)
      if (tmp[j][0] == tmp[i][1]) break;
    if (j >= loop->size) {
      fprintf(stderr, "oracle: broken edge loop\n");
      abort();
    }
    int t0 = tmp[i + 1][0], t1 = tmp[i + 1][1];
    tmp[i + 1][0] = tmp[j][0];
    tmp[i + 1][1] = tmp[j][1];
    tmp[j][0] = t0;
    tmp[j][1] = t1;
  }
  for (int lv = 0; lv < loop->size; ++lv) loop->loop_verts[lv] = tmp[lv][0];
}

int om_kat_find_loop(om_mesh* m, int edge, int* eev2v, int* loop_verts) {
  swap_loop loop;
  find_loop(m, edge, &loop);
  eev2v[0] = loop.eev2v[0];
  eev2v[1] = loop.eev2v[1];
  for (int i = 0; i < loop.size && i < OM_MAX_EDGE_SWAP; ++i) loop_verts[i] = loop.loop_verts[i];
  return loop.size;
}

static double tet_quality_verts(om_mesh* m, int const* v) {
  int nc = m->metric_ncomps;
  int mdim = (nc == 1) ? 1 : 3;
  double const* p[4];
  double const* ms[4];
  for (int k = 0; k < 4; ++k) {
    p[k] = m->coords + v[k] * 3;
    ms[k] = m->metric + v[k] * nc;
  }
  return om_element_quality(3, mdim, p, ms);
}

static double edge_length_verts(om_mesh* m, int a, int b) {
  int nc = m->metric_ncomps;
  int sdim = m->dim;
  int mdim = (nc == 1) ? 1 : sdim;
  return om_metric_edge_length(sdim, mdim, m->coords + a * sdim,
      m->coords + b * sdim, m->metric + a * nc, m->metric + b * nc);
}

static void swap_choose(om_mesh* m, swap_loop const* loop, double max_length,
    int* mesh_out, double* qual_out) {
  int n = loop->size;
  int nmeshes = om_swap_mesh_counts[n];
  int nmesh_tris = om_swap_mesh_sizes[n];
  int tri_cached[OM_MAX_UNIQUE_TRIS] = {0};
  double tri_qual[OM_MAX_UNIQUE_TRIS] = {0};
  int edge_cached[OM_MAX_UNIQUE_EDGES] = {0};
  double edge_len[OM_MAX_UNIQUE_EDGES] = {0};
  int choice_mesh = -1;
  double choice_quality = 0.0;
  for (int mesh = 0; mesh < nmeshes; ++mesh) {
    double mesh_minqual = 1.0;
    for (int mt = 0; mt < nmesh_tris; ++mt) {
      int ut = om_swap_mesh_tri(n, mesh, mt);
      if (!tri_cached[ut]) {
        int const* tlv = om_swap_triangle(n, ut);
        int tv[4];
        for (int k = 0; k < 3; ++k) tv[k] = loop->loop_verts[tlv[k]];
        double tri_minqual = 1.0;
        for (int tri_tet = 0; tri_tet < 2; ++tri_tet) {
          tv[3] = loop->eev2v[1 - tri_tet];
          double q = tet_quality_verts(m, tv);
          tri_minqual = min2(tri_minqual, q);
          int t = tv[1];
          tv[1] = tv[2];
          tv[2] = t;
        }
        tri_cached[ut] = 1;
        tri_qual[ut] = tri_minqual;
      }
      mesh_minqual = min2(mesh_minqual, tri_qual[ut]);
      if (mesh_minqual < 0.0) break;
    }
    if (mesh_minqual > choice_quality) {
      int overshoot = 0;
      for (int me = 0; me < om_swap_nedges[n]; ++me) {
        int ue = om_swap_mesh_edge(n, mesh, me);
        if (!edge_cached[ue]) {
          int const* elv = om_swap_unique_edge(n, ue);
          edge_len[ue] = edge_length_verts(m, loop->loop_verts[elv[0]], loop->loop_verts[elv[1]]);
          edge_cached[ue] = 1;
        }
        if (edge_len[ue] > max_length) {
          overshoot = 1;
          break;
        }
      }
      if (!overshoot) {
        choice_mesh = mesh;
        choice_quality = mesh_minqual;
      }
    }
  }
  *mesh_out = choice_mesh;
  *qual_out = choice_quality;
}

static signed char* swap_candidates(om_mesh* m, om_adapt_opts const* o) {
  int dim = m->dim;
  signed char* elems = mark_sliver_layers(m, o->min_quality_desired, o->nsliver_layers);
  signed char* edges = mark_down(m, dim, 1, elems);
  free(elems);
  int ne = m->nents[1];
  int any = 0;
  for (int e = 0; e < ne; ++e) {
    edges[e] = (signed char)(edges[e] && (m->class_dim[1][e] == dim));
    if (edges[e] > 0) any = 1;
  }
  if (!any) {
    free(edges);
    return NULL;
  }
  return edges;
}

/* keep candidates whose new cavity quality beats the current worst element
 * around the edge (filter_swap_improve, src/Omega_h_swap.cpp:29-37) */
static double edge_old_quality(om_mesh* m, int e) {
  om_graph const* e2c = om_ask_up(m, 1, m->dim);
  double const* q = om_ask_qualities(m);
  double r = 1.79769313486231570815e+308;
  for (int j = e2c->a2ab[e]; j < e2c->a2ab[e + 1]; ++j) r = min2(r, q[e2c->ab2b[j]]);
  return r;
}

/* the per-candidate part of swap_edges_3d (find_loop, then
 * choose_int_edges over the loop's triangulations,
 * src/Omega_h_swap3d_qualities.cpp:11-75, src/Omega_h_swap3d_choice.hpp):
 * the chosen triangulation, and its quality when it improves on the
 * elements around the edge (-1 otherwise) */
void om_kat_swap3d_qualities(om_mesh* m, double max_length, int n, int const* edges,
    double* quals, signed char* configs) {
  om_ask_lengths(m);
  for (int i = 0; i < n; ++i) {
    int e = edges[i];
    swap_loop loop;
    find_loop(m, e, &loop);
    int config = -1;
    double q = -1.0;
    if (loop.size <= OM_MAX_EDGE_SWAP) swap_choose(m, &loop, max_length, &config, &q);
    configs[i] = (signed char)config;
    quals[i] = (config >= 0 && q > edge_old_quality(m, e)) ? q : -1.0;
  }
}

static void swap_rebuild(om_mesh* m, int nkeys, int const* keys2edges,
    int* const* k2p, int* const* pv) {
  int dim = m->dim;
  om_mesh* nw = om_mesh_new(dim);
  int nv = m->nents[0];
  int nc = m->metric_ncomps;
  nw->metric_ncomps = nc;
  nw->nents[0] = nv;
  nw->coords = (double*)om_malloc(sizeof(double) * (size_t)nv * dim);
  memcpy(nw->coords, m->coords, sizeof(double) * (size_t)nv * dim);
  nw->metric = (double*)om_malloc(sizeof(double) * (size_t)nv * nc);
  memcpy(nw->metric, m->metric, sizeof(double) * (size_t)nv * nc);
  copy_vtags(m, nw, NULL);
  nw->nitags = m->nitags;
  for (int t = 0; t < m->nitags; ++t) {
    nw->itags[t] = m->itags[t];
    for (int d = 0; d < 4; ++d) nw->itags[t].data[d] = NULL;
    size_t n = (size_t)nv * m->itags[t].ncomps;
    nw->itags[t].data[0] = (double*)om_malloc(sizeof(double) * (n + 1));
    memcpy(nw->itags[t].data[0], m->itags[t].data[0], sizeof(double) * n);
  }
  nw->class_id[0] = (int*)om_malloc(sizeof(int) * (size_t)nv);
  memcpy(nw->class_id[0], m->class_id[0], sizeof(int) * (size_t)nv);
  nw->class_dim[0] = (signed char*)om_malloc((size_t)nv);
  memcpy(nw->class_dim[0], m->class_dim[0], (size_t)nv);
  int* old_lows2new_lows = (int*)om_malloc(sizeof(int) * (size_t)nv);
  for (int i = 0; i < nv; ++i) old_lows2new_lows[i] = i;
  for (int d = 1; d <= dim; ++d) {
    mod_out mo;
    memset(&mo, 0, sizeof(mo));
    modify_ents_adapt(m, nw, d, 1, nkeys, keys2edges, k2p[d], pv[d], old_lows2new_lows, &mo);
    int* src = (int*)om_malloc(sizeof(int) * (size_t)(mo.nprods + 1));
    for (int k = 0; k < nkeys; ++k)
      for (int p = k2p[d][k]; p < k2p[d][k + 1]; ++p) src[p] = keys2edges[k];
    transfer_class_from(m, nw, d, 1, src, &mo);
    inherit_itags(m, nw, d, &mo, 1, NULL, src);
    if (d == dim) {
      etags_pointwise(m, nw, &mo, nkeys, keys2edges, 1, k2p[d], pv[d]);
      om_conserve_pass(m, nw, 2, 1, nkeys, keys2edges, k2p[d], mo.prods2new, mo.nsame,
          mo.same2old, mo.same2new);
    }
    free(src);
    transfer_lengths_qualities(m, nw, d, &mo);
    free(old_lows2new_lows);
    old_lows2new_lows = (int*)om_malloc(sizeof(int) * (size_t)(m->nents[d] + 1));
    memcpy(old_lows2new_lows, mo.old2new, sizeof(int) * (size_t)m->nents[d]);
    mod_out_free(&mo);
  }
  free(old_lows2new_lows);
  replace_mesh(m, nw);
}

static int swap_edges_3d(om_mesh* m, om_adapt_opts const* o) {
  om_ask_lengths(m);
  signed char* cand = swap_candidates(m, o);
  if (!cand) return 0;
  int ne = m->nents[1];
  int ncands;
  int* cands = collect_marked(ne, cand, &ncands);
  free(cand);
  signed char* edge_configs = (signed char*)om_malloc((size_t)ne);
  memset(edge_configs, -1, (size_t)ne);
  signed char* keep = (signed char*)om_calloc((size_t)ne, 1);
  double* edge_quals = (double*)om_malloc(sizeof(double) * (size_t)ne);
  for (int e = 0; e < ne; ++e) edge_quals[e] = -1.0;
  int any = 0;
  for (int c = 0; c < ncands; ++c) {
    int e = cands[c];
    swap_loop loop;
    find_loop(m, e, &loop);
    int config = -1;
    double q = -1.0;
    if (loop.size <= OM_MAX_EDGE_SWAP) swap_choose(m, &loop, o->max_length_allowed, &config, &q);
    edge_configs[e] = (signed char)config;
    if (config >= 0 && q > edge_old_quality(m, e)) {
      keep[e] = 1;
      edge_quals[e] = q;
      any = 1;
    }
  }
  free(cands);
  if (!any) {
    free(edge_configs);
    free(keep);
    free(edge_quals);
    return 0;
  }
  signed char* keys_mark = (signed char*)om_malloc((size_t)ne);
  om_find_indset(m, 1, edge_quals, keep, keys_mark);
  free(keep);
  free(edge_quals);
  int nkeys;
  int* keys2edges = collect_marked(ne, keys_mark, &nkeys);
  free(keys_mark);
  /* swap3d_keys_to_prods + swap3d_topology */
  int* k2p[4] = {NULL, NULL, NULL, NULL};
  int* pv[4] = {NULL, NULL, NULL, NULL};
  om_graph const* e2t = om_ask_up(m, 1, 3);
  for (int d = 1; d <= 3; ++d) {
    int* cnt = (int*)om_malloc(sizeof(int) * (size_t)(nkeys + 1));
    for (int k = 0; k < nkeys; ++k) {
      int e = keys2edges[k];
      int ls = e2t->a2ab[e + 1] - e2t->a2ab[e];
      int npt = om_swap_mesh_sizes[ls], npe = om_swap_nedges[ls];
      cnt[k] = (d == 1) ? npe : (d == 2) ? (npt + 2 * npe) : (2 * npt);
    }
    k2p[d] = om_offset_scan(nkeys, cnt);
    free(cnt);
    pv[d] = (int*)om_malloc(sizeof(int) * (size_t)(k2p[d][nkeys] + 1) * (d + 1));
  }
  for (int k = 0; k < nkeys; ++k) {
    int e = keys2edges[k];
    int config = edge_configs[e];
    swap_loop loop;
    find_loop(m, e, &loop);
    int n = loop.size;
    int npt = om_swap_mesh_sizes[n], npe = om_swap_nedges[n];
    for (int pe = 0; pe < npe; ++pe) {
      int ue = om_swap_mesh_edge(n, config, pe);
      int const* elv = om_swap_unique_edge(n, ue);
      int ev[2] = {loop.loop_verts[elv[0]], loop.loop_verts[elv[1]]};
      int pedge = k2p[1][k] + pe;
      pv[1][pedge * 2] = ev[0];
      pv[1][pedge * 2 + 1] = ev[1];
      for (int eev = 0; eev < 2; ++eev) {
        int ptri = k2p[2][k] + 2 * pe + eev;
        pv[2][ptri * 3 + 0] = ev[0];
        pv[2][ptri * 3 + 1] = ev[1];
        pv[2][ptri * 3 + 2] = loop.eev2v[eev];
      }
    }
    for (int pt = 0; pt < npt; ++pt) {
      int ut = om_swap_mesh_tri(n, config, pt);
      int const* tlv = om_swap_triangle(n, ut);
      int tv[3] = {loop.loop_verts[tlv[0]], loop.loop_verts[tlv[1]], loop.loop_verts[tlv[2]]};
      int ptri = k2p[2][k] + 2 * npe + pt;
      for (int j = 0; j < 3; ++j) pv[2][ptri * 3 + j] = tv[j];
      for (int eev = 0; eev < 2; ++eev) {
        int ptet = k2p[3][k] + 2 * pt + eev;
        int nt[4] = {tv[0], tv[1], tv[2], loop.eev2v[eev]};
        if (eev == 0) {
          int t = nt[1];
          nt[1] = nt[2];
          nt[2] = t;
        }
        for (int j = 0; j < 4; ++j) pv[3][ptet * 4 + j] = nt[j];
      }
    }
  }
  free(edge_configs);
  swap_rebuild(m, nkeys, keys2edges, k2p, pv);
  for (int d = 0; d < 4; ++d) {
    free(k2p[d]);
    free(pv[d]);
  }
  free(keys2edges);
  return 1;
}

/* ------------------------------------------------------------ swap 2D */
/* swap2d_topology, src/Omega_h_swap2d_topology.cpp:62-79: per key edge one
 * new edge (the quadrilateral's other diagonal, its vertices in the order
 * of the rotations of the edge's triangles) and two new triangles */
static void swap2d_products(om_mesh* m, int nkeys, int const* keys2edges, int* edge_verts,
    int* tri_verts) {
  om_graph const* e2t = om_ask_up(m, 1, 2);
  int const* tv2v = om_verts_of(m, 2);
  int const* ev2v = m->down[1];
  for (int k = 0; k < nkeys; ++k) {
    int e = keys2edges[k];
    int* ov = edge_verts + k * 2;
    for (int i = 0; i < 2; ++i) {
      int et = e2t->a2ab[e] + i;
      int tte = om_code_which_down(e2t->codes[et]);
      int rot = om_code_rotation(e2t->codes[et]);
      int t = e2t->ab2b[et];
      ov[rot] = tv2v[t * 3 + om_opposite_template(2, 1, tte)];
    }
    for (int i = 0; i < 2; ++i) {
      int* ntv = tri_verts + (k * 2 + i) * 3;
      ntv[0] = ev2v[e * 2 + (1 - i)];
      ntv[1] = ov[i];
      ntv[2] = ov[1 - i];
    }
  }
}

void om_kat_swap2d_topology(om_mesh* m, int nkeys, int const* keys2edges, int* edge_verts,
    int* tri_verts) {
  swap2d_products(m, nkeys, keys2edges, edge_verts, tri_verts);
}

static int swap_edges_2d(om_mesh* m, om_adapt_opts const* o) {
  om_ask_lengths(m);
  signed char* cand = swap_candidates(m, o);
  if (!cand) return 0;
  int ne = m->nents[1];
  om_graph const* e2t = om_ask_up(m, 1, 2);
  int const* tv2v = om_verts_of(m, 2);
  int const* ev2v = m->down[1];
  signed char* keep = (signed char*)om_calloc((size_t)ne, 1);
  double* edge_quals = (double*)om_malloc(sizeof(double) * (size_t)ne);
  int any = 0;
  int nc = m->metric_ncomps;
  int mdim = (nc == 1) ? 1 : 2;
  for (int e = 0; e < ne; ++e) {
    edge_quals[e] = -1.0;
    if (!cand[e]) continue;
    int ov[2];
    for (int i = 0; i < 2; ++i) {
      int et = e2t->a2ab[e] + i;
      int tte = om_code_which_down(e2t->codes[et]);
      int rot = om_code_rotation(e2t->codes[et]);
      int t = e2t->ab2b[et];
      int ttv = om_opposite_template(2, 1, tte);
      ov[rot] = tv2v[t * 3 + ttv];
    }
    double q;
    double l = edge_length_verts(m, ov[0], ov[1]);
    if (l > o->max_length_allowed) {
      q = -1.0;
    } else {
      double minqual = 1.0;
      for (int i = 0; i < 2; ++i) {
        int ntv[3] = {ev2v[e * 2 + (1 - i)], ov[i], ov[1 - i]};
        double const* p[3];
        double const* ms[3];
        for (int k = 0; k < 3; ++k) {
          p[k] = m->coords + ntv[k] * 2;
          ms[k] = m->metric + ntv[k] * nc;
        }
        minqual = min2(minqual, om_element_quality(2, mdim, p, ms));
      }
      q = minqual;
    }
    if (q > edge_old_quality(m, e)) {
      keep[e] = 1;
      edge_quals[e] = q;
      any = 1;
    }
  }
  free(cand);
  if (!any) {
    free(keep);
    free(edge_quals);
    return 0;
  }
  signed char* keys_mark = (signed char*)om_malloc((size_t)ne);
  om_find_indset(m, 1, edge_quals, keep, keys_mark);
  free(keep);
  free(edge_quals);
  int nkeys;
  int* keys2edges = collect_marked(ne, keys_mark, &nkeys);
  free(keys_mark);
  int* k2p[4] = {NULL, NULL, NULL, NULL};
  int* pv[4] = {NULL, NULL, NULL, NULL};
  k2p[1] = (int*)om_malloc(sizeof(int) * (size_t)(nkeys + 1));
  k2p[2] = (int*)om_malloc(sizeof(int) * (size_t)(nkeys + 1));
  for (int k = 0; k <= nkeys; ++k) {
    k2p[1][k] = k;
    k2p[2][k] = 2 * k;
  }
  pv[1] = (int*)om_malloc(sizeof(int) * (size_t)(nkeys + 1) * 2);
  pv[2] = (int*)om_malloc(sizeof(int) * (size_t)(nkeys + 1) * 6);
  swap2d_products(m, nkeys, keys2edges, pv[1], pv[2]);
  swap_rebuild(m, nkeys, keys2edges, k2p, pv);
  for (int d = 0; d < 4; ++d) {
    free(k2p[d]);
    free(pv[d]);
  }
  free(keys2edges);
  return 1;
}

int om_swap_edges(om_mesh* m, om_adapt_opts const* o) {
  if (m->dim == 3) return swap_edges_3d(m, o);
  return swap_edges_2d(m, o);
}

/* ------------------------------------------------------------ adapt */
static double min_quality(om_mesh* m) {
  double const* q = om_ask_qualities(m);
  double r = 1.79769313486231570815e+308;
  for (int i = 0; i < m->nents[m->dim]; ++i) r = min2(r, q[i]);
  return r;
}

static double now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void record(om_adapt_stats* s, int kind, int nin, om_mesh* m, double t) {
  if (!s || s->npasses >= OM_MAX_PASSES) return;
  s->kind[s->npasses] = kind;
  s->nelems_in[s->npasses] = nin;
  s->nelems_out[s->npasses] = m->nents[m->dim];
  s->seconds[s->npasses] = t;
  s->npasses++;
}

/* FNV-1a (64-bit) of the mesh after a pass, four digests:
 *   0: vertex lists of every dimension 1..dim (int32), then class_id (int32)
 *      and class_dim (int8) of every dimension 0..dim
 *   1: coordinate bits, 2: metric bits,
 *   3: edge length bits, then element quality bits */
static uint64_t fnv(uint64_t h, void const* p, size_t n) {
  unsigned char const* b = (unsigned char const*)p;
  for (size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= 1099511628211ull;
  }
  return h;
}

void om_mesh_digests(om_mesh* m, uint64_t* out) {
  uint64_t const basis = 1469598103934665603ull;
  uint64_t h = basis;
  for (int d = 1; d <= m->dim; ++d)
    h = fnv(h, om_verts_of(m, d), sizeof(int) * (size_t)m->nents[d] * (size_t)(d + 1));
  for (int d = 0; d <= m->dim; ++d) {
    h = fnv(h, m->class_id[d], sizeof(int) * (size_t)m->nents[d]);
    h = fnv(h, m->class_dim[d], (size_t)m->nents[d]);
  }
  out[0] = h;
  out[1] = fnv(basis, m->coords, sizeof(double) * (size_t)m->nents[0] * (size_t)m->dim);
  out[2] = fnv(basis, m->metric, sizeof(double) * (size_t)m->nents[0] * (size_t)m->metric_ncomps);
  h = fnv(basis, om_ask_lengths(m), sizeof(double) * (size_t)m->nents[1]);
  out[3] = fnv(h, om_ask_qualities(m), sizeof(double) * (size_t)m->nents[m->dim]);
}

static int run_pass(om_mesh* m, om_adapt_opts const* o, om_adapt_stats* s,
    int kind, uint64_t* digests) {
  if (s && s->npasses >= OM_MAX_PASSES) {
    fprintf(stderr, "oracle: adapt exceeded %d passes\n", OM_MAX_PASSES);
    abort();
  }
  int nin = m->nents[m->dim];
  double t0 = now_seconds();
  int r;
  if (kind == 0)
    r = om_refine_by_size(m, o);
  else if (kind == 1)
    r = om_coarsen_by_size(m, o);
  else if (kind == 2)
    r = om_swap_edges(m, o);
  else
    r = om_coarsen_slivers(m, o);
  double t1 = now_seconds();
  if (r) {
    if (digests && s && s->npasses < OM_MAX_PASSES)
      om_mesh_digests(m, digests + 4 * (size_t)s->npasses);
    record(s, kind, nin, m, t1 - t0);
  }
  return r;
}

static int adapt_impl(om_mesh* m, om_adapt_opts const* o, om_adapt_stats* s,
    uint64_t* digests) {
  if (s) s->npasses = 0;
  /* pre_adapt: nothing to do if quality and lengths already satisfy goals */
  double const* l = om_ask_lengths(m);
  double lmin = 1.79769313486231570815e+308, lmax = -1.79769313486231570815e+308;
  for (int e = 0; e < m->nents[1]; ++e) {
    lmin = min2(lmin, l[e]);
    lmax = (lmax < l[e]) ? l[e] : lmax;
  }
  if (min_quality(m) >= o->min_quality_desired && lmin >= o->min_length_desired &&
      lmax <= o->max_length_desired)
    return 0;
  /* setup_conservation_tags */
  om_conserve_setup(m);
  /* satisfy_lengths */
  int did;
  do {
    did = 0;
    if (o->should_refine && run_pass(m, o, s, 0, digests)) did = 1;
    if (o->should_coarsen && run_pass(m, o, s, 1, digests)) did = 1;
  } while (did);
  /* satisfy_quality */
  if (min_quality(m) < o->min_quality_desired) {
    do {
      if (o->should_swap && run_pass(m, o, s, 2, digests)) continue;
      if (o->should_coarsen_slivers && run_pass(m, o, s, 3, digests)) continue;
      break;
    } while (min_quality(m) < o->min_quality_desired);
  }
  /* correct_integral_errors */
  if (om_conserve_correct(m, OM_MAX_DIFFUSION_SWEEPS)) {
    fprintf(stderr, "oracle: correct_integral_errors did not finish\n");
    abort();
  }
  return 1;
}

int om_adapt(om_mesh* m, om_adapt_opts const* o, om_adapt_stats* s) {
  return adapt_impl(m, o, s, NULL);
}

int om_adapt_digests(om_mesh* m, om_adapt_opts const* o, om_adapt_stats* s,
    uint64_t* digests) {
  return adapt_impl(m, o, s, digests);
}

/* Oracle (test infrastructure only): mesh data model and adjacency
 * derivation of Omega_h, restated serially in C.
 *
 *   Mesh::ask_adj / derive_adj      src/Omega_h_mesh.cpp:307-357
 *   invert_adj (+ sort_by_high_index, separate_upward_*) src/Omega_h_adj.cpp:178-263
 *   transit                         src/Omega_h_adj.cpp:443-510
 *   form_uses / find_matches / reflect_down src/Omega_h_adj.cpp:155-176,297-441
 *   verts_across_edges / edges_across_tris / edges_across_tets
 *                                   src/Omega_h_adj.cpp:512-589
 *   add_edges                       src/Omega_h_graph.cpp:14-40
 *   elements_across_sides / mark_exposed_sides src/Omega_h_adj.cpp:591-657,
 *                                   src/Omega_h_mark.cpp:3-10
 *   find_unique                     src/Omega_h_adj.cpp:103-153
 *   measure_edges_metric / measure_qualities src/Omega_h_shape.cpp:7-48,
 *                                   src/Omega_h_quality.cpp:7-41
 * Up-adjacency rows are sorted by the high entity index, which makes every
 * derived structure independent of atomic ordering.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "om_internal.h"

void* om_malloc(size_t n) {
  void* p = malloc(n ? n : 1);
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    abort();
  }
  return p;
}

void* om_calloc(size_t n, size_t sz) {
  void* p = calloc(n ? n : 1, sz ? sz : 1);
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    abort();
  }
  return p;
}

int* om_offset_scan(int n, int const* counts) {
  int* o = (int*)om_malloc(sizeof(int) * (size_t)(n + 1));
  o[0] = 0;
  for (int i = 0; i < n; ++i) o[i + 1] = o[i] + counts[i];
  return o;
}

void om_graph_free(om_graph* g) {
  if (!g) return;
  free(g->a2ab);
  free(g->ab2b);
  free(g->codes);
  free(g);
}

/* ------------------------------------------------------------ templates */
static int const tet_faces[4][3] = {{0, 2, 1}, {0, 1, 3}, {1, 2, 3}, {2, 0, 3}};
static int const tet_edges[6][2] = {
    {0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
static int const tri_edges[3][2] = {{0, 1}, {1, 2}, {2, 0}};

int om_down_template(int elem_dim, int bdry_dim, int which_bdry, int which_vert) {
  if (bdry_dim == 0) return which_bdry;
  if (elem_dim == 3 && bdry_dim == 2) return tet_faces[which_bdry][which_vert];
  if (elem_dim == 3 && bdry_dim == 1) return tet_edges[which_bdry][which_vert];
  if (elem_dim == 2 && bdry_dim == 1) return tri_edges[which_bdry][which_vert];
  return -1;
}

int om_opposite_template(int elem_dim, int bdry_dim, int which_bdry) {
  static int const t30[4] = {2, 3, 1, 0};
  static int const t31[6] = {5, 3, 4, 1, 2, 0};
  static int const t32[4] = {3, 2, 0, 1};
  static int const t20[3] = {1, 2, 0};
  static int const t21[3] = {2, 0, 1};
  if (elem_dim == 3 && bdry_dim == 0) return t30[which_bdry];
  if (elem_dim == 3 && bdry_dim == 1) return t31[which_bdry];
  if (elem_dim == 3 && bdry_dim == 2) return t32[which_bdry];
  if (elem_dim == 2 && bdry_dim == 0) return t20[which_bdry];
  if (elem_dim == 2 && bdry_dim == 1) return t21[which_bdry];
  if (elem_dim == 1 && bdry_dim == 0) return 1 - which_bdry;
  return -1;
}

int om_simplex_degree(int from_dim, int to_dim) {
  static int const deg[4][4] = {
      {1, -1, -1, -1}, {2, 1, -1, -1}, {3, 3, 1, -1}, {4, 6, 4, 1}};
  return deg[from_dim][to_dim];
}

/* first entry (which_up == 0) of simplex_up_template, src/Omega_h_simplex.hpp:91-229 */
static void up_template0(int elem_dim, int bdry_dim, int which_bdry, int* up,
    int* which_down, int* is_flipped) {
  static int const t30[4][3] = {{0, 0, 0}, {1, 0, 0}, {2, 0, 0}, {5, 1, 0}};
  static int const t31[6][3] = {
      {0, 2, 1}, {0, 1, 1}, {0, 0, 1}, {1, 2, 1}, {2, 2, 1}, {3, 2, 1}};
  static int const t20[3][3] = {{0, 0, 0}, {1, 0, 0}, {2, 0, 0}};
  int const* t;
  if (elem_dim == 3 && bdry_dim == 0)
    t = t30[which_bdry];
  else if (elem_dim == 3 && bdry_dim == 1)
    t = t31[which_bdry];
  else
    t = t20[which_bdry];
  *up = t[0];
  *which_down = t[1];
  *is_flipped = t[2];
}

/* ------------------------------------------------------------ alignment */
static int rotate_index(int n, int index, int rotation) {
  return (index + rotation) % n;
}
static int invert_rotation(int n, int rotation) { return (n - rotation) % n; }
static int flip_vert_index(int n, int i) {
  return (i == 0 ? 0 : (1 + ((n - 1) - 1 - (i - 1))));
}
static int align_vert_index(int n, int index, signed char code) {
  int r = rotate_index(n, index, om_code_rotation(code));
  return om_code_is_flipped(code) ? flip_vert_index(n, r) : r;
}
static int align_edge_index(int n, int index, signed char code) {
  int r = rotate_index(n, index, om_code_rotation(code));
  return om_code_is_flipped(code) ? (n - 1 - r) : r;
}
static signed char invert_alignment(int n, signed char code) {
  if (om_code_is_flipped(code)) return code;
  return om_make_code(0, invert_rotation(n, om_code_rotation(code)), 0);
}

/* ------------------------------------------------------------ mesh */
om_mesh* om_mesh_new(int dim) {
  om_mesh* m = (om_mesh*)om_calloc(1, sizeof(om_mesh));
  m->dim = dim;
  for (int d = 0; d < 4; ++d) m->nents[d] = -1;
  return m;
}

void om_mesh_clear_caches(om_mesh* m) {
  for (int d = 0; d < 4; ++d) {
    if (d >= 2) free(m->verts_of[d]);
    m->verts_of[d] = NULL;
    for (int e = 0; e < 4; ++e) {
      om_graph_free(m->up[d][e]);
      m->up[d][e] = NULL;
    }
    om_graph_free(m->star[d]);
    m->star[d] = NULL;
  }
  om_graph_free(m->dual);
  m->dual = NULL;
  free(m->tet_edges_cache);
  free(m->tet_edge_codes_cache);
  m->tet_edges_cache = NULL;
  m->tet_edge_codes_cache = NULL;
}

void om_mesh_free(om_mesh* m) {
  if (!m) return;
  om_mesh_clear_caches(m);
  for (int d = 0; d < 4; ++d) {
    free(m->down[d]);
    free(m->down_codes[d]);
    free(m->class_id[d]);
    free(m->class_dim[d]);
  }
  free(m->coords);
  free(m->metric);
  for (int t = 0; t < m->nvtags; ++t) {
    free(m->vtags[t].data);
    free(m->vtags[t].err);
  }
  for (int t = 0; t < m->nitags; ++t)
    for (int d = 0; d < 4; ++d) free(m->itags[t].data[d]);
  for (int t = 0; t < m->netags; ++t) {
    free(m->etags[t].data);
    free(m->etags[t].err);
    free(m->etags[t].old);
  }
  free(m->length);
  free(m->quality);
  free(m);
}

/* transit, src/Omega_h_adj.cpp:443-510 */
static void transit(int nhigh, int high_dim, int low_dim, int const* hm2m,
    signed char const* m2hm_codes, int const* ml2l, signed char const* ml_codes,
    int* hl2l, signed char* codes) {
  int mid_dim = low_dim + 1;
  int nmids_per_high = om_simplex_degree(high_dim, mid_dim);
  int nlows_per_mid = om_simplex_degree(mid_dim, low_dim);
  int nlows_per_high = om_simplex_degree(high_dim, low_dim);
  for (int h = 0; h < nhigh; ++h) {
    int hl_begin = h * nlows_per_high;
    int hm_begin = h * nmids_per_high;
    for (int hl = 0; hl < nlows_per_high; ++hl) {
      int hm, hml, canon_flipped;
      up_template0(high_dim, low_dim, hl, &hm, &hml, &canon_flipped);
      int mm = hm2m[hm_begin + hm];
      signed char m2hm_code = m2hm_codes[hm_begin + hm];
      signed char hm2m_code = invert_alignment(nlows_per_mid, m2hm_code);
      int ml = (low_dim == 0) ? align_vert_index(nlows_per_mid, hml, hm2m_code)
                              : align_edge_index(nlows_per_mid, hml, hm2m_code);
      int ml_begin = mm * nlows_per_mid;
      int l = ml2l[ml_begin + ml];
      hl2l[hl_begin + hl] = l;
      if (low_dim == 1) {
        int region_face_flipped = om_code_is_flipped(hm2m_code);
        int face_edge_flipped = (om_code_rotation(ml_codes[ml_begin + ml]) == 1);
        int region_edge_flipped =
            region_face_flipped ^ face_edge_flipped ^ canon_flipped;
        codes[hl_begin + hl] = om_make_code(0, region_edge_flipped, 0);
      }
    }
  }
}

/* tet -> edge (with codes) derived by transit(tet->tri, tri->edge) */
static void ask_tet_edges(om_mesh* m, int const** t2e, signed char const** codes) {
  if (!m->tet_edges_cache) {
    int nt = m->nents[3];
    m->tet_edges_cache = (int*)om_malloc(sizeof(int) * (size_t)nt * 6);
    m->tet_edge_codes_cache = (signed char*)om_malloc((size_t)nt * 6);
    transit(nt, 3, 1, m->down[3], m->down_codes[3], m->down[2],
        m->down_codes[2], m->tet_edges_cache, m->tet_edge_codes_cache);
  }
  *t2e = m->tet_edges_cache;
  *codes = m->tet_edge_codes_cache;
}

int const* om_verts_of(om_mesh* m, int d) {
  if (d == 1) return m->down[1];
  if (!m->verts_of[d]) {
    int n = m->nents[d];
    int* out = (int*)om_malloc(sizeof(int) * (size_t)n * (d + 1));
    if (d == 2) {
      transit(n, 2, 0, m->down[2], m->down_codes[2], m->down[1], NULL, out, NULL);
    } else {
      int const* t2e;
      signed char const* codes;
      ask_tet_edges(m, &t2e, &codes);
      transit(n, 3, 0, t2e, codes, m->down[1], NULL, out, NULL);
    }
    m->verts_of[d] = out;
  }
  return m->verts_of[d];
}

int const* om_ask_down(om_mesh* m, int high, int low, signed char const** codes) {
  if (low == 0) {
    if (codes) *codes = NULL;
    return om_verts_of(m, high);
  }
  if (low == high - 1) {
    if (codes) *codes = m->down_codes[high];
    return m->down[high];
  }
  /* 3 -> 1 */
  int const* t2e;
  signed char const* c;
  ask_tet_edges(m, &t2e, &c);
  if (codes) *codes = c;
  return t2e;
}

/* invert_adj, src/Omega_h_adj.cpp:231-263 */
static om_graph* invert_adj(int nhigh, int deg, int nlow, int const* hl2l,
    signed char const* down_codes) {
  om_graph* g = (om_graph*)om_calloc(1, sizeof(om_graph));
  g->n = nlow;
  int* counts = (int*)om_calloc((size_t)nlow, sizeof(int));
  int nhl = nhigh * deg;
  for (int i = 0; i < nhl; ++i) counts[hl2l[i]]++;
  g->a2ab = om_offset_scan(nlow, counts);
  g->ab2b = (int*)om_malloc(sizeof(int) * (size_t)nhl);
  g->codes = (signed char*)om_malloc((size_t)nhl);
  memset(counts, 0, sizeof(int) * (size_t)nlow);
  /* filling in high order keeps each row sorted by high index, which is
     what sort_by_high_index establishes after the atomic fill */
  for (int hl = 0; hl < nhl; ++hl) {
    int l = hl2l[hl];
    int pos = g->a2ab[l] + counts[l]++;
    int h = hl / deg;
    int which_down = hl % deg;
    g->ab2b[pos] = h;
    if (down_codes) {
      signed char dc = down_codes[hl];
      g->codes[pos] = om_make_code(
          om_code_is_flipped(dc), om_code_rotation(dc), which_down);
    } else {
      g->codes[pos] = om_make_code(0, 0, which_down);
    }
  }
  free(counts);
  return g;
}

om_graph const* om_ask_up(om_mesh* m, int low, int high) {
  if (!m->up[low][high]) {
    signed char const* codes = NULL;
    int const* down = om_ask_down(m, high, low, &codes);
    m->up[low][high] = invert_adj(m->nents[high],
        om_simplex_degree(high, low), m->nents[low], down, codes);
  }
  return m->up[low][high];
}

/* IsMatch + find_matches_deg, src/Omega_h_adj.cpp:297-396 */
static int is_match(int deg, int const* av, int const* bv, int which_down,
    signed char* code) {
  if (deg == 2) {
    if (av[1] == bv[1 - which_down]) {
      *code = om_make_code(0, which_down, 0);
      return 1;
    }
    return 0;
  }
  if (av[1] == bv[(which_down + 1) % 3] && av[2] == bv[(which_down + 2) % 3]) {
    *code = om_make_code(0, invert_rotation(3, which_down), 0);
    return 1;
  }
  if (av[1] == bv[(which_down + 2) % 3] && av[2] == bv[(which_down + 1) % 3]) {
    *code = om_make_code(1, invert_rotation(3, which_down), 0);
    return 1;
  }
  return 0;
}

static void find_matches(int deg, int na, int const* av2v, int const* bv2v,
    om_graph const* v2b, int* a2b, signed char* codes) {
  for (int a = 0; a < na; ++a) {
    int fv = av2v[a * deg];
    int found = 0;
    for (int vb = v2b->a2ab[fv]; vb < v2b->a2ab[fv + 1]; ++vb) {
      int b = v2b->ab2b[vb];
      int which_down = om_code_which_down(v2b->codes[vb]);
      signed char code;
      if (is_match(deg, av2v + a * deg, bv2v + b * deg, which_down, &code)) {
        if (found) {
          fprintf(stderr, "oracle: duplicate match in reflect_down\n");
          abort();
        }
        a2b[a] = b;
        codes[a] = code;
        found = 1;
      }
    }
    if (!found) {
      fprintf(stderr, "oracle: no match in reflect_down\n");
      abort();
    }
  }
}

/* form_uses, src/Omega_h_adj.cpp:155-176: the vertex lists of the
 * low_dim-dimensional sides of each high entity, in template order */
static int* form_uses(int nhigh, int high_dim, int low_dim, int const* hv2v) {
  int nvh = high_dim + 1;
  int nlh = om_simplex_degree(high_dim, low_dim);
  int nvl = low_dim + 1;
  int* uv2v = (int*)om_malloc(sizeof(int) * ((size_t)nhigh * nlh * nvl + 1));
  for (int h = 0; h < nhigh; ++h)
    for (int u = 0; u < nlh; ++u)
      for (int uv = 0; uv < nvl; ++uv)
        uv2v[(h * nlh + u) * nvl + uv] =
            hv2v[h * nvh + om_down_template(high_dim, low_dim, u, uv)];
  return uv2v;
}

void om_kat_form_uses(int nhigh, int high_dim, int low_dim, int const* hv2v, int* out) {
  int* u = form_uses(nhigh, high_dim, low_dim, hv2v);
  size_t n = (size_t)nhigh * om_simplex_degree(high_dim, low_dim) * (low_dim + 1);
  memcpy(out, u, sizeof(int) * n);
  free(u);
}

/* reflect_down(hv2v, lv2v, nverts, high_dim, low_dim),
 * src/Omega_h_adj.cpp:424-441: the low entities (given by vertex lists)
 * on each high entity's sides, with alignment codes */
void om_kat_reflect_down(int high_dim, int low_dim, int nhigh, int const* hv2v, int nlow,
    int const* lv2v, int nverts, int* ab2b, signed char* codes) {
  int nvl = low_dim + 1;
  int nuses = nhigh * om_simplex_degree(high_dim, low_dim);
  int* uv2v = form_uses(nhigh, high_dim, low_dim, hv2v);
  om_graph* v2l = invert_adj(nlow, nvl, nverts, lv2v, NULL);
  find_matches(nvl, nuses, uv2v, lv2v, v2l, ab2b, codes);
  om_graph_free(v2l);
  free(uv2v);
}

/* invert_map_by_atomics / invert_map_by_sorting (src/Omega_h_map.cpp:238-280):
 * rows of the inverse sorted by the high index */
void om_kat_invert_adj(int nhigh, int deg, int nlow, int const* hl2l,
    signed char const* down_codes, int* a2ab, int* ab2b, signed char* codes) {
  om_graph* g = invert_adj(nhigh, deg, nlow, hl2l, down_codes);
  memcpy(a2ab, g->a2ab, sizeof(int) * (size_t)(nlow + 1));
  memcpy(ab2b, g->ab2b, sizeof(int) * (size_t)nhigh * (size_t)deg);
  memcpy(codes, g->codes, (size_t)nhigh * (size_t)deg);
  om_graph_free(g);
}

void om_kat_invert_map(int n, int const* hl2l, int nlow, int* a2ab, int* ab2b) {
  om_graph* g = invert_adj(n, 1, nlow, hl2l, NULL);
  memcpy(a2ab, g->a2ab, sizeof(int) * (size_t)(nlow + 1));
  memcpy(ab2b, g->ab2b, sizeof(int) * (size_t)n);
  om_graph_free(g);
}

/* index alignment of src/Omega_h_align.hpp:65-85 */
int om_kat_align_vert_index(int n, int index, int code) {
  return align_vert_index(n, index, (signed char)code);
}

/* reflect_down of new entities of dimension high_dim (>= 2) given by their
 * vertex lists, against the already-built lower dimension of m_new. */
void om_reflect_down(om_mesh* m_new, int high_dim, int nprods,
    int const* prod_verts, int* out_down, signed char* out_codes) {
  int low_dim = high_dim - 1;
  int nlh = high_dim + 1;
  int nvl = low_dim + 1;
  int nuses = nprods * nlh;
  int* uv2v = form_uses(nprods, high_dim, low_dim, prod_verts);
  int const* lv2v = om_verts_of(m_new, low_dim);
  om_graph const* v2l = om_ask_up(m_new, 0, low_dim);
  find_matches(nvl, nuses, uv2v, lv2v, v2l, out_down, out_codes);
  free(uv2v);
}

/* ------------------------------------------------------------ stars */
static om_graph* verts_across_edges(om_mesh* m) {
  om_graph const* v2e = om_ask_up(m, 0, 1);
  int const* ev2v = m->down[1];
  om_graph* g = (om_graph*)om_calloc(1, sizeof(om_graph));
  int nv = m->nents[0];
  g->n = nv;
  g->a2ab = (int*)om_malloc(sizeof(int) * (size_t)(nv + 1));
  memcpy(g->a2ab, v2e->a2ab, sizeof(int) * (size_t)(nv + 1));
  int nve = v2e->a2ab[nv];
  g->ab2b = (int*)om_malloc(sizeof(int) * (size_t)nve);
  for (int ve = 0; ve < nve; ++ve) {
    int e = v2e->ab2b[ve];
    int eev = om_code_which_down(v2e->codes[ve]);
    g->ab2b[ve] = ev2v[e * 2 + (1 - eev)];
  }
  return g;
}

static om_graph* edges_across(om_mesh* m) {
  int ne = m->nents[1];
  om_graph const* e2f = om_ask_up(m, 1, 2);
  int const* fe2e = m->down[2];
  om_graph const* e2t = NULL;
  int const* te2e = NULL;
  if (m->dim == 3) {
    e2t = om_ask_up(m, 1, 3);
    signed char const* c;
    te2e = om_ask_down(m, 3, 1, &c);
  }
  int* deg = (int*)om_malloc(sizeof(int) * (size_t)ne);
  for (int e = 0; e < ne; ++e) {
    deg[e] = 2 * (e2f->a2ab[e + 1] - e2f->a2ab[e]);
    if (e2t) deg[e] += e2t->a2ab[e + 1] - e2t->a2ab[e];
  }
  om_graph* g = (om_graph*)om_calloc(1, sizeof(om_graph));
  g->n = ne;
  g->a2ab = om_offset_scan(ne, deg);
  g->ab2b = (int*)om_malloc(sizeof(int) * (size_t)g->a2ab[ne]);
  for (int e = 0; e < ne; ++e) {
    int k = g->a2ab[e];
    for (int ef = e2f->a2ab[e]; ef < e2f->a2ab[e + 1]; ++ef) {
      int f = e2f->ab2b[ef];
      int ffe = om_code_which_down(e2f->codes[ef]);
      g->ab2b[k++] = fe2e[f * 3 + ((ffe + 1) % 3)];
      g->ab2b[k++] = fe2e[f * 3 + ((ffe + 2) % 3)];
    }
    if (e2t) {
      for (int er = e2t->a2ab[e]; er < e2t->a2ab[e + 1]; ++er) {
        int r = e2t->ab2b[er];
        int rre = om_code_which_down(e2t->codes[er]);
        int rre_opp = om_opposite_template(3, 1, rre);
        g->ab2b[k++] = te2e[r * 6 + rre_opp];
      }
    }
  }
  free(deg);
  return g;
}

om_graph const* om_ask_star(om_mesh* m, int d) {
  if (!m->star[d]) {
    m->star[d] = (d == 0) ? verts_across_edges(m) : edges_across(m);
  }
  return m->star[d];
}

om_graph const* om_ask_dual(om_mesh* m) {
  if (!m->dual) {
    int dim = m->dim;
    int ne = m->nents[dim];
    int ns = m->nents[dim - 1];
    om_graph const* s2e = om_ask_up(m, dim - 1, dim);
    int const* es2s = m->down[dim];
    int nspe = dim + 1;
    int* deg = (int*)om_calloc((size_t)ne, sizeof(int));
    for (int e = 0; e < ne; ++e)
      for (int k = 0; k < nspe; ++k) {
        int s = es2s[e * nspe + k];
        if ((s2e->a2ab[s + 1] - s2e->a2ab[s]) >= 2) deg[e]++;
      }
    (void)ns;
    om_graph* g = (om_graph*)om_calloc(1, sizeof(om_graph));
    g->n = ne;
    g->a2ab = om_offset_scan(ne, deg);
    g->ab2b = (int*)om_malloc(sizeof(int) * (size_t)g->a2ab[ne]);
    for (int e = 0; e < ne; ++e) {
      int k = g->a2ab[e];
      for (int j = 0; j < nspe; ++j) {
        int s = es2s[e * nspe + j];
        if ((s2e->a2ab[s + 1] - s2e->a2ab[s]) >= 2) {
          int se = s2e->a2ab[s];
          if (s2e->ab2b[se] == e) ++se;
          g->ab2b[k++] = s2e->ab2b[se];
        }
      }
    }
    free(deg);
    m->dual = g;
  }
  return m->dual;
}

/* ------------------------------------------------------------ measures */
void om_measure_edges(om_mesh* m, int n, int const* edges, double* out) {
  int sdim = m->dim;
  int nc = m->metric_ncomps;
  int mdim = (nc == 1) ? 1 : sdim;
  int const* ev2v = m->down[1];
  for (int i = 0; i < n; ++i) {
    int e = edges ? edges[i] : i;
    int a = ev2v[e * 2 + 0], b = ev2v[e * 2 + 1];
    out[i] = om_metric_edge_length(sdim, mdim, m->coords + a * sdim,
        m->coords + b * sdim, m->metric + a * nc, m->metric + b * nc);
  }
}

void om_measure_elems(om_mesh* m, int n, int const* elems, double* out) {
  int dim = m->dim;
  int nc = m->metric_ncomps;
  int mdim = (nc == 1) ? 1 : dim;
  int const* ev2v = om_verts_of(m, dim);
  for (int i = 0; i < n; ++i) {
    int e = elems ? elems[i] : i;
    double const* p[4];
    double const* ms[4];
    for (int k = 0; k < dim + 1; ++k) {
      int v = ev2v[e * (dim + 1) + k];
      p[k] = m->coords + v * dim;
      ms[k] = m->metric + v * nc;
    }
    out[i] = om_element_quality(dim, mdim, p, ms);
  }
}

double const* om_ask_lengths(om_mesh* m) {
  if (!m->length) {
    m->length = (double*)om_malloc(sizeof(double) * (size_t)m->nents[1]);
    om_measure_edges(m, m->nents[1], NULL, m->length);
  }
  return m->length;
}

double const* om_ask_qualities(om_mesh* m) {
  if (!m->quality) {
    m->quality = (double*)om_malloc(sizeof(double) * (size_t)m->nents[m->dim]);
    om_measure_elems(m, m->nents[m->dim], NULL, m->quality);
  }
  return m->quality;
}

/* ------------------------------------------------------------ sorting */
static int64_t const* g_sort_keys;
static int g_sort_width;

static int cmp_keys(int a, int b) {
  for (int i = 0; i < g_sort_width; ++i) {
    int64_t x = g_sort_keys[(size_t)a * g_sort_width + i];
    int64_t y = g_sort_keys[(size_t)b * g_sort_width + i];
    if (x != y) return x < y ? -1 : 1;
  }
  return 0;
}

static void merge_sort(int* a, int* tmp, int n) {
  if (n < 2) return;
  int h = n / 2;
  merge_sort(a, tmp, h);
  merge_sort(a + h, tmp, n - h);
  int i = 0, j = h, k = 0;
  while (i < h && j < n) {
    if (cmp_keys(a[j], a[i]) < 0)
      tmp[k++] = a[j++];
    else
      tmp[k++] = a[i++];
  }
  while (i < h) tmp[k++] = a[i++];
  while (j < n) tmp[k++] = a[j++];
  memcpy(a, tmp, sizeof(int) * (size_t)n);
}

/* stable sort of tuple keys, src/Omega_h_sort.cpp:43-92 */
void om_stable_sort_keys(int n, int width, int64_t const* keys, int* perm) {
  for (int i = 0; i < n; ++i) perm[i] = i;
  int* tmp = (int*)om_malloc(sizeof(int) * (size_t)n);
  g_sort_keys = keys;
  g_sort_width = width;
  merge_sort(perm, tmp, n);
  free(tmp);
}

/* find_unique: canonical keys, stable sort, last use of each group,
 * src/Omega_h_adj.cpp:55-153 */
int* om_find_unique(int nhigh, int high_dim, int low_dim, int const* hv2v,
    int* nlow_out) {
  int nvh = high_dim + 1;
  int nlh = om_simplex_degree(high_dim, low_dim);
  int deg = low_dim + 1;
  int nuses = nhigh * nlh;
  int* uv2v = (int*)om_malloc(sizeof(int) * (size_t)nuses * deg);
  for (int h = 0; h < nhigh; ++h)
    for (int u = 0; u < nlh; ++u)
      for (int uv = 0; uv < deg; ++uv)
        uv2v[(h * nlh + u) * deg + uv] =
            hv2v[h * nvh + om_down_template(high_dim, low_dim, u, uv)];
  int64_t* canon = (int64_t*)om_malloc(sizeof(int64_t) * (size_t)nuses * deg);
  for (int e = 0; e < nuses; ++e) {
    int const* ev = uv2v + e * deg;
    int min_j = 0;
    int min_v = ev[0];
    for (int j = 1; j < deg; ++j)
      if (ev[j] < min_v) {
        min_j = j;
        min_v = ev[j];
      }
    int rotation = invert_rotation(deg, min_j);
    int tmp[3];
    for (int j = 0; j < deg; ++j) tmp[rotate_index(deg, j, rotation)] = ev[j];
    int flipped = (deg == 3) ? (tmp[2] < tmp[1]) : 0;
    if (flipped) {
      int t = tmp[1];
      tmp[1] = tmp[2];
      tmp[2] = t;
    }
    for (int j = 0; j < deg; ++j) canon[(size_t)e * deg + j] = tmp[j];
  }
  int* sorted2u = (int*)om_malloc(sizeof(int) * (size_t)nuses);
  om_stable_sort_keys(nuses, deg, canon, sorted2u);
  int nlow = 0;
  for (int s = 0; s < nuses; ++s) {
    int last = (s == nuses - 1);
    if (!last) {
      int a = sorted2u[s], b = sorted2u[s + 1];
      for (int j = 0; j < deg; ++j)
        if (canon[(size_t)a * deg + j] != canon[(size_t)b * deg + j]) last = 1;
    }
    if (last) ++nlow;
  }
  int* out = (int*)om_malloc(sizeof(int) * (size_t)nlow * deg);
  int k = 0;
  for (int s = 0; s < nuses; ++s) {
    int last = (s == nuses - 1);
    if (!last) {
      int a = sorted2u[s], b = sorted2u[s + 1];
      for (int j = 0; j < deg; ++j)
        if (canon[(size_t)a * deg + j] != canon[(size_t)b * deg + j]) last = 1;
    }
    if (last) {
      int u = sorted2u[s];
      for (int j = 0; j < deg; ++j) out[k * deg + j] = uv2v[u * deg + j];
      ++k;
    }
  }
  free(uv2v);
  free(canon);
  free(sorted2u);
  *nlow_out = nlow;
  return out;
}

/* Oracle (test infrastructure only): build_box, restated serially in C.
 *
 *   make_2d_box / make_3d_box       src/Omega_h_box.cpp:25-102
 *   tris_from_quads / tets_from_hexes src/Omega_h_simplify.cpp:14-228
 *   build_from_elems2verts          src/Omega_h_build.cpp:24-105
 *   reorder_by_hilbert              src/Omega_h_reorder.cpp:10-64
 *   hilbert::from_spatial (Skilling 2004, "Programming the Hilbert curve")
 *                                   src/Omega_h_hilbert.hpp, src/Omega_h_hilbert.cpp:23-55
 *   unmap_mesh                      src/Omega_h_unmap_mesh.cpp:108-134
 *   classify_box                    src/Omega_h_box.cpp:104-160
 *   build_box                       src/Omega_h_build.cpp:136-149
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "om_internal.h"

/* ------------------------------------------------------------ hex split */
static int find_min(int const* v, int n) {
  int min_i = 0;
  int min_v = v[0];
  for (int i = 1; i < n; ++i)
    if (v[i] < min_v) {
      min_i = i;
      min_v = v[i];
    }
  return min_i;
}

static int lt4(int vi, int vj, int vk, int vl) {
  int a = vi < vj ? vi : vj;
  int b = vk < vl ? vk : vl;
  return a < b;
}

/* the reference's rot() shifts with a forward loop (src/Omega_h_simplify.cpp:37-41);
 * restated as written */
static void rot(int* v, int n) {
  int tmp = v[n - 1];
  for (int i = 0; i < n - 1; ++i) v[i + 1] = v[i];
  v[0] = tmp;
}
static void rot_ntimes(int* v, int nv, int ntimes) {
  for (int i = 0; i < ntimes; ++i) rot(v, nv);
}
static void rot_to_first(int* v, int nv, int first) {
  rot_ntimes(v, nv, ((nv - first) % nv));
}

static int const htv_0[5][4] = {
    {0, 1, 2, 5}, {0, 2, 7, 5}, {0, 2, 3, 7}, {0, 5, 7, 4}, {2, 7, 5, 6}};
static int const htv_1[6][4] = {{0, 5, 7, 4}, {0, 1, 7, 5}, {1, 6, 7, 5},
    {0, 7, 2, 3}, {0, 7, 1, 2}, {1, 7, 6, 2}};
static int const htv_2[6][4] = {{0, 4, 5, 6}, {0, 3, 7, 6}, {0, 7, 4, 6},
    {0, 1, 2, 5}, {0, 3, 6, 2}, {0, 6, 5, 2}};
static int const htv_3[6][4] = {{0, 2, 3, 6}, {0, 3, 7, 6}, {0, 7, 4, 6},
    {0, 5, 6, 4}, {1, 5, 6, 0}, {1, 6, 2, 0}};
static int const hex_flip_pairs[4][2] = {{0, 4}, {3, 5}, {1, 7}, {2, 6}};
static int const hex_bur_faces[3][4] = {
    {6, 5, 1, 2}, {6, 2, 3, 7}, {6, 7, 4, 5}};
static int const hex_bur_ring[6] = {1, 2, 3, 7, 5, 4};

static void hex_bur_rot_ntimes(int* h, int ntimes) {
  int tmp[6];
  for (int i = 0; i < 6; ++i) tmp[i] = h[hex_bur_ring[i]];
  rot_ntimes(tmp, 6, ntimes * 2);
  for (int i = 0; i < 6; ++i) h[hex_bur_ring[i]] = tmp[i];
}

static void hex_prepare(int const* hv, int* h, int* diags, int* ndiags) {
  for (int i = 0; i < 8; ++i) h[i] = hv[i];
  int min_i = find_min(h, 8);
  rot_to_first(h + 0, 4, min_i % 4);
  rot_to_first(h + 4, 4, min_i % 4);
  if (min_i >= 4) {
    for (int i = 0; i < 4; ++i) {
      int a = hex_flip_pairs[i][0], b = hex_flip_pairs[i][1];
      int t = h[a];
      h[a] = h[b];
      h[b] = t;
    }
  }
  for (int i = 0; i < 3; ++i)
    diags[i] = lt4(h[hex_bur_faces[i][0]], h[hex_bur_faces[i][2]],
        h[hex_bur_faces[i][1]], h[hex_bur_faces[i][3]]);
  *ndiags = diags[0] + diags[1] + diags[2];
}

static int* tets_from_hexes(int nh, int const* hv2v, int* nt_out) {
  int* deg = (int*)om_malloc(sizeof(int) * (size_t)nh);
  for (int hh = 0; hh < nh; ++hh) {
    int h[8], d[3], nd;
    hex_prepare(hv2v + hh * 8, h, d, &nd);
    deg[hh] = nd == 0 ? 5 : 6;
  }
  int* h2ht = om_offset_scan(nh, deg);
  int nt = h2ht[nh];
  int* tv2v = (int*)om_malloc(sizeof(int) * (size_t)nt * 4);
  for (int hh = 0; hh < nh; ++hh) {
    int h[8], d[3], nd;
    hex_prepare(hv2v + hh * 8, h, d, &nd);
    int const(*tmpl)[4];
    int n;
    if (nd == 0) {
      tmpl = htv_0;
      n = 5;
    } else if (nd == 1) {
      int face = -1;
      for (int i = 0; i < 3; ++i)
        if (d[i]) face = i;
      hex_bur_rot_ntimes(h, ((3 - face) % 3));
      tmpl = htv_1;
      n = 6;
    } else if (nd == 2) {
      int face = -1;
      for (int i = 0; i < 3; ++i)
        if (!d[i]) face = i;
      hex_bur_rot_ntimes(h, ((3 - face) % 3));
      tmpl = htv_2;
      n = 6;
    } else {
      tmpl = htv_3;
      n = 6;
    }
    int t = h2ht[hh];
    for (int i = 0; i < n; ++i, ++t)
      for (int j = 0; j < 4; ++j) tv2v[t * 4 + j] = h[tmpl[i][j]];
  }
  free(deg);
  free(h2ht);
  *nt_out = nt;
  return tv2v;
}

static int* tris_from_quads(int nq, int const* qv2v) {
  static int const qtv[2][3] = {{0, 1, 2}, {2, 3, 0}};
  int* tv2v = (int*)om_malloc(sizeof(int) * (size_t)nq * 6);
  for (int q = 0; q < nq; ++q) {
    int v[4];
    for (int i = 0; i < 4; ++i) v[i] = qv2v[q * 4 + i];
    int min_i = find_min(v, 4);
    rot_to_first(v, 4, min_i);
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 3; ++j) tv2v[(q * 2 + i) * 3 + j] = v[qtv[i][j]];
  }
  return tv2v;
}

/* ------------------------------------------------------------ assembly */
static void add_ents(om_mesh* m, int d, int n, int* ev2v) {
  if (d == 1) {
    m->down[1] = ev2v;
    m->down_codes[1] = NULL;
    m->nents[1] = n;
    return;
  }
  int deg = d + 1;
  int* down = (int*)om_malloc(sizeof(int) * (size_t)n * deg);
  signed char* codes = (signed char*)om_malloc((size_t)n * deg);
  om_reflect_down(m, d, n, ev2v, down, codes);
  m->down[d] = down;
  m->down_codes[d] = codes;
  m->nents[d] = n;
  free(ev2v);
}

static void build_from_elems(om_mesh* m, int nelems, int* ev2v) {
  int dim = m->dim;
  for (int md = 1; md < dim; ++md) {
    int nlow;
    int* mv2v = om_find_unique(nelems, dim, md, ev2v, &nlow);
    add_ents(m, md, nlow, mv2v);
  }
  add_ents(m, dim, nelems, ev2v);
}

/* ------------------------------------------------------------ hilbert */
typedef uint64_t hcoord;

static void axes_to_transpose(hcoord* X, int b, int n) {
  hcoord M = (hcoord)1 << (b - 1), P, Q, t;
  for (Q = M; Q > 1; Q >>= 1) {
    P = Q - 1;
    for (int i = 0; i < n; i++) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < n; i++) X[i] ^= X[i - 1];
  t = 0;
  for (Q = M; Q > 1; Q >>= 1)
    if (X[n - 1] & Q) t ^= Q - 1;
  for (int i = 0; i < n; i++) X[i] ^= t;
}

static void untranspose(hcoord const* in, hcoord* out, int b, int n) {
  for (int i = 0; i < n; ++i) out[i] = 0;
  for (int i = 0; i < (b * n); ++i)
    out[i / b] |= (((in[i % n] >> (b - 1 - (i / n))) & 1) << (b - 1 - (i % b)));
}

static void reorder_by_hilbert(om_mesh* m) {
  int dim = m->dim;
  int nv = m->nents[0];
  double bmin[3], bmax[3];
  for (int j = 0; j < dim; ++j) {
    bmin[j] = 1.79769313486231570815e+308;
    bmax[j] = -1.79769313486231570815e+308;
  }
  for (int v = 0; v < nv; ++v)
    for (int j = 0; j < dim; ++j) {
      double c = m->coords[v * dim + j];
      bmin[j] = (c < bmin[j]) ? c : bmin[j];
      bmax[j] = (bmax[j] < c) ? c : bmax[j];
    }
  double maxl = bmax[0] - bmin[0];
  for (int j = 1; j < dim; ++j) {
    double l = bmax[j] - bmin[j];
    maxl = (maxl < l) ? l : maxl;
  }
  for (int j = 0; j < dim; ++j) bmax[j] = bmin[j] + maxl;
  double s[3], t[3];
  for (int j = 0; j < dim; ++j) s[j] = 1.0 / (bmax[j] - bmin[j]);
  /* t = -(diag(s) * bmin), a full column-major product as in the reference */
  for (int i = 0; i < dim; ++i) {
    double acc = (i == 0 ? s[0] : 0.0) * bmin[0];
    for (int j = 1; j < dim; ++j) acc = acc + (i == j ? s[j] : 0.0) * bmin[j];
    t[i] = -acc;
  }
  int64_t* keys = (int64_t*)om_malloc(sizeof(int64_t) * (size_t)nv * dim);
  double const scale = 4503599627370496.0; /* 2^52 */
  for (int v = 0; v < nv; ++v) {
    double const* c = m->coords + v * dim;
    hcoord X[3];
    for (int i = 0; i < dim; ++i) {
      double u = (i == 0 ? s[0] : 0.0) * c[0];
      for (int j = 1; j < dim; ++j) u = u + (i == j ? s[j] : 0.0) * c[j];
      u = u + t[i];
      double z = (u < 0.0) ? 0.0 : u;
      z = (1.0 < z) ? 1.0 : z;
      X[i] = (hcoord)(z * scale);
      if (X[i] >= ((hcoord)1 << 52)) X[i] = ((hcoord)1 << 52) - 1;
    }
    axes_to_transpose(X, 52, dim);
    hcoord Y[3];
    untranspose(X, Y, 52, dim);
    for (int j = 0; j < dim; ++j) keys[(size_t)v * dim + j] = (int64_t)Y[j];
  }
  int* new2old[4];
  new2old[0] = (int*)om_malloc(sizeof(int) * (size_t)nv);
  om_stable_sort_keys(nv, dim, keys, new2old[0]);
  free(keys);
  /* entities follow the new order of their first vertex, ties by old index */
  for (int d = 1; d <= dim; ++d) {
    int n = m->nents[d];
    int const* ev = om_verts_of(m, d);
    int deg = d + 1;
    int* cnt = (int*)om_calloc((size_t)nv, sizeof(int));
    for (int e = 0; e < n; ++e) cnt[ev[e * deg]]++;
    int* off = om_offset_scan(nv, cnt);
    int* fill = (int*)om_calloc((size_t)nv, sizeof(int));
    int* byv = (int*)om_malloc(sizeof(int) * (size_t)n);
    for (int e = 0; e < n; ++e) {
      int v = ev[e * deg];
      byv[off[v] + fill[v]++] = e;
    }
    new2old[d] = (int*)om_malloc(sizeof(int) * (size_t)n);
    int k = 0;
    for (int i = 0; i < nv; ++i) {
      int ov = new2old[0][i];
      for (int j = off[ov]; j < off[ov + 1]; ++j) new2old[d][k++] = byv[j];
    }
    free(cnt);
    free(off);
    free(fill);
    free(byv);
  }
  /* unmap_mesh */
  int* old2new_low = NULL;
  for (int d = 0; d <= dim; ++d) {
    int n = m->nents[d];
    int* old2new = (int*)om_malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; ++i) old2new[new2old[d][i]] = i;
    if (d == 0) {
      double* c = (double*)om_malloc(sizeof(double) * (size_t)n * dim);
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < dim; ++j)
          c[i * dim + j] = m->coords[new2old[0][i] * dim + j];
      free(m->coords);
      m->coords = c;
    } else {
      int deg = d + 1;
      int* down = (int*)om_malloc(sizeof(int) * (size_t)n * deg);
      signed char* codes = NULL;
      if (m->down_codes[d]) codes = (signed char*)om_malloc((size_t)n * deg);
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < deg; ++j) {
          down[i * deg + j] = old2new_low[m->down[d][new2old[d][i] * deg + j]];
          if (codes) codes[i * deg + j] = m->down_codes[d][new2old[d][i] * deg + j];
        }
      free(m->down[d]);
      free(m->down_codes[d]);
      m->down[d] = down;
      m->down_codes[d] = codes;
    }
    free(old2new_low);
    old2new_low = old2new;
  }
  free(old2new_low);
  for (int d = 0; d <= dim; ++d) free(new2old[d]);
  om_mesh_clear_caches(m);
}

/* ------------------------------------------------------------ classify */
static void classify_box(om_mesh* m, double const* l, int const* nel) {
  int dim = m->dim;
  double dists[3];
  for (int i = 0; i < dim; ++i) dists[i] = l[i] / (double)(nel[i] * 32);
  for (int d = 0; d <= dim; ++d) {
    int n = m->nents[d];
    int* ids = (int*)om_malloc(sizeof(int) * (size_t)n);
    signed char* dims = (signed char*)om_malloc((size_t)n);
    int const* ev = d ? om_verts_of(m, d) : NULL;
    for (int e = 0; e < n; ++e) {
      double x[3];
      if (d == 0) {
        for (int j = 0; j < dim; ++j) x[j] = m->coords[e * dim + j];
      } else {
        for (int j = 0; j < dim; ++j) {
          double comp = 0;
          for (int k = 0; k < d + 1; ++k) comp += m->coords[ev[e * (d + 1) + k] * dim + j];
          comp /= (double)(d + 1);
          x[j] = comp;
        }
      }
      int id = 0;
      int cd = 0;
      for (int j = dim - 1; j >= 0; --j) {
        id *= 3;
        if (x[j] > (l[j] - dists[j])) {
          id += 2;
        } else if (x[j] > dists[j]) {
          id += 1;
          ++cd;
        }
      }
      ids[e] = id;
      dims[e] = (signed char)cd;
    }
    m->class_id[d] = ids;
    m->class_dim[d] = dims;
  }
}

om_mesh* om_build_from_elems(int dim, int nelems, int const* ev2v, int nverts) {
  om_mesh* m = om_mesh_new(dim);
  m->nents[0] = nverts;
  int* copy = (int*)om_malloc(sizeof(int) * (size_t)(nelems * (dim + 1) + 1));
  memcpy(copy, ev2v, sizeof(int) * (size_t)nelems * (dim + 1));
  build_from_elems(m, nelems, copy);
  return m;
}

/* build_box_internal (no Hilbert reorder, no classification),
 * src/Omega_h_build.cpp:107-134 */
om_mesh* om_build_box_internal(int dim, double x, double y, double z, int nx,
    int ny, int nz) {
  om_mesh* m = om_mesh_new(dim);
  if (dim == 2) {
    int nvx = nx + 1, nvy = ny + 1;
    int nv = nvx * nvy;
    double dx = x / nx, dy = y / ny;
    m->coords = (double*)om_malloc(sizeof(double) * (size_t)nv * 2);
    for (int v = 0; v < nv; ++v) {
      int i = v % nvx, j = v / nvx;
      m->coords[v * 2 + 0] = i * dx;
      m->coords[v * 2 + 1] = j * dy;
    }
    int nq = nx * ny;
    int* qv2v = (int*)om_malloc(sizeof(int) * (size_t)nq * 4);
    for (int q = 0; q < nq; ++q) {
      int i = q % nx, j = q / nx;
      qv2v[q * 4 + 0] = (j + 0) * nvx + (i + 0);
      qv2v[q * 4 + 1] = (j + 0) * nvx + (i + 1);
      qv2v[q * 4 + 2] = (j + 1) * nvx + (i + 1);
      qv2v[q * 4 + 3] = (j + 1) * nvx + (i + 0);
    }
    int* tv2v = tris_from_quads(nq, qv2v);
    free(qv2v);
    m->nents[0] = nv;
    build_from_elems(m, nq * 2, tv2v);
  } else {
    int nvx = nx + 1, nvy = ny + 1, nvz = nz + 1;
    int nvxy = nvx * nvy;
    int nv = nvxy * nvz;
    double dx = x / nx, dy = y / ny, dz = z / nz;
    m->coords = (double*)om_malloc(sizeof(double) * (size_t)nv * 3);
    for (int v = 0; v < nv; ++v) {
      int ij = v % nvxy, k = v / nvxy;
      int i = ij % nvx, j = ij / nvx;
      m->coords[v * 3 + 0] = i * dx;
      m->coords[v * 3 + 1] = j * dy;
      m->coords[v * 3 + 2] = k * dz;
    }
    int nxy = nx * ny;
    int nh = nxy * nz;
    int* hv2v = (int*)om_malloc(sizeof(int) * (size_t)nh * 8);
    for (int h = 0; h < nh; ++h) {
      int ij = h % nxy, k = h / nxy;
      int i = ij % nx, j = ij / nx;
      int* o = hv2v + h * 8;
      o[0] = (k + 0) * nvxy + (j + 0) * nvx + (i + 0);
      o[1] = (k + 0) * nvxy + (j + 0) * nvx + (i + 1);
      o[2] = (k + 0) * nvxy + (j + 1) * nvx + (i + 1);
      o[3] = (k + 0) * nvxy + (j + 1) * nvx + (i + 0);
      o[4] = (k + 1) * nvxy + (j + 0) * nvx + (i + 0);
      o[5] = (k + 1) * nvxy + (j + 0) * nvx + (i + 1);
      o[6] = (k + 1) * nvxy + (j + 1) * nvx + (i + 1);
      o[7] = (k + 1) * nvxy + (j + 1) * nvx + (i + 0);
    }
    int nt;
    int* tv2v = tets_from_hexes(nh, hv2v, &nt);
    free(hv2v);
    m->nents[0] = nv;
    build_from_elems(m, nt, tv2v);
  }
  return m;
}

om_mesh* om_build_box(int dim, double x, double y, double z, int nx, int ny,
    int nz) {
  om_mesh* m = om_build_box_internal(dim, x, y, z, nx, ny, nz);
  reorder_by_hilbert(m);
  double l[3] = {x, y, z};
  int nel[3] = {nx, ny, nz};
  classify_box(m, l, nel);
  return m;
}

void om_kat_hilbert(uint64_t* X, int b, int n, uint64_t* Y) {
  axes_to_transpose(X, b, n);
  untranspose(X, Y, b, n);
}

// build_box on the device: structured box, hex/quad split into simplices,
// derivation of edges and faces, Hilbert reordering and box classification.
//
//   make_2d_box / make_3d_box            src/Omega_h_box.cpp:25-102
//   tris_from_quads / tets_from_hexes    src/Omega_h_simplify.cpp:14-228
//   build_from_elems2verts               src/Omega_h_build.cpp:24-105
//   reorder_by_hilbert / unmap_mesh      src/Omega_h_reorder.cpp:10-64,
//                                        src/Omega_h_unmap_mesh.cpp:108-134
//   hilbert::from_spatial                src/Omega_h_hilbert.hpp (Skilling 2004)
//   classify_box                         src/Omega_h_box.cpp:104-160
#include "osh_mesh.hpp"

namespace osh {

namespace {

__global__ void box_coords_2d(LO nv, LO nvx, Real dx, Real dy, Real* coords) {
  LO v = thread_index();
  if (v >= nv) return;
  LO i = v % nvx, j = v / nvx;
  coords[v * 2 + 0] = i * dx;
  coords[v * 2 + 1] = j * dy;
}

__global__ void box_coords_3d(LO nv, LO nvx, LO nvxy, Real dx, Real dy, Real dz, Real* coords) {
  LO v = thread_index();
  if (v >= nv) return;
  LO ij = v % nvxy, k = v / nvxy;
  LO i = ij % nvx, j = ij / nvx;
  coords[v * 3 + 0] = i * dx;
  coords[v * 3 + 1] = j * dy;
  coords[v * 3 + 2] = k * dz;
}

// the reference's rot() moves entries with a forward loop
// (src/Omega_h_simplify.cpp:37-41); restated as written
__device__ void rot_fwd(LO* v, int n) {
  LO tmp = v[n - 1];
  for (int i = 0; i < n - 1; ++i) v[i + 1] = v[i];
  v[0] = tmp;
}
__device__ void rot_ntimes(LO* v, int nv, int ntimes) {
  for (int i = 0; i < ntimes; ++i) rot_fwd(v, nv);
}

__device__ int find_min(LO const* v, int n) {
  int min_i = 0;
  LO min_v = v[0];
  for (int i = 1; i < n; ++i)
    if (v[i] < min_v) {
      min_i = i;
      min_v = v[i];
    }
  return min_i;
}

__global__ void tris_from_quads_kernel(LO nq, LO nx, LO nvx, LO* tv2v) {
  LO q = thread_index();
  if (q >= nq) return;
  LO i = q % nx, j = q / nx;
  LO v[4] = {(j + 0) * nvx + (i + 0), (j + 0) * nvx + (i + 1), (j + 1) * nvx + (i + 1),
      (j + 1) * nvx + (i + 0)};
  int min_i = find_min(v, 4);
  rot_ntimes(v, 4, (4 - min_i) % 4);
  int const t[2][3] = {{0, 1, 2}, {2, 3, 0}};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 3; ++b) tv2v[(q * 2 + a) * 3 + b] = v[t[a][b]];
}

__device__ void hex_verts(LO h, LO nx, LO nxy, LO nvx, LO nvxy, LO* o) {
  LO ij = h % nxy, k = h / nxy;
  LO i = ij % nx, j = ij / nx;
  o[0] = (k + 0) * nvxy + (j + 0) * nvx + (i + 0);
  o[1] = (k + 0) * nvxy + (j + 0) * nvx + (i + 1);
  o[2] = (k + 0) * nvxy + (j + 1) * nvx + (i + 1);
  o[3] = (k + 0) * nvxy + (j + 1) * nvx + (i + 0);
  o[4] = (k + 1) * nvxy + (j + 0) * nvx + (i + 0);
  o[5] = (k + 1) * nvxy + (j + 0) * nvx + (i + 1);
  o[6] = (k + 1) * nvxy + (j + 1) * nvx + (i + 1);
  o[7] = (k + 1) * nvxy + (j + 1) * nvx + (i + 0);
}

__device__ void hex_prepare(LO* h, int* diags, int* ndiags) {
  int const flip_pairs[4][2] = {{0, 4}, {3, 5}, {1, 7}, {2, 6}};
  int const bur_faces[3][4] = {{6, 5, 1, 2}, {6, 2, 3, 7}, {6, 7, 4, 5}};
  int min_i = find_min(h, 8);
  rot_ntimes(h + 0, 4, (4 - min_i % 4) % 4);
  rot_ntimes(h + 4, 4, (4 - min_i % 4) % 4);
  if (min_i >= 4)
    for (int i = 0; i < 4; ++i) {
      LO t = h[flip_pairs[i][0]];
      h[flip_pairs[i][0]] = h[flip_pairs[i][1]];
      h[flip_pairs[i][1]] = t;
    }
  for (int i = 0; i < 3; ++i) {
    LO a = h[bur_faces[i][0]], b = h[bur_faces[i][2]];
    LO cc = h[bur_faces[i][1]], d = h[bur_faces[i][3]];
    diags[i] = ((a < b ? a : b) < (cc < d ? cc : d)) ? 1 : 0;
  }
  *ndiags = diags[0] + diags[1] + diags[2];
}

__device__ void hex_bur_rot(LO* h, int ntimes) {
  int const ring[6] = {1, 2, 3, 7, 5, 4};
  LO tmp[6];
  for (int i = 0; i < 6; ++i) tmp[i] = h[ring[i]];
  rot_ntimes(tmp, 6, ntimes * 2);
  for (int i = 0; i < 6; ++i) h[ring[i]] = tmp[i];
}

__global__ void hex_degrees(LO nh, LO nx, LO nxy, LO nvx, LO nvxy, LO* deg) {
  LO hh = thread_index();
  if (hh >= nh) return;
  LO h[8];
  int d[3], nd;
  hex_verts(hh, nx, nxy, nvx, nvxy, h);
  hex_prepare(h, d, &nd);
  deg[hh] = (nd == 0) ? 5 : 6;
}

__global__ void tets_from_hexes_kernel(LO nh, LO nx, LO nxy, LO nvx, LO nvxy,
    LO const* h2ht, LO* tv2v) {
  LO hh = thread_index();
  if (hh >= nh) return;
  int const t0[5][4] = {{0, 1, 2, 5}, {0, 2, 7, 5}, {0, 2, 3, 7}, {0, 5, 7, 4}, {2, 7, 5, 6}};
  int const t1[6][4] = {{0, 5, 7, 4}, {0, 1, 7, 5}, {1, 6, 7, 5}, {0, 7, 2, 3}, {0, 7, 1, 2},
      {1, 7, 6, 2}};
  int const t2[6][4] = {{0, 4, 5, 6}, {0, 3, 7, 6}, {0, 7, 4, 6}, {0, 1, 2, 5}, {0, 3, 6, 2},
      {0, 6, 5, 2}};
  int const t3[6][4] = {{0, 2, 3, 6}, {0, 3, 7, 6}, {0, 7, 4, 6}, {0, 5, 6, 4}, {1, 5, 6, 0},
      {1, 6, 2, 0}};
  LO h[8];
  int d[3], nd;
  hex_verts(hh, nx, nxy, nvx, nvxy, h);
  hex_prepare(h, d, &nd);
  int const(*tmpl)[4] = t3;
  int n = 6;
  if (nd == 0) {
    tmpl = t0;
    n = 5;
  } else if (nd == 1) {
    int face = -1;
    for (int i = 0; i < 3; ++i)
      if (d[i]) face = i;
    hex_bur_rot(h, (3 - face) % 3);
    tmpl = t1;
  } else if (nd == 2) {
    int face = -1;
    for (int i = 0; i < 3; ++i)
      if (!d[i]) face = i;
    hex_bur_rot(h, (3 - face) % 3);
    tmpl = t2;
  }
  LO t = h2ht[hh];
  for (int i = 0; i < n; ++i, ++t)
    for (int j = 0; j < 4; ++j) tv2v[t * 4 + j] = h[tmpl[i][j]];
}

// ---- Hilbert keys, Skilling's AxestoTranspose + untranspose ----
__device__ void axes_to_transpose(std::uint64_t* X, int b, int n) {
  std::uint64_t M = std::uint64_t(1) << (b - 1), P, Q, t;
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

template <int D>
__global__ void hilbert_keys(LO nv, Real const* coords, Real const* s, Real const* t, I64* keys) {
  LO v = thread_index();
  if (v >= nv) return;
  std::uint64_t X[D];
  for (int i = 0; i < D; ++i) {
    double u = (i == 0 ? s[0] : 0.0) * coords[v * D + 0];
    for (int j = 1; j < D; ++j) u = u + (i == j ? s[j] : 0.0) * coords[v * D + j];
    u = u + t[i];
    double z = (u < 0.0) ? 0.0 : u;
    z = (1.0 < z) ? 1.0 : z;
    X[i] = std::uint64_t(z * 4503599627370496.0);
    if (X[i] >= (std::uint64_t(1) << 52)) X[i] = (std::uint64_t(1) << 52) - 1;
  }
  axes_to_transpose(X, 52, D);
  std::uint64_t Y[D];
  for (int i = 0; i < D; ++i) Y[i] = 0;
  for (int i = 0; i < 52 * D; ++i)
    Y[i / 52] |= (((X[i % D] >> (52 - 1 - (i / D))) & 1) << (52 - 1 - (i % 52)));
  for (int j = 0; j < D; ++j) keys[I64(v) * D + j] = I64(Y[j]);
}

__global__ void coord_column(LO nv, int dim, int j, Real const* coords, Real* out) {
  LO v = thread_index();
  if (v < nv) out[v] = coords[v * dim + j];
}

__global__ void first_vertex_rank(LO n, int deg, LO const* ev2v, LO const* old2new_v, I64* keys) {
  LO e = thread_index();
  if (e < n) keys[e] = old2new_v[ev2v[e * deg]];
}

__global__ void invert_perm(LO n, LO const* new2old, LO* old2new) {
  LO i = thread_index();
  if (i < n) old2new[new2old[i]] = i;
}

__global__ void permute_coords(LO n, int dim, LO const* new2old, Real const* in, Real* out) {
  LO i = thread_index();
  if (i >= n) return;
  for (int j = 0; j < dim; ++j) out[i * dim + j] = in[new2old[i] * dim + j];
}

__global__ void unmap_down_kernel(LO n, int deg, LO const* new2old, LO const* old_down,
    I8 const* old_codes, LO const* old_lows2new, LO* down, I8* codes) {
  LO i = thread_index();
  if (i >= n) return;
  LO o = new2old[i];
  for (int j = 0; j < deg; ++j) {
    down[i * deg + j] = old_lows2new[old_down[o * deg + j]];
    if (codes) codes[i * deg + j] = old_codes[o * deg + j];
  }
}

template <int D>
__global__ void classify_kernel(LO n, int deg, LO const* ev2v, Real const* coords,
    Real l0, Real l1, Real l2, Real d0, Real d1, Real d2, LO* ids, I8* dims) {
  LO e = thread_index();
  if (e >= n) return;
  double x[3];
  for (int j = 0; j < D; ++j) {
    if (deg == 1) {
      x[j] = coords[e * D + j];
    } else {
      double comp = 0;
      for (int k = 0; k < deg; ++k) comp += coords[ev2v[e * deg + k] * D + j];
      comp /= double(deg);
      x[j] = comp;
    }
  }
  double l[3] = {l0, l1, l2};
  double dists[3] = {d0, d1, d2};
  int id = 0, cd = 0;
  for (int j = D - 1; j >= 0; --j) {
    id *= 3;
    if (x[j] > (l[j] - dists[j])) {
      id += 2;
    } else if (x[j] > dists[j]) {
      id += 1;
      ++cd;
    }
  }
  ids[e] = id;
  dims[e] = I8(cd);
}

void add_ents(Ctx& c, DMesh& m, int d, LO n, LOs const& ev2v) {
  if (d == 1) {
    m.down[1] = ev2v;
    m.nents[1] = n;
    return;
  }
  LOs down;
  Bytes codes;
  reflect_down(c, m, d, n, ev2v, down, codes);
  m.down[d] = down;
  m.down_codes[d] = codes;
  m.nents[d] = n;
}

void reorder_by_hilbert(Ctx& c, DMesh& m) {
  int dim = m.dim;
  LO nv = m.nents[0];
  std::vector<Real> bmin(dim), bmax(dim);
  Reals col(c, nv);
  for (int j = 0; j < dim; ++j) {
    OSH_LAUNCH(coord_column, nv, c, dim, j, m.coords.data(), col.data());
    bmin[j] = min_real(c, col);
    bmax[j] = max_real(c, col);
  }
  Real maxl = bmax[0] - bmin[0];
  for (int j = 1; j < dim; ++j) {
    Real l = bmax[j] - bmin[j];
    maxl = (maxl < l) ? l : maxl;
  }
  for (int j = 0; j < dim; ++j) bmax[j] = bmin[j] + maxl;
  std::vector<Real> s(dim), t(dim);
  for (int j = 0; j < dim; ++j) s[j] = 1.0 / (bmax[j] - bmin[j]);
  for (int i = 0; i < dim; ++i) {
    Real acc = (i == 0 ? s[0] : 0.0) * bmin[0];
    for (int j = 1; j < dim; ++j) acc = acc + (i == j ? s[j] : 0.0) * bmin[j];
    t[i] = -acc;
  }
  Reals ds = to_device(c, s), dt = to_device(c, t);
  DArray<I64> keys(c, nv * dim);
  if (dim == 3)
    OSH_LAUNCH(hilbert_keys<3>, nv, c, m.coords.data(), ds.data(), dt.data(), keys.data());
  else
    OSH_LAUNCH(hilbert_keys<2>, nv, c, m.coords.data(), ds.data(), dt.data(), keys.data());
  LOs new2old[4];
  new2old[0] = sort_by_keys(c, keys, dim);
  LOs old2new_v(c, nv);
  OSH_LAUNCH(invert_perm, nv, c, new2old[0].data(), old2new_v.data());
  for (int d = 1; d <= dim; ++d) {
    LO n = m.nents[d];
    LOs const& ev = verts_of(c, m, d);
    DArray<I64> k(c, n);
    OSH_LAUNCH(first_vertex_rank, n, c, d + 1, ev.data(), old2new_v.data(), k.data());
    new2old[d] = sort_by_keys(c, k, 1);
  }
  Reals coords(c, nv * dim);
  OSH_LAUNCH(permute_coords, nv, c, dim, new2old[0].data(), m.coords.data(), coords.data());
  m.coords = coords;
  LOs old2new_low = old2new_v;
  for (int d = 1; d <= dim; ++d) {
    LO n = m.nents[d];
    int deg = d + 1;
    LOs down(c, n * deg);
    Bytes codes;
    if (d >= 2) codes = Bytes(c, n * deg);
    OSH_LAUNCH(unmap_down_kernel, n, c, deg, new2old[d].data(), m.down[d].data(),
        d >= 2 ? m.down_codes[d].data() : (I8 const*)nullptr, old2new_low.data(),
        down.data(), d >= 2 ? codes.data() : (I8*)nullptr);
    LOs old2new(c, n);
    OSH_LAUNCH(invert_perm, n, c, new2old[d].data(), old2new.data());
    m.down[d] = down;
    m.down_codes[d] = codes;
    old2new_low = old2new;
  }
  m.clear_caches();
}

void classify_box(Ctx& c, DMesh& m, Real const* l, LO const* nel) {
  int dim = m.dim;
  Real dists[3] = {0, 0, 0};
  for (int i = 0; i < dim; ++i) dists[i] = l[i] / double(nel[i] * 32);
  for (int d = 0; d <= dim; ++d) {
    LO n = m.nents[d];
    m.class_id[d] = LOs(c, n);
    m.class_dim[d] = Bytes(c, n);
    LO const* ev = d ? verts_of(c, m, d).data() : nullptr;
    if (dim == 3)
      OSH_LAUNCH(classify_kernel<3>, n, c, d + 1, ev, m.coords.data(), l[0], l[1], l[2],
          dists[0], dists[1], dists[2], m.class_id[d].data(), m.class_dim[d].data());
    else
      OSH_LAUNCH(classify_kernel<2>, n, c, d + 1, ev, m.coords.data(), l[0], l[1], l[2],
          dists[0], dists[1], dists[2], m.class_id[d].data(), m.class_dim[d].data());
  }
}

}  // namespace

void build_from_elems(Ctx& c, DMesh& m, int dim, LO nelems, LOs const& ev2v, LO nverts) {
  m.dim = dim;
  m.nents[0] = nverts;
  for (int md = 1; md < dim; ++md) {
    LOs mv2v = find_unique(c, nelems, dim, md, ev2v);
    add_ents(c, m, md, mv2v.size() / (md + 1), mv2v);
  }
  add_ents(c, m, dim, nelems, ev2v);
}

void build_box(Ctx& c, DMesh& m, int dim, Real x, Real y, Real z, LO nx, LO ny, LO nz,
    bool reorder_and_classify) {
  m = DMesh();
  m.dim = dim;
  if (dim == 2) {
    LO nvx = nx + 1, nvy = ny + 1, nv = nvx * nvy;
    m.coords = Reals(c, nv * 2);
    OSH_LAUNCH(box_coords_2d, nv, c, nvx, x / nx, y / ny, m.coords.data());
    LO nq = nx * ny;
    LOs tv2v(c, nq * 6);
    OSH_LAUNCH(tris_from_quads_kernel, nq, c, nx, nvx, tv2v.data());
    build_from_elems(c, m, 2, nq * 2, tv2v, nv);
  } else {
    LO nvx = nx + 1, nvy = ny + 1, nvz = nz + 1;
    LO nvxy = nvx * nvy, nv = nvxy * nvz;
    m.coords = Reals(c, nv * 3);
    OSH_LAUNCH(box_coords_3d, nv, c, nvx, nvxy, x / nx, y / ny, z / nz, m.coords.data());
    LO nxy = nx * ny, nh = nxy * nz;
    LOs deg(c, nh);
    OSH_LAUNCH(hex_degrees, nh, c, nx, nxy, nvx, nvxy, deg.data());
    LO nt;
    LOs h2ht = offset_scan(c, deg, &nt);
    LOs tv2v(c, nt * 4);
    OSH_LAUNCH(tets_from_hexes_kernel, nh, c, nx, nxy, nvx, nvxy, h2ht.data(), tv2v.data());
    build_from_elems(c, m, 3, nt, tv2v, nv);
  }
  if (reorder_and_classify) {
    reorder_by_hilbert(c, m);
    Real l[3] = {x, y, z};
    LO nel[3] = {nx, ny, nz};
    classify_box(c, m, l, nel);
  }
}

}  // namespace osh

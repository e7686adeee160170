// Conservative and momentum transfer through adapt() on the device.
//
// Element densities whose integral is conserved (OMEGA_H_CONSERVE) and
// velocities whose momentum is conserved (OMEGA_H_MOMENTUM_VELOCITY) carry an
// element array of integral errors through the passes of one adapt():
//
//   setup_conservation_tags              src/Omega_h_conserve.cpp:578-606
//   cavities by boundary status and the
//   sub-cavities of one class_id each    src/Omega_h_conserve.cpp:65-181
//   carry_class_bdry_integ_error         src/Omega_h_conserve.cpp:183-207
//   introduce_class_integ_error          src/Omega_h_conserve.cpp:209-243
//   transfer_integ_error, density and
//   momentum errors per operation        src/Omega_h_conserve.cpp:245-336,
//                                        372-389,460-537
//   transfer_by_intersection             src/Omega_h_conserve.cpp:391-455
//   diffusion and correct_integral_errors src/Omega_h_conserve.cpp:608-829
//
// A cavity is a key's old elements (the key's upward row, in row order) and
// its products (in product order).  Cavities whose key is classified below
// the mesh dimension split into sub-cavities of one element class each, in
// the order in which a class first appears among the old elements; products
// of a class that no old element has form a last sub-cavity without old
// elements.  Every sum starts from 0.0 and runs in list order, so each value
// is bit-identical to a serial evaluation of the same formulas.
//
// Layout: one thread per key for the error sums (cavities hold tens of
// elements) and one thread per product for the intersection densities; the
// clipped polytope of one product lives in that thread's private memory
// (32 vertices with their neighbours).  The diffusion is a Jacobi sweep per
// launch over the element graph across interior sides, with the "all
// errors within tolerance" test folded into the sweep.
#include "osh_math.hpp"
#include "osh_passes.hpp"

namespace osh {
namespace {

enum { kNotBdry = 0, kTouchBdry = 1, kKeyBdry = 2 };

OSH_HD double min2(double a, double b) { return (b < a) ? b : a; }
OSH_HD double max2(double a, double b) { return (a < b) ? b : a; }

// ------------------------------------------------------------ intersection
// The size of the intersection of two simplices: the first is a polytope
// clipped by the planes of the second's sides, then measured (src/r3d.hpp:
// init :459-495, faces_from_verts :509-560, clip :373-441 with ClipHelper
// :277-361, reduce<0> :629-735,776-852, intersect_simplices :868-886).
// A polytope is a vertex buffer in which every vertex lists its D
// neighbours.  A tetrahedron clipped by four planes stays convex with at most
// eight faces, hence at most twelve vertices, plus at most eighteen created
// while one plane is applied: 32 vertices hold every case.
constexpr int kPolyVerts = 32;

template <int D>
struct Poly {
  double pos[kPolyVerts][D];
  int nbr[kPolyVerts][D];
  int n;
};

struct Plane {
  double n[3];
  double d;
};

template <int D>
__device__ __forceinline__ double pdot(double const* a, double const* b) {
  double c = a[0] * b[0];
  for (int i = 1; i < D; ++i) c += a[i] * b[i];
  return c;
}

template <int D>
__device__ __forceinline__ void pnormalize(double* v) {
  double l = math::sqrt_(pdot<D>(v, v));
  for (int i = 0; i < D; ++i) v[i] = v[i] / l;
}

// the plane through a, b, c with unit normal (b - a) x (c - a), placed
// through the triangle's centroid
__device__ void face_plane(double const* a, double const* b, double const* c, Plane& f) {
  double ctr[3], u[3], w[3];
  for (int i = 0; i < 3; ++i) {
    ctr[i] = (a[i] + b[i] + c[i]) * (1.0 / 3.0);
    u[i] = b[i] - a[i];
    w[i] = c[i] - a[i];
  }
  f.n[0] = u[1] * w[2] - u[2] * w[1];
  f.n[1] = u[2] * w[0] - u[0] * w[2];
  f.n[2] = u[0] * w[1] - u[1] * w[0];
  pnormalize<3>(f.n);
  f.d = -pdot<3>(f.n, ctr);
}

template <int D>
__device__ void side_planes(double const (*p)[D], Plane* f);

template <>
__device__ void side_planes<3>(double const (*p)[3], Plane* f) {
  face_plane(p[3], p[2], p[1], f[0]);
  face_plane(p[0], p[2], p[3], f[1]);
  face_plane(p[0], p[3], p[1], f[2]);
  face_plane(p[0], p[1], p[2], f[3]);
}

template <>
__device__ void side_planes<2>(double const (*p)[2], Plane* f) {
  for (int k = 0; k < 3; ++k) {
    double const* p0 = p[k];
    double const* p1 = p[(k + 1) % 3];
    f[k].n[0] = p0[1] - p1[1];
    f[k].n[1] = p1[0] - p0[0];
    pnormalize<2>(f[k].n);
    f[k].d = -pdot<2>(f[k].n, p0);
  }
}

template <int D>
__device__ void poly_init(Poly<D>& poly, double const (*p)[D]) {
  if constexpr (D == 3) {
    // each vertex's neighbours, counterclockwise seen from outside
    constexpr int nb[4][3] = {{1, 3, 2}, {2, 3, 0}, {0, 3, 1}, {1, 2, 0}};
    poly.n = 4;
    for (int v = 0; v < 4; ++v)
      for (int k = 0; k < D; ++k) {
        poly.nbr[v][k] = nb[v][k];
        poly.pos[v][k] = p[v][k];
      }
  } else {
    poly.n = 3;
    for (int v = 0; v < 3; ++v) {
      poly.pos[v][0] = p[v][0];
      poly.pos[v][1] = p[v][1];
      poly.nbr[v][0] = (v + 1) % 3;
      poly.nbr[v][1] = (3 + v - 1) % 3;
    }
  }
}

// clips the polytope by the D + 1 planes; false if the buffer would overflow
// (or the neighbour lists do not describe a polytope: the walks are bounded)
template <int D>
__device__ bool poly_clip(Poly<D>& poly, Plane const* planes) {
  if (poly.n <= 0) return true;
  double sd[kPolyVerts];
  int clipped[kPolyVerts];
  for (int p = 0; p <= D; ++p) {
    int onv = poly.n;
    double smin = 1.79769313486231570815e+308, smax = -1.79769313486231570815e+308;
    for (int v = 0; v < kPolyVerts; ++v) clipped[v] = 0;
    for (int v = 0; v < onv; ++v) {
      sd[v] = planes[p].d + pdot<D>(poly.pos[v], planes[p].n);
      if (sd[v] < smin) smin = sd[v];
      if (sd[v] > smax) smax = sd[v];
      if (sd[v] < 0.0) clipped[v] = 1;
    }
    if (smin >= 0.0) continue;
    if (smax <= 0.0) {
      poly.n = 0;
      return true;
    }
    // a new vertex on every edge from a kept vertex to a clipped one
    for (int vcur = 0; vcur < onv; ++vcur) {
      if (clipped[vcur]) continue;
      for (int np = 0; np < D; ++np) {
        int vnext = poly.nbr[vcur][np];
        if (!clipped[vnext]) continue;
        if (poly.n >= kPolyVerts) return false;
        int nv = poly.n;
        if constexpr (D == 3) {
          poly.nbr[nv][0] = vcur;
        } else {
          poly.nbr[nv][1 - np] = vcur;
          poly.nbr[nv][np] = -1;
        }
        poly.nbr[vcur][np] = nv;
        double wa = -sd[vnext], wb = sd[vcur];
        for (int i = 0; i < D; ++i)
          poly.pos[nv][i] = (wa * poly.pos[vcur][i] + wb * poly.pos[vnext][i]) / (wa + wb);
        ++poly.n;
      }
    }
    // link the new vertices around the cut face (3D) or segment (2D)
    if constexpr (D == 3) {
      for (int vstart = onv; vstart < poly.n; ++vstart) {
        int vcur = vstart;
        int vnext = poly.nbr[vcur][0];
        int steps = 0;
        do {
          int np;
          for (np = 0; np < 3; ++np)
            if (poly.nbr[vnext][np] == vcur) break;
          if (np == 3 || ++steps > 3 * kPolyVerts) return false;  // not a polytope
          vcur = vnext;
          vnext = poly.nbr[vcur][(np + 1) % 3];
        } while (vcur < onv);
        poly.nbr[vstart][2] = vcur;
        poly.nbr[vcur][1] = vstart;
      }
    } else {
      for (int vstart = onv; vstart < poly.n; ++vstart) {
        if (poly.nbr[vstart][1] >= 0) continue;
        int vcur = poly.nbr[vstart][0];
        int steps = 0;
        do {
          if (vcur < 0 || ++steps > kPolyVerts) return false;  // not a polygon
          vcur = poly.nbr[vcur][0];
        } while (vcur < onv);
        poly.nbr[vstart][1] = vcur;
        poly.nbr[vcur][0] = vstart;
      }
    }
    // drop the clipped vertices
    int nun = 0;
    for (int v = 0; v < poly.n; ++v) {
      if (!clipped[v]) {
        for (int i = 0; i < D; ++i) {
          poly.pos[nun][i] = poly.pos[v][i];
          poly.nbr[nun][i] = poly.nbr[v][i];
        }
        clipped[v] = nun++;
      }
    }
    poly.n = nun;
    for (int v = 0; v < poly.n; ++v)
      for (int np = 0; np < D; ++np) {
        int w = poly.nbr[v][np];
        if (w < 0 || w >= kPolyVerts) return false;  // a neighbour left unlinked
        poly.nbr[v][np] = clipped[w];
      }
  }
  return true;
}

// the polytope's volume (3D: a fan of triangles over every face, each with
// the origin) or area (2D: the shoelace sum over its edges)
template <int D>
__device__ double poly_measure(Poly<D> const& poly, bool* ok) {
  double m0 = 0.0;
  if (poly.n > 0) {
    if constexpr (D == 3) {
      unsigned char emarks[kPolyVerts][3];
      for (int v = 0; v < kPolyVerts; ++v)
        for (int k = 0; k < 3; ++k) emarks[v][k] = 0;
      for (int vstart = 0; vstart < poly.n; ++vstart) {
        for (int pstart = 0; pstart < 3; ++pstart) {
          if (emarks[vstart][pstart]) continue;
          int pnext = pstart;
          int vcur = vstart;
          emarks[vcur][pnext] = 1;
          int vnext = poly.nbr[vcur][pnext];
          double const* v0 = poly.pos[vcur];
          int np;
          for (np = 0; np < 3; ++np)
            if (poly.nbr[vnext][np] == vcur) break;
          vcur = vnext;
          pnext = (np + 1) % 3;
          emarks[vcur][pnext] = 1;
          vnext = poly.nbr[vcur][pnext];
          int steps = 0;
          while (vnext != vstart && ++steps <= 3 * kPolyVerts) {
            double const* v2 = poly.pos[vcur];
            double const* v1 = poly.pos[vnext];
            double sixv = (-v2[0] * v1[1] * v0[2] + v1[0] * v2[1] * v0[2] +
                           v2[0] * v0[1] * v1[2] - v0[0] * v2[1] * v1[2] -
                           v1[0] * v0[1] * v2[2] + v0[0] * v1[1] * v2[2]);
            m0 += (1.0 / 6.0) * sixv;
            for (np = 0; np < 3; ++np)
              if (poly.nbr[vnext][np] == vcur) break;
            vcur = vnext;
            pnext = (np + 1) % 3;
            emarks[vcur][pnext] = 1;
            vnext = poly.nbr[vcur][pnext];
          }
          if (vnext != vstart) *ok = false;  // a face that does not close
        }
      }
    } else {
      for (int vcur = 0; vcur < poly.n; ++vcur) {
        int vnext = poly.nbr[vcur][0];
        double const* v0 = poly.pos[vcur];
        double const* v1 = poly.pos[vnext];
        double twoa = (v0[0] * v1[1] - v0[1] * v1[0]);
        m0 += 0.5 * twoa;
      }
    }
  }
  double result = 0;
  result += m0 * 1.0;
  return result;
}

// the size of (simplex a) ∩ (simplex b); false if the buffer overflowed
template <int D>
__device__ bool intersect_size(double const (*a)[D], double const (*b)[D], double* out) {
  Poly<D> poly;
  Plane f[D + 1];
  poly_init<D>(poly, a);
  side_planes<D>(b, f);
  if (!poly_clip<D>(poly, f)) return false;
  bool ok = true;
  *out = poly_measure<D>(poly, &ok);
  return ok;
}

template <int D>
__device__ __forceinline__ void elem_points(LO const* ev2v, Real const* coords, LO e,
    double (*p)[D]) {
  for (int k = 0; k <= D; ++k) {
    LO v = ev2v[std::size_t(e) * (D + 1) + k];
    for (int x = 0; x < D; ++x) p[k][x] = coords[std::size_t(v) * D + x];
  }
}

// the key of product p: the last key whose products start at or before p
__device__ __forceinline__ LO key_of(LO p, LO nkeys, LO const* k2p) {
  LO lo = 0, hi = nkeys - 1;
  while (lo < hi) {
    LO mid = lo + (hi - lo + 1) / 2;
    if (k2p[mid] <= p)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// ------------------------------------------------------------- cavity kinds
// NOT_BDRY: no old element has a vertex classified below the mesh
// dimension; KEY_BDRY: the key is classified below it; TOUCH_BDRY: the rest
// (separate_cavities, get_elems_are_bdry src/Omega_h_conserve.cpp:65-131)
template <int D>
__global__ void cavity_kinds(LO nkeys, LO const* keys2kds, I8 const* key_class_dim,
    LO const* up_a2ab, LO const* up_ab2b, LO const* ev2v, I8 const* vert_class_dim,
    I8* kinds) {
  LO k = thread_index();
  if (k >= nkeys) return;
  LO kd = keys2kds[k];
  if (key_class_dim[kd] < D) {
    kinds[k] = kKeyBdry;
    return;
  }
  I8 kind = kNotBdry;
  for (LO i = up_a2ab[kd]; i < up_a2ab[kd + 1]; ++i) {
    LO e = up_ab2b[i];
    for (int j = 0; j <= D; ++j)
      if (vert_class_dim[ev2v[std::size_t(e) * (D + 1) + j]] != D) kind = kTouchBdry;
  }
  kinds[k] = kind;
}

// transfer_by_intersection_dim: every product of a cavity takes the
// intersection-weighted mean of the old elements of its (sub-)cavity; with
// by_class, a key-boundary cavity's old elements are those of the product's
// class.  One thread per product.
template <int D>
__global__ void intersect_densities(LO nprods, LO nkeys, LO const* k2p, LO const* keys2kds,
    LO const* up_a2ab, LO const* up_ab2b, I8 const* kinds, int by_class, LO const* prods2new,
    LO const* old_ev2v, Real const* old_coords, LO const* new_ev2v, Real const* new_coords,
    LO const* old_cls, LO const* new_cls, int nc, Real const* odata, Real* ndata, int* error) {
  LO p = thread_index();
  if (p >= nprods) return;
  LO k = key_of(p, nkeys, k2p);
  LO kd = keys2kds[k];
  LO ne = prods2new[p];
  bool filter = by_class && kinds[k] == kKeyBdry;
  LO color = new_cls[ne];
  double npts[D + 1][D], opts[D + 1][D];
  elem_points<D>(new_ev2v, new_coords, ne, npts);
  Real* out = ndata + std::size_t(ne) * nc;
  for (int q = 0; q < nc; ++q) out[q] = 0;
  double total = 0.0;
  for (LO i = up_a2ab[kd]; i < up_a2ab[kd + 1]; ++i) {
    LO oe = up_ab2b[i];
    if (filter && old_cls[oe] != color) continue;
    elem_points<D>(old_ev2v, old_coords, oe, opts);
    double isz = 0.0;
    if (!intersect_size<D>(npts, opts, &isz)) {
      *error = kPolytopeOverflow;
      return;
    }
    for (int q = 0; q < nc; ++q) out[q] += isz * odata[std::size_t(oe) * nc + q];
    total += isz;
  }
  for (int q = 0; q < nc; ++q) out[q] /= total;
}

// ------------------------------------------------------------ error sums
// the conserved quantity of element e, component q: a density, or the
// element-mean velocity times the element density (a momentum density)
template <int D, bool MOMENTUM>
struct Quantity {
  Real const* data;   // densities (nc per element) or velocities (nc per vertex)
  Real const* rho;    // momentum: the element densities
  LO const* ev2v;     // momentum: element vertices
  int nc;
  __device__ __forceinline__ double operator()(LO e, int q) const {
    if (!MOMENTUM) return data[std::size_t(e) * nc + q];
    double comp = 0;
    for (int k = 0; k <= D; ++k) comp += data[std::size_t(ev2v[std::size_t(e) * (D + 1) + k]) * nc + q];
    comp /= (D + 1);
    return comp * rho[e];
  }
};

// the error densities of one sub-cavity added into the new elements' errors:
// the error it introduces (new minus old integral) when `intro`, then the
// old elements' error it carries when `carry`, each spread over the new
// elements in proportion to their sizes (introduce_class_integ_error,
// carry_class_bdry_integ_error)
template <int D, bool MOMENTUM>
__device__ void subcavity_errors(LO ob, LO oe_end, LO const* olist, LO pb, LO pe,
    LO const* prods2new, bool filter, LO color, LO const* old_cls, LO const* new_cls,
    bool any_old_color, bool intro, bool carry, Quantity<D, MOMENTUM> const& oq,
    Quantity<D, MOMENTUM> const& nq, Real const* oerr, Real const* osize, Real const* nsize,
    int nc, Real* err, int* error) {
  // a product belongs here when it has the color (or, in the last
  // sub-cavity of a colored cavity, when no old element has its color)
  auto in_new = [&](LO ne) {
    if (!filter) return true;
    if (any_old_color) return new_cls[ne] == color;
    for (LO i = ob; i < oe_end; ++i)
      if (old_cls[olist[i]] == new_cls[ne]) return false;
    return true;
  };
  auto in_old = [&](LO e) { return !filter || (any_old_color && old_cls[e] == color); };
  double nsum = 0.0;
  for (LO p = pb; p < pe; ++p) {
    LO ne = prods2new[p];
    if (in_new(ne)) nsum = nsum + nsize[ne];
  }
  if (intro) {
    for (int q = 0; q < nc; ++q) {
      double oint = 0.0, nint = 0.0;
      for (LO i = ob; i < oe_end; ++i) {
        LO e = olist[i];
        if (in_old(e)) oint = oint + oq(e, q) * osize[e];
      }
      for (LO p = pb; p < pe; ++p) {
        LO ne = prods2new[p];
        if (in_new(ne)) nint = nint + nq(ne, q) * nsize[ne];
      }
      double cerr = nint - oint;
      double d = 0.0;
      if (nsum != 0.0)
        d = cerr / nsum;
      else if (cerr != 0.0)
        *error = kZeroSizeCavity;
      for (LO p = pb; p < pe; ++p) {
        LO ne = prods2new[p];
        if (in_new(ne)) err[std::size_t(ne) * nc + q] += d * nsize[ne];
      }
    }
  }
  if (carry) {
    for (int q = 0; q < nc; ++q) {
      double oe = 0.0;
      for (LO i = ob; i < oe_end; ++i) {
        LO e = olist[i];
        if (in_old(e)) oe = oe + oerr[std::size_t(e) * nc + q];
      }
      double d = 0.0;
      if (nsum != 0.0)
        d = oe / nsum;
      else if (oe != 0.0)
        *error = kZeroSizeCavity;
      for (LO p = pb; p < pe; ++p) {
        LO ne = prods2new[p];
        if (in_new(ne)) err[std::size_t(ne) * nc + q] += d * nsize[ne];
      }
    }
  }
}

// transfer_integ_error for one key: `intro_mask` and `carry_mask` hold a bit
// per cavity kind; key-boundary cavities split by class
template <int D, bool MOMENTUM>
__global__ void cavity_errors(LO nkeys, LO const* keys2kds, LO const* up_a2ab,
    LO const* up_ab2b, LO const* k2p, LO const* prods2new, I8 const* kinds, LO const* old_cls,
    LO const* new_cls, int intro_mask, int carry_mask, Quantity<D, MOMENTUM> oq,
    Quantity<D, MOMENTUM> nq, Real const* oerr, Real const* osize, Real const* nsize, int nc,
    Real* err, int* error) {
  LO k = thread_index();
  if (k >= nkeys) return;
  int kind = kinds[k];
  bool intro = (intro_mask >> kind) & 1;
  bool carry = (carry_mask >> kind) & 1;
  if (!intro && !carry) return;
  LO kd = keys2kds[k];
  LO ob = up_a2ab[kd], oe = up_a2ab[kd + 1];
  LO pb = k2p[k], pe = k2p[k + 1];
  if (kind != kKeyBdry) {
    subcavity_errors<D, MOMENTUM>(ob, oe, up_ab2b, pb, pe, prods2new, false, 0, old_cls,
        new_cls, true, intro, carry, oq, nq, oerr, osize, nsize, nc, err, error);
    return;
  }
  // one sub-cavity per class, in the order classes first appear among the
  // old elements
  for (LO i = ob; i < oe; ++i) {
    LO color = old_cls[up_ab2b[i]];
    bool seen = false;
    for (LO j = ob; j < i; ++j) seen = seen || old_cls[up_ab2b[j]] == color;
    if (seen) continue;
    subcavity_errors<D, MOMENTUM>(ob, oe, up_ab2b, pb, pe, prods2new, true, color, old_cls,
        new_cls, true, intro, carry, oq, nq, oerr, osize, nsize, nc, err, error);
  }
  // products of classes no old element has
  bool rest = false;
  for (LO p = pb; p < pe && !rest; ++p) {
    LO c = new_cls[prods2new[p]];
    bool found = false;
    for (LO i = ob; i < oe; ++i) found = found || old_cls[up_ab2b[i]] == c;
    rest = !found;
  }
  if (rest)
    subcavity_errors<D, MOMENTUM>(ob, oe, up_ab2b, pb, pe, prods2new, true, 0, old_cls,
        new_cls, false, intro, carry, oq, nq, oerr, osize, nsize, nc, err, error);
}

// the new mesh's error array: same elements keep theirs, products start at 0
__global__ void errors_start(LO n, LO nsame, int nc, LO const* same2old, LO const* same2new,
    LO nprods, LO const* prods2new, Real const* oerr, Real* err) {
  LO i = thread_index();
  if (i >= n) return;
  if (i < nsame) {
    for (int q = 0; q < nc; ++q)
      err[std::size_t(same2new[i]) * nc + q] = oerr[std::size_t(same2old[i]) * nc + q];
  } else {
    LO ne = prods2new[i - nsame];
    for (int q = 0; q < nc; ++q) err[std::size_t(ne) * nc + q] = 0.0;
  }
}

// ------------------------------------------------------------- correction
// the element graph across sides classified in the interior
// (get_elem_diffusion_graph): slot k of element e holds the element across
// its side k, or -1; slots keep side order
template <int D>
__global__ void diffusion_graph(LO ne, LO const* es2s, I8 const* side_class_dim,
    unsigned long long const* side_elems, LO* nbrs) {
  LO e = thread_index();
  if (e >= ne) return;
  for (int k = 0; k <= D; ++k) {
    LO s = es2s[std::size_t(e) * (D + 1) + k];
    LO o = -1;
    if (side_class_dim[s] == D) o = LO(unsigned(side_elems[s] & 0xffffffffull) - unsigned(e));
    nbrs[std::size_t(e) * (D + 1) + k] = o;
  }
}

// weighted sizes max(|quantity|, floor) and weighted error densities of
// component q (diffuse_integrals_weighted).  flags[0] becomes `tag` when
// some density is above tol; flags[1] becomes 1 when a nonzero error sits
// over a zero weight (plain stores of one value: no atomics needed)
__global__ void diffusion_start(LO ne, int nc, int q, Real const* quantity, Real const* err,
    double floor, double tol, Real* ws, Real* wd, int* flags, int tag) {
  LO e = thread_index();
  if (e >= ne) return;
  double w = max2(fabs(quantity[std::size_t(e) * nc + q]), floor);
  double x = err[std::size_t(e) * nc + q];
  double d = 0.0;
  if (w != 0.0)
    d = x / w;
  else if (x != 0.0)
    flags[1] = 1;
  ws[e] = w;
  wd[e] = d;
  if (!(fabs(d) <= tol)) flags[0] = tag;
}

// diffuse_densities_once, plus the bound test of its output
template <int D>
__global__ void diffusion_sweep(LO ne, LO const* nbrs, Real const* sizes, Real const* dens,
    double tol, Real* out, int* flags, int tag) {
  LO e = thread_index();
  if (e >= ne) return;
  double s = sizes[e];
  double o = dens[e];
  for (int k = 0; k <= D; ++k) {
    LO oe = nbrs[std::size_t(e) * (D + 1) + k];
    if (oe < 0) continue;
    double os = sizes[oe];
    double mins = min2(s, os);
    double factor = mins / (s * double(D + 1));
    double x = dens[e];
    double ox = dens[oe];
    o += (ox - x) * factor;
  }
  out[e] = o;
  if (!(fabs(o) <= tol)) flags[0] = tag;
}

__global__ void diffusion_finish(LO ne, int nc, int q, Real const* ws, Real const* wd,
    Real* err) {
  LO e = thread_index();
  if (e >= ne) return;
  err[std::size_t(e) * nc + q] = wd[e] * ws[e];
}

// integrals of a conserved density (data times size)
__global__ void density_integrals(LO n, int nc, Real const* data, Real const* sizes,
    Real* out) {
  LO i = thread_index();
  if (i >= n) return;
  out[i] = data[i] * sizes[i / nc];
}

// correct_density_error: the density whose integral is the old integral
// minus the diffused error
__global__ void density_correct(LO n, int nc, Real const* oint, Real const* err,
    Real const* sizes, Real* data) {
  LO i = thread_index();
  if (i >= n) return;
  double nint = oint[i] - err[i];
  data[i] = nint / sizes[i / nc];
}

// correct_momentum_error, elementwise part: the element masses before and
// after the density correction, the element momenta, and the momentum
// change the density correction made, added to the error
template <int D>
__global__ void momentum_elems(LO ne, int nc, LO const* ev2v, Real const* vel,
    Real const* rho, Real const* rho_old, Real const* sizes, Real* mass, Real* newm,
    Real* err) {
  LO e = thread_index();
  if (e >= ne) return;
  double ms = rho[e] * sizes[e];
  double oms = rho_old[e] * sizes[e];
  mass[e] = ms;
  for (int q = 0; q < nc; ++q) {
    double comp = 0;
    for (int k = 0; k <= D; ++k) comp += vel[std::size_t(ev2v[std::size_t(e) * (D + 1) + k]) * nc + q];
    comp /= (D + 1);
    std::size_t i = std::size_t(e) * nc + q;
    double om = comp * oms;
    double nm = comp * ms;
    newm[i] = nm;
    err[i] = err[i] + (nm - om);
  }
}

// correct_momentum_error, vertex part: each vertex's velocity loses, per
// adjacent element (in upward-row order), that element's momentum error
// shared among its free vertices and divided by the vertex's mass (a third
// or a quarter of its elements' masses); components fixed on a vertex
// (bit q of `fixed`) stay
template <int D>
__global__ void momentum_verts(LO nv, int nc, LO const* v2e_a2ab, LO const* v2e_ab2b,
    LO const* ev2v, Real const* mass, Real const* fixed, Real const* err, Real const* vel,
    Real* out) {
  LO v = thread_index();
  if (v >= nv) return;
  double s = 0.0;
  for (LO j = v2e_a2ab[v]; j < v2e_a2ab[v + 1]; ++j) s = s + mass[v2e_ab2b[j]];
  double vm = s / double(D + 1);
  int vf = fixed ? int(fixed[v]) : 0;
  double o[3];
  for (int q = 0; q < nc; ++q) o[q] = vel[std::size_t(v) * nc + q];
  for (LO j = v2e_a2ab[v]; j < v2e_a2ab[v + 1]; ++j) {
    LO e = v2e_ab2b[j];
    double nfree[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k <= D; ++k) {
      LO v2 = ev2v[std::size_t(e) * (D + 1) + k];
      int f2 = fixed ? int(fixed[v2]) : 0;
      for (int q = 0; q < nc; ++q)
        if (!(f2 & (1 << q))) nfree[q] += 1.0;
    }
    for (int q = 0; q < nc; ++q)
      if (!(vf & (1 << q))) o[q] -= err[std::size_t(e) * nc + q] / (nfree[q] * vm);
  }
  for (int q = 0; q < nc; ++q) out[std::size_t(v) * nc + q] = o[q];
}

DMesh::ElemTag const* find_elem_tag(DMesh const& m, std::string const& name) {
  for (auto const& t : m.etags)
    if (t.name == name) return &t;
  return nullptr;
}

template <int D>
Quantity<D, false> density_quantity(Real const* data, int nc) {
  return Quantity<D, false>{data, nullptr, nullptr, nc};
}

template <int D>
void pass_errors(Ctx& c, DMesh& old, DMesh& nw, LO nkeys, LOs const& keys2kds, Adj const& up,
    LOs const& k2p, ModOut const& mo, Bytes const& kinds, int op) {
  LO nsame = mo.same2old.size();
  LO nprods = mo.prods2new.size();
  Reals const& osize = ask_sizes(c, old);
  Reals const& nsize = ask_sizes(c, nw);
  // OpConservation (src/Omega_h_conserve.cpp:294-305): a density is always
  // conserved in interior cavities and, by this operation, everywhere but
  // in key-boundary cavities of coarsening and swap; momentum is conserved
  // by refinement only
  int d_intro = op == 0 ? 0 : (1 << kKeyBdry);
  int d_carry = (1 << kTouchBdry) | (1 << kKeyBdry);
  int m_intro = op == 0 ? 0 : 7;
  int m_carry = 7;
  for (std::size_t t = 0; t < old.etags.size(); ++t) {
    auto const& a = old.etags[t];
    if (a.kind != 1 || !a.err.exists()) continue;
    auto& b = nw.etags[t];
    int nc = a.ncomps;
    Reals err(c, nw.nents[D] * nc);
    OSH_LAUNCH(errors_start, nsame + nprods, c, nsame, nc, mo.same2old.data(),
        mo.same2new.data(), nprods, mo.prods2new.data(), a.err.data(), err.data());
    OSH_LAUNCH((cavity_errors<D, false>), nkeys, c, keys2kds.data(), up.a2ab.data(),
        up.ab2b.data(), k2p.data(), mo.prods2new.data(), kinds.data(), old.class_id[D].data(),
        nw.class_id[D].data(), d_intro, d_carry, density_quantity<D>(a.data.data(), nc),
        density_quantity<D>(b.data.data(), nc), a.err.data(), osize.data(), nsize.data(), nc,
        err.data(), c.error_word());
    b.err = err;
  }
  LOs const& oev2v = verts_of(c, old, D);
  LOs const& nev2v = verts_of(c, nw, D);
  for (std::size_t t = 0; t < old.vtags.size(); ++t) {
    auto const& a = old.vtags[t];
    auto& b = nw.vtags[t];
    b.err = Reals();
    if (a.kind != 2 || !a.err.exists()) continue;
    auto const* orho = find_elem_tag(old, a.density);
    auto const* nrho = find_elem_tag(nw, b.density);
    OSH_CHECK(orho && nrho, "momentum velocity without its density");
    int nc = a.ncomps;
    Reals err(c, nw.nents[D] * nc);
    OSH_LAUNCH(errors_start, nsame + nprods, c, nsame, nc, mo.same2old.data(),
        mo.same2new.data(), nprods, mo.prods2new.data(), a.err.data(), err.data());
    Quantity<D, true> oq{a.data.data(), orho->data.data(), oev2v.data(), nc};
    Quantity<D, true> nq{b.data.data(), nrho->data.data(), nev2v.data(), nc};
    OSH_LAUNCH((cavity_errors<D, true>), nkeys, c, keys2kds.data(), up.a2ab.data(),
        up.ab2b.data(), k2p.data(), mo.prods2new.data(), kinds.data(), old.class_id[D].data(),
        nw.class_id[D].data(), m_intro, m_carry, oq, nq, a.err.data(), osize.data(),
        nsize.data(), nc, err.data(), c.error_word());
    b.err = err;
  }
}

template <int D>
void pass_tmpl(Ctx& c, DMesh& old, DMesh& nw, int op, int key_dim, LOs const& keys2kds,
    LOs const& k2p, ModOut const& mo) {
  LO nkeys = keys2kds.size();
  LO nprods = mo.prods2new.size();
  Adj const& up = ask_up(c, old, key_dim, D);
  LOs const& oev2v = verts_of(c, old, D);
  Bytes kinds(c, nkeys);
  OSH_LAUNCH(cavity_kinds<D>, nkeys, c, keys2kds.data(), old.class_dim[key_dim].data(),
      up.a2ab.data(), up.ab2b.data(), oev2v.data(), old.class_dim[0].data(), kinds.data());
  if (op != 0 && nprods > 0) {
    // densities by intersection: coarsening over whole interior cavities and
    // class by class in key-boundary cavities, swap over whole cavities;
    // refinement products already took their parent's value
    LOs const& nev2v = verts_of(c, nw, D);
    for (std::size_t t = 0; t < old.etags.size(); ++t) {
      auto const& a = old.etags[t];
      if (a.kind != 1 && a.kind != 2) continue;
      OSH_LAUNCH(intersect_densities<D>, nprods, c, nkeys, k2p.data(), keys2kds.data(),
          up.a2ab.data(), up.ab2b.data(), kinds.data(), op == 1 ? 1 : 0, mo.prods2new.data(),
          oev2v.data(), old.coords.data(), nev2v.data(), nw.coords.data(),
          old.class_id[D].data(), nw.class_id[D].data(), a.ncomps, a.data.data(),
          nw.etags[t].data.data(), c.error_word());
    }
  }
  pass_errors<D>(c, old, nw, nkeys, keys2kds, up, k2p, mo, kinds, op);
}

// diffuse_integrals_weighted over every component: sweeps until every
// weighted error density is within tolerance, one launch and one flag read
// per sweep (the reference loops without a limit; more than max_sweeps
// sweeps is an error here)
template <int D>
void diffuse_weighted(Ctx& c, LO ne, LOs const& nbrs, Diffusion const& opt, int nc, Reals& err,
    Reals const& quantity, int max_sweeps) {
  if (opt.type == 0) return;
  Reals ws(c, ne), a(c, ne), b(c, ne);
  LOs flags = fill_lo(c, 2, 0);
  int tag = 0;
  auto read_flags = [&](int* out) {
    OSH_HIP(hipMemcpyAsync(c.pinned, flags.data(), 2 * sizeof(int), hipMemcpyDeviceToHost,
        c.stream));
    c.sync();
    std::memcpy(out, c.pinned, 2 * sizeof(int));
  };
  for (int q = 0; q < nc; ++q) {
    int f[2];
    OSH_LAUNCH(diffusion_start, ne, c, nc, q, quantity.data(), err.data(), opt.floor, opt.tol,
        ws.data(), a.data(), flags.data(), ++tag);
    read_flags(f);
    OSH_CHECK(f[1] == 0, "conservation: an integral error where the weight is zero");
    int sweeps = 0;
    while (f[0] == tag) {
      OSH_CHECK(sweeps < max_sweeps, "conservation: the error diffusion did not converge");
      OSH_LAUNCH(diffusion_sweep<D>, ne, c, nbrs.data(), ws.data(), a.data(), opt.tol, b.data(),
          flags.data(), ++tag);
      read_flags(f);
      std::swap(a, b);
      ++sweeps;
    }
    OSH_LAUNCH(diffusion_finish, ne, c, nc, q, ws.data(), a.data(), err.data());
  }
}

template <int D>
void correct_tmpl(Ctx& c, DMesh& m, int max_sweeps) {
  LO ne = m.nents[D];
  LO nv = m.nents[0];
  LOs nbrs(c, ne * (D + 1));
  auto const& se = ask_side_elems(c, m);
  OSH_LAUNCH(diffusion_graph<D>, ne, c, m.down[D].data(), m.class_dim[D - 1].data(), se.data(),
      nbrs.data());
  Reals const& sizes = ask_sizes(c, m);
  for (auto& a : m.etags) {
    if (a.kind != 1 || !a.err.exists()) continue;
    int nc = a.ncomps;
    LO n = ne * nc;
    Reals old(c, n);
    OSH_HIP(hipMemcpyAsync(old.data(), a.data.data(), sizeof(Real) * std::size_t(n),
        hipMemcpyDeviceToDevice, c.stream));
    a.old = old;
    Reals oint(c, n);
    OSH_LAUNCH(density_integrals, n, c, nc, a.data.data(), sizes.data(), oint.data());
    diffuse_weighted<D>(c, ne, nbrs, a.diffuse, nc, a.err, oint, max_sweeps);
    Reals data(c, n);
    OSH_LAUNCH(density_correct, n, c, nc, oint.data(), a.err.data(), sizes.data(), data.data());
    a.data = data;
    a.err = Reals();
  }
  Adj const* v2e = nullptr;
  LOs const& ev2v = verts_of(c, m, D);
  for (auto& a : m.vtags) {
    if (a.kind != 2 || !a.err.exists()) continue;
    auto const* rho = find_elem_tag(m, a.density);
    OSH_CHECK(rho && rho->old.exists(), "momentum velocity without its conserved density");
    if (!v2e) v2e = &ask_up(c, m, 0, D);
    int nc = a.ncomps;
    Reals mass(c, ne), newm(c, ne * nc);
    OSH_LAUNCH(momentum_elems<D>, ne, c, nc, ev2v.data(), a.data.data(), rho->data.data(),
        rho->old.data(), sizes.data(), mass.data(), newm.data(), a.err.data());
    diffuse_weighted<D>(c, ne, nbrs, a.diffuse, nc, a.err, newm, max_sweeps);
    Real const* fixed = nullptr;
    for (auto const& it : m.itags)
      if (it.name == "momentum_velocity_fixed" && it.ncomps == 1) fixed = it.data[0].data();
    Reals out(c, nv * nc);
    OSH_LAUNCH(momentum_verts<D>, nv, c, nc, v2e->a2ab.data(), v2e->ab2b.data(), ev2v.data(),
        mass.data(), fixed, a.err.data(), a.data.data(), out.data());
    a.data = out;
    a.err = Reals();
  }
  for (auto& a : m.etags) a.old = Reals();
  c.sync();
}

}  // namespace

bool conserve_any(DMesh const& m) {
  for (auto const& t : m.etags)
    if (t.kind == 1) return true;
  return false;
}

void conserve_setup(Ctx& c, DMesh& m) {
  LO ne = m.nents[m.dim];
  for (auto& t : m.etags)
    if (t.kind == 1) t.err = fill_real(c, ne * t.ncomps, 0.0);
  for (auto& t : m.vtags)
    if (t.kind == 2) t.err = fill_real(c, ne * t.ncomps, 0.0);
}

void conserve_pass(Ctx& c, DMesh& old, DMesh& nw, int op, int key_dim, LOs const& keys2kds,
    LOs const& k2p, ModOut const& mo) {
  bool any = false;
  for (auto const& t : old.etags) any = any || t.kind == 1 || t.kind == 2;
  if (!any) {
    for (auto& t : nw.vtags) t.err = Reals();
    return;
  }
  // the reference's names for the three transfers (src/Omega_h_conserve.cpp:
  // 372, 490, 457)
  Range range(c, op == 0 ? "transfer_conserve_refine"
                         : (op == 1 ? "transfer_densities_and_conserve_coarsen"
                                    : "transfer_densities_and_conserve_swap"));
  if (old.dim == 3)
    pass_tmpl<3>(c, old, nw, op, key_dim, keys2kds, k2p, mo);
  else
    pass_tmpl<2>(c, old, nw, op, key_dim, keys2kds, k2p, mo);
}

void conserve_correct(Ctx& c, DMesh& m, int max_sweeps) {
  if (!conserve_any(m)) return;
  Range range(c, "correct_integral_errors");
  if (m.dim == 3)
    correct_tmpl<3>(c, m, max_sweeps);
  else
    correct_tmpl<2>(c, m, max_sweeps);
}

}  // namespace osh

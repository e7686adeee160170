/* Oracle (test infrastructure only): the intersection volume of two
 * simplices, restating the polytope clipping and moment code the reference
 * vendors (src/r3d.hpp, Devon Powell's r3d as modified for Omega_h):
 *
 *   init (tetrahedron, triangle)      src/r3d.hpp:459-495
 *   faces_from_verts                  src/r3d.hpp:509-560
 *   clip                              src/r3d.hpp:373-441, ClipHelper :277-361
 *   reduce<0> (volume, area)          src/r3d.hpp:629-735, 776-852
 *   measure / intersect_simplices     src/r3d.hpp:868-886
 *
 * A polytope is a vertex buffer in which every vertex lists its dim
 * neighbours (a simple polytope: three edges at a vertex in 3D, two in 2D).
 * A tetrahedron clipped by four planes stays convex with at most eight faces,
 * so at most twelve vertices, plus at most eighteen made while one plane is
 * applied: a buffer of 32 holds every case (the reference's holds 128). */
#include <math.h>
#include <string.h>

#include "om_internal.h"

#define R3D_MAXV 32

typedef struct {
  int nbr[3];
  double pos[3];
} r3d_vert;

typedef struct {
  r3d_vert v[R3D_MAXV];
  int n;
} r3d_poly;

typedef struct {
  double n[3];
  double d;
} r3d_plane;

static double r3d_dot(int dim, double const* a, double const* b) {
  double c = a[0] * b[0];
  for (int i = 1; i < dim; ++i) c += a[i] * b[i];
  return c;
}

static void r3d_normalize(int dim, double* v) {
  double l = om_sqrt(r3d_dot(dim, v, v));
  for (int i = 0; i < dim; ++i) v[i] = v[i] / l;
}

/* init(Polytope<3>&, ...) and init(Polytope<2>&, ...) */
static void r3d_init(int dim, r3d_poly* p, double const* pts) {
  if (dim == 3) {
    static int const nb[4][3] = {{1, 3, 2}, {2, 3, 0}, {0, 3, 1}, {1, 2, 0}};
    p->n = 4;
    for (int v = 0; v < 4; ++v) {
      for (int k = 0; k < 3; ++k) p->v[v].nbr[k] = nb[v][k];
      for (int k = 0; k < 3; ++k) p->v[v].pos[k] = pts[v * 3 + k];
    }
  } else {
    p->n = 3;
    for (int v = 0; v < 3; ++v) {
      p->v[v].pos[0] = pts[v * 2];
      p->v[v].pos[1] = pts[v * 2 + 1];
      p->v[v].nbr[0] = (v + 1) % 3;
      p->v[v].nbr[1] = (3 + v - 1) % 3;
    }
  }
}

/* tet_face_from_verts: the plane through a, b, c with the unit normal
 * (b - a) x (c - a), through the centroid (a + b + c) / 3 */
static void tet_face(double const* a, double const* b, double const* c, r3d_plane* f) {
  double ctr[3], u[3], w[3];
  for (int i = 0; i < 3; ++i) {
    ctr[i] = (a[i] + b[i] + c[i]) * (1.0 / 3.0);
    u[i] = b[i] - a[i];
    w[i] = c[i] - a[i];
  }
  f->n[0] = u[1] * w[2] - u[2] * w[1];
  f->n[1] = u[2] * w[0] - u[0] * w[2];
  f->n[2] = u[0] * w[1] - u[1] * w[0];
  r3d_normalize(3, f->n);
  f->d = -r3d_dot(3, f->n, ctr);
}

static void r3d_faces(int dim, double const* pts, r3d_plane* f) {
  if (dim == 3) {
    double const *v0 = pts, *v1 = pts + 3, *v2 = pts + 6, *v3 = pts + 9;
    tet_face(v3, v2, v1, &f[0]);
    tet_face(v0, v2, v3, &f[1]);
    tet_face(v0, v3, v1, &f[2]);
    tet_face(v0, v1, v2, &f[3]);
  } else {
    for (int k = 0; k < 3; ++k) {
      double const* p0 = pts + k * 2;
      double const* p1 = pts + ((k + 1) % 3) * 2;
      f[k].n[0] = p0[1] - p1[1];
      f[k].n[1] = p1[0] - p0[0];
      r3d_normalize(2, f[k].n);
      f[k].d = -r3d_dot(2, f[k].n, p0);
    }
  }
}

/* clip, with ClipHelper<3> and ClipHelper<2>; returns 1 if the buffer
 * would overflow (never for two simplices) */
static int r3d_clip(int dim, r3d_poly* poly, r3d_plane const* planes, int nplanes) {
  if (poly->n <= 0) return 0;
  double sd[R3D_MAXV];
  for (int p = 0; p < nplanes; ++p) {
    int onv = poly->n;
    double smin = 1.79769313486231570815e+308, smax = -1.79769313486231570815e+308;
    int clipped[R3D_MAXV];
    memset(clipped, 0, sizeof(clipped));
    for (int v = 0; v < onv; ++v) {
      sd[v] = planes[p].d + r3d_dot(dim, poly->v[v].pos, planes[p].n);
      if (sd[v] < smin) smin = sd[v];
      if (sd[v] > smax) smax = sd[v];
      if (sd[v] < 0.0) clipped[v] = 1;
    }
    if (smin >= 0.0) continue;
    if (smax <= 0.0) {
      poly->n = 0;
      return 0;
    }
    for (int vcur = 0; vcur < onv; ++vcur) {
      if (clipped[vcur]) continue;
      for (int np = 0; np < dim; ++np) {
        int vnext = poly->v[vcur].nbr[np];
        if (!clipped[vnext]) continue;
        if (poly->n >= R3D_MAXV) return 1;
        r3d_vert* nv = &poly->v[poly->n];
        if (dim == 3) {
          nv->nbr[0] = vcur;
        } else {
          nv->nbr[1 - np] = vcur;
          nv->nbr[np] = -1;
        }
        poly->v[vcur].nbr[np] = poly->n;
        /* wav(va, wa, vb, wb) = (wa va + wb vb) / (wa + wb) */
        double wa = -sd[vnext], wb = sd[vcur];
        for (int i = 0; i < dim; ++i)
          nv->pos[i] = (wa * poly->v[vcur].pos[i] + wb * poly->v[vnext].pos[i]) / (wa + wb);
        ++poly->n;
      }
    }
    if (dim == 3) {
      for (int vstart = onv; vstart < poly->n; ++vstart) {
        int vcur = vstart;
        int vnext = poly->v[vcur].nbr[0];
        do {
          int np;
          for (np = 0; np < 3; ++np)
            if (poly->v[vnext].nbr[np] == vcur) break;
          vcur = vnext;
          vnext = poly->v[vcur].nbr[(np + 1) % 3];
        } while (vcur < onv);
        poly->v[vstart].nbr[2] = vcur;
        poly->v[vcur].nbr[1] = vstart;
      }
    } else {
      for (int vstart = onv; vstart < poly->n; ++vstart) {
        if (poly->v[vstart].nbr[1] >= 0) continue;
        int vcur = poly->v[vstart].nbr[0];
        do {
          vcur = poly->v[vcur].nbr[0];
        } while (vcur < onv);
        poly->v[vstart].nbr[1] = vcur;
        poly->v[vcur].nbr[0] = vstart;
      }
    }
    int nun = 0;
    for (int v = 0; v < poly->n; ++v) {
      if (!clipped[v]) {
        poly->v[nun] = poly->v[v];
        clipped[v] = nun++;
      }
    }
    poly->n = nun;
    for (int v = 0; v < poly->n; ++v)
      for (int np = 0; np < dim; ++np) poly->v[v].nbr[np] = clipped[poly->v[v].nbr[np]];
  }
  return 0;
}

/* reduce<0>: the volume by a fan of triangles over every face (3D), the
 * area by the shoelace sum over edges (2D); integrate() then adds it to 0 */
static double r3d_measure(int dim, r3d_poly const* poly) {
  double m0 = 0.0;
  if (poly->n > 0) {
    if (dim == 3) {
      int emarks[R3D_MAXV][3];
      memset(emarks, 0, sizeof(emarks));
      for (int vstart = 0; vstart < poly->n; ++vstart) {
        for (int pstart = 0; pstart < 3; ++pstart) {
          if (emarks[vstart][pstart]) continue;
          int pnext = pstart;
          int vcur = vstart;
          emarks[vcur][pnext] = 1;
          int vnext = poly->v[vcur].nbr[pnext];
          double const* v0 = poly->v[vcur].pos;
          int np;
          for (np = 0; np < 3; ++np)
            if (poly->v[vnext].nbr[np] == vcur) break;
          vcur = vnext;
          pnext = (np + 1) % 3;
          emarks[vcur][pnext] = 1;
          vnext = poly->v[vcur].nbr[pnext];
          while (vnext != vstart) {
            double const* v2 = poly->v[vcur].pos;
            double const* v1 = poly->v[vnext].pos;
            double sixv = (-v2[0] * v1[1] * v0[2] + v1[0] * v2[1] * v0[2] +
                           v2[0] * v0[1] * v1[2] - v0[0] * v2[1] * v1[2] -
                           v1[0] * v0[1] * v2[2] + v0[0] * v1[1] * v2[2]);
            m0 += (1.0 / 6.0) * sixv;
            for (np = 0; np < 3; ++np)
              if (poly->v[vnext].nbr[np] == vcur) break;
            vcur = vnext;
            pnext = (np + 1) % 3;
            emarks[vcur][pnext] = 1;
            vnext = poly->v[vcur].nbr[pnext];
          }
        }
      }
    } else {
      for (int vcur = 0; vcur < poly->n; ++vcur) {
        int vnext = poly->v[vcur].nbr[0];
        double const* v0 = poly->v[vcur].pos;
        double const* v1 = poly->v[vnext].pos;
        double twoa = (v0[0] * v1[1] - v0[1] * v1[0]);
        m0 += 0.5 * twoa;
      }
    }
  }
  double result = 0;
  result += m0 * 1.0;
  return result;
}

/* intersect_simplices then measure: the size of the intersection of the
 * simplex a (the polytope) with the simplex b (the clipping planes); points
 * are (dim + 1) x dim.  *bad is set if the buffer would overflow. */
double om_intersect_size(int dim, double const* a, double const* b, int* bad) {
  r3d_poly poly;
  r3d_plane f[4];
  r3d_init(dim, &poly, a);
  r3d_faces(dim, b, f);
  if (r3d_clip(dim, &poly, f, dim + 1)) {
    *bad = 1;
    return 0.0;
  }
  return r3d_measure(dim, &poly);
}

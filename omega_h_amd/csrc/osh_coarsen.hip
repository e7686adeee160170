// Edge-collapse (coarsening) pass on the device.
//
//   coarsen_by_size / coarsen_slivers / coarsen / coarsen_verts
//                                    src/Omega_h_coarsen.cpp:29-220
//   check_collapse_class / check_collapse_exposure
//                                    src/Omega_h_check_collapse.cpp:10-117
//   prevent_coarsen_overshoot        src/Omega_h_coarsen_overshoot.cpp:12-49
//   prevent_coarsen_flip             src/Omega_h_coarsen_flip.cpp:14-150,
//                                    src/Omega_h_most_normal.hpp:15-93
//   coarsen_qualities, filter_coarsen_min_qual / _improve
//                                    src/Omega_h_coarsen_qualities.cpp:16-122
//   choose_rails / find_rails        src/Omega_h_collapse_rail.cpp:15-101
//   mark_dead_ents / find_coarsen_domains / coarsen_topology
//                                    src/Omega_h_coarsen_topology.cpp:15-122
//   transfer_coarsen                 src/Omega_h_transfer.cpp:585-619
#include <cstdio>

#include "osh_math.hpp"
#include "osh_passes.hpp"

namespace osh {

namespace {

constexpr I8 kDont = 0;

__device__ __forceinline__ bool collapses(I8 code, int col) { return (code & (1 << col)) != 0; }
__device__ __forceinline__ I8 dont_collapse(I8 code, int col) { return I8(code & ~(1 << col)); }

__global__ void lt_kernel(LO n, Real const* a, Real v, I8* out) {
  LO i = thread_index();
  if (i < n) out[i] = a[i] < v;
}

__global__ void edge_codes_from_verts(LO ne, LO const* ev2v, I8 const* vmarks, I8* codes) {
  LO e = thread_index();
  if (e >= ne) return;
  I8 code = kDont;
  for (int eev = 0; eev < 2; ++eev)
    if (vmarks[ev2v[e * 2 + eev]]) code = I8(code | (1 << eev));
  codes[e] = code;
}

__global__ void check_class_kernel(LO ne, LO const* ev2v, I8 const* vdim, I8 const* edim,
    I8* codes) {
  LO e = thread_index();
  if (e >= ne) return;
  I8 code = codes[e];
  if (code == kDont) return;
  int cd0 = vdim[ev2v[e * 2]], cd1 = vdim[ev2v[e * 2 + 1]];
  int ecd = edim[e];
  I8 orig = code;
  if (cd0 == cd1) {
    if (cd0 != ecd) code = kDont;
  } else {
    int cds[2] = {cd0, cd1};
    for (int col = 0; col < 2; ++col)
      if (collapses(orig, col) && cds[col] != ecd) code = dont_collapse(code, col);
  }
  codes[e] = code;
}

// check_collapse_exposure from the cells' side: every cell of dimension CD
// looks at its own edges, and a collapse side whose cell's side opposite the
// target vertex is classified differently from the cell is vetoed (the
// order of the cells does not matter, so no edge -> cell rows are built)
template <int CD>
__device__ __forceinline__ int local_vertex(LO const* cv, LO v) {
  int k = 0;
#pragma unroll
  for (int j = 1; j < CD + 1; ++j)
    if (cv[j] == v) k = j;
  return k;
}

template <int CD>
__global__ void check_exposure_cells(LO ncells, LO const* c2e, LO const* cv2v, LO const* cs2s,
    I8 const* cdim, I8 const* sdim, LO const* ev2v, I8 const* codes, I8* veto) {
  LO c = thread_index();
  if (c >= ncells) return;
  constexpr int nepc = (CD == 3) ? 6 : 3;
  LO const* cv = cv2v + c * (CD + 1);
  for (int j = 0; j < nepc; ++j) {
    LO e = c2e[c * nepc + j];
    I8 code = codes[e];
    if (code == kDont) continue;
    for (int col = 0; col < 2; ++col) {
      if (!collapses(code, col)) continue;
      int ccv_onto = local_vertex<CD>(cv, ev2v[e * 2 + (1 - col)]);
      int ccs_opp = opposite_template(CD, 0, ccv_onto);
      if (sdim[cs2s[c * (CD + 1) + ccs_opp]] != cdim[c]) veto[e * 2 + col] = 1;
    }
  }
}

// the vetoed sides of the candidate edges leave their codes
__global__ void apply_edge_vetoes(LO ncands, LO const* cands, I8 const* veto, I8* codes) {
  LO i = thread_index();
  if (i >= ncands) return;
  LO e = cands[i];
  I8 code = codes[e];
  for (int col = 0; col < 2; ++col)
    if (veto[e * 2 + col]) code = dont_collapse(code, col);
  codes[e] = code;
}

// Each collapse side (candidate edge, direction) gets 2^lshift threads
// (Tuning::coarsen_lanes), which split the collapsing vertex's row between
// them, so that passes with few candidates still spread their work over the
// device.  Minima do not depend on how a row is split, so the lane count
// changes no result.  The product builders below use kLanes threads per key.
constexpr int kLanes = 8;

// prevent_coarsen_overshoot: a side is vetoed when any new edge would reach
// max_length (the order of the row does not matter).  F32 (mixed
// precision): the lengths are FP32 estimates; a side is vetoed when one is
// over max_length by more than the screening band, and sides with a length
// inside the band are marked in `need` for the FP64 sweep.  With `need`
// given, the FP64 sweep looks only at marked sides not already vetoed.
template <int D, int M, bool F32>
__global__ void overshoot_lanes(LO nt, int lshift, LO const* cands, LO const* ev2v,
    LO const* v2e_a2ab, LO const* v2e_ab2b, I8 const* v2e_codes, Real const* coords,
    Real const* metric, Real max_length, I8 const* codes, I8* veto, I8* need) {
  LO t = thread_index();
  if (t >= nt) return;
  LO side = t >> lshift;
  int lane = t & ((1 << lshift) - 1);
  LO e = cands[side / 2];
  int col = side % 2;
  if (!collapses(codes[e], col)) return;
  if (!F32 && need && (!need[side] || veto[side])) return;
  constexpr int nc = (M == 1) ? 1 : (M == 2 ? 3 : 6);
  LO v_col = ev2v[e * 2 + col];
  LO v_onto = ev2v[e * 2 + (1 - col)];
  for (LO ve = v2e_a2ab[v_col] + lane; ve < v2e_a2ab[v_col + 1]; ve += (1 << lshift)) {
    LO e2 = v2e_ab2b[ve];
    if (e2 == e) continue;
    int eev_in = code_which_down(v2e_codes[ve]);
    int eev_out = 1 - eev_in;
    LO nv[2];
    nv[eev_in] = v_onto;
    nv[eev_out] = ev2v[e2 * 2 + eev_out];
    double p0[D], p1[D], m0[nc], m1[nc];
    for (int k = 0; k < D; ++k) {
      p0[k] = coords[nv[0] * D + k];
      p1[k] = coords[nv[1] * D + k];
    }
    for (int k = 0; k < nc; ++k) {
      m0[k] = metric[nv[0] * nc + k];
      m1[k] = metric[nv[1] * nc + k];
    }
    if constexpr (F32) {
      float l = f32::edge_length<D, M>(p0, p1, m0, m1);
      float lmax = float(max_length);
      if (l >= lmax * (1.0f + f32::kScreenBand)) {
        veto[side] = 1;
        return;
      }
      if (!(l < lmax * (1.0f - f32::kScreenBand))) need[side] = 1;
    } else {
      if (edge_length<D, M>(p0, p1, m0, m1) >= max_length) {
        veto[side] = 1;
        return;
      }
    }
  }
}

__global__ void apply_vetoes(LO ncands, LO const* cands, I8 const* veto, I8* codes) {
  LO i = thread_index();
  if (i >= ncands) return;
  LO e = cands[i];
  I8 code = codes[e];
  for (int col = 0; col < 2; ++col)
    if (veto[i * 2 + col]) code = dont_collapse(code, col);
  codes[e] = code;
}

// coarsen_qualities: each lane folds the qualities of its share of the
// collapsing vertex's cells (those that survive the collapse, with v_col
// moved onto v_onto) into a partial minimum.  F32 (mixed precision): the
// qualities are FP32 estimates, and a lane that meets an element whose
// max-determinant metric is not clear marks its side in `unsure`.  With
// `need` given, the FP64 sweep evaluates only the marked sides.
template <int D, int M, bool F32>
__global__ void __launch_bounds__(kBlock, 6) coarsen_quality_lanes(LO nt, int lshift,
    LO const* cands, LO const* ev2v, LO const* cv2v, LO const* v2c_a2ab, LO const* v2c_ab2b,
    I8 const* v2c_codes, Real const* coords, Real const* metric, I8 const* codes,
    I8 const* need, I8* unsure, Real* partial) {
  LO t = thread_index();
  if (t >= nt) return;
  LO side = t >> lshift;
  int lane = t & ((1 << lshift) - 1);
  LO e = cands[side / 2];
  int col = side % 2;
  double minqual = 1.0;
  float minq32 = 1.0f;
  bool unclear = false;
  if (collapses(codes[e], col) && (F32 || !need || need[side])) {
    constexpr int nc = (M == 1) ? 1 : (M == 2 ? 3 : 6);
    LO v_col = ev2v[e * 2 + col];
    LO v_onto = ev2v[e * 2 + (1 - col)];
    for (LO vc = v2c_a2ab[v_col] + lane; vc < v2c_a2ab[v_col + 1]; vc += (1 << lshift)) {
      LO c = v2c_ab2b[vc];
      int ccv_col = code_which_down(v2c_codes[vc]);
      LO cv[D + 1];
      bool will_die = false;
      for (int k = 0; k < D + 1; ++k) {
        cv[k] = cv2v[c * (D + 1) + k];
        if (k != ccv_col && cv[k] == v_onto) will_die = true;
      }
      if (will_die) continue;
      cv[ccv_col] = v_onto;
      // the quality reads the vertices' coordinates and metrics where they
      // lie, so that they are loaded as they are used rather than all held
      // at once (fewer live registers, more waves per SIMD)
      double const* pp[D + 1];
      double const* ms[D + 1];
      for (int k = 0; k < D + 1; ++k) {
        pp[k] = coords + cv[k] * D;
        ms[k] = metric + cv[k] * nc;
      }
      if constexpr (F32) {
        float q = f32::element_quality<D, M>(pp, ms, &unclear);
        minq32 = (q < minq32) ? q : minq32;
      } else {
        minqual = math::min2(minqual, element_quality<D, M>(pp, ms));
      }
    }
  }
  if constexpr (F32) {
    if (unclear) unsure[side] = 1;
    partial[t] = double(minq32);
  } else {
    partial[t] = minqual;
  }
}

// the minimum over the lanes: the minimum over the row, as the reference's
// loop computes it (min2 never takes a NaN, and a zero's sign decides
// nothing downstream: the filters, rails and independent set compare it
// only).  After a mixed-precision screen (`need` given) a side the screen
// rejected gets -1, below every quality threshold.
__global__ void combine_quality_lanes(LO ncands, int lanes, LO const* cands, I8 const* codes,
    I8 const* need, Real const* partial, Real* quals) {
  LO i = thread_index();
  if (i >= ncands) return;
  LO e = cands[i];
  I8 code = codes[e];
  for (int col = 0; col < 2; ++col) {
    double q = -1.0;
    if (collapses(code, col) && (!need || need[i * 2 + col])) {
      q = 1.0;
      for (int l = 0; l < lanes; ++l) q = math::min2(q, partial[(i * 2 + col) * lanes + l]);
    }
    quals[e * 2 + col] = q;
  }
}

// the screen's verdict per side: FP64 is needed unless the FP32 minimum is
// below qmin by more than the band and every element's metric was clear
// (a NaN estimate compares false and is evaluated again)
__global__ void screen_quality_sides(LO nsides, int lanes, LO const* cands, I8 const* codes,
    I8 const* unsure, Real const* partial, Real qmin, I8* need) {
  LO side = thread_index();
  if (side >= nsides) return;
  LO e = cands[side / 2];
  int col = side % 2;
  if (!collapses(codes[e], col)) {
    need[side] = 0;
    return;
  }
  double q = 1.0;
  for (int l = 0; l < lanes; ++l) q = math::min2(q, partial[side * lanes + l]);
  bool reject = !unsure[side] && q < qmin * (1.0 - double(f32::kScreenBand));
  need[side] = reject ? 0 : 1;
}

// combine_quality_lanes with one output row per candidate (the layout of
// the reference's coarsen_qualities, two values per candidate)
__global__ void combine_to_cands(LO ncands, int lanes, LO const* cands, I8 const* codes,
    I8 const* need, Real const* partial, Real* out) {
  LO i = thread_index();
  if (i >= ncands) return;
  I8 code = codes[cands[i]];
  for (int col = 0; col < 2; ++col) {
    double q = -1.0;
    if (collapses(code, col) && (!need || need[i * 2 + col])) {
      q = 1.0;
      for (int l = 0; l < lanes; ++l) q = math::min2(q, partial[(i * 2 + col) * lanes + l]);
    }
    out[i * 2 + col] = q;
  }
}

__global__ void scatter_codes(LO n, LO const* cands, I8 const* cand_codes, I8* codes) {
  LO i = thread_index();
  if (i < n) codes[cands[i]] = cand_codes[i];
}

// the vertices that would collapse along some candidate edge
__global__ void mark_collapsing_verts(LO ncands, LO const* cands, LO const* ev2v,
    I8 const* codes, I8* vmarks) {
  LO i = thread_index();
  if (i >= ncands) return;
  LO e = cands[i];
  I8 code = codes[e];
  for (int col = 0; col < 2; ++col)
    if (collapses(code, col)) vmarks[ev2v[e * 2 + col]] = 1;
}

// the vertex star (verts_across_edges, src/Omega_h_adj.cpp:512-530) of the
// vertices whose vertex -> edge rows are built
__global__ void star_from_v2e(LO nv, LO const* a2ab, LO const* ab2b, I8 const* codes,
    LO const* ev2v, LO* out) {
  LO v = thread_index();
  if (v >= nv) return;
  for (LO ve = a2ab[v]; ve < a2ab[v + 1]; ++ve)
    out[ve] = ev2v[ab2b[ve] * 2 + (1 - code_which_down(codes[ve]))];
}

__global__ void vert_min_quality(LO nv, LO const* a2ab, LO const* ab2b, Real const* q, Real* out) {
  LO v = thread_index();
  if (v >= nv) return;
  double r = 1.79769313486231570815e+308;
  for (LO j = a2ab[v]; j < a2ab[v + 1]; ++j) r = math::min2(r, q[ab2b[j]]);
  out[v] = r;
}

__global__ void filter_quals_kernel(LO ncands, LO const* cands, LO const* ev2v,
    Real const* quals, Real qmin, Real const* vold, I8* codes) {
  LO i = thread_index();
  if (i >= ncands) return;
  LO e = cands[i];
  I8 code = codes[e];
  if (code == kDont) return;
  for (int col = 0; col < 2; ++col)
    if (!(quals[e * 2 + col] >= qmin)) code = dont_collapse(code, col);
  if (vold)
    for (int col = 0; col < 2; ++col)
      if (!(quals[e * 2 + col] > vold[ev2v[e * 2 + col]])) code = dont_collapse(code, col);
  codes[e] = code;
}

__global__ void choose_rails_kernel(LO nv, LO const* v2e_a2ab, LO const* v2e_ab2b,
    I8 const* v2e_codes, I8 const* codes, Real const* lengths, Real const* quals, I8* vcand,
    Real* vqual, LO* vrail) {
  LO v = thread_index();
  if (v >= nv) return;
  LO best = -1;
  double best_len = -1, best_q = -1;
  I8 is_cand = 0;
  for (LO ve = v2e_a2ab[v]; ve < v2e_a2ab[v + 1]; ++ve) {
    LO e = v2e_ab2b[ve];
    I8 code = codes[e];
    if (code == kDont) continue;
    int eev = code_which_down(v2e_codes[ve]);
    if (!collapses(code, eev)) continue;
    double len = lengths[e];
    double q = quals[e * 2 + eev];
    if (best == -1 || len < best_len || (len == best_len && q > best_q)) {
      is_cand = 1;
      best = e;
      best_len = len;
      best_q = q;
    }
  }
  vcand[v] = is_cand;
  vqual[v] = best_q;
  vrail[v] = best;
}

__global__ void find_rails_kernel(LO nkeys, LO const* keys2verts, LO const* vrail,
    LO const* v2e_a2ab, LO const* v2e_ab2b, I8 const* v2e_codes, LO const* ev2v,
    LO* rails2edges, LO* col_dirs, LO* keys2onto) {
  LO k = thread_index();
  if (k >= nkeys) return;
  LO v = keys2verts[k];
  for (LO ve = v2e_a2ab[v]; ve < v2e_a2ab[v + 1]; ++ve) {
    if (v2e_ab2b[ve] == vrail[v]) {
      LO e = v2e_ab2b[ve];
      int eev = code_which_down(v2e_codes[ve]);
      rails2edges[k] = e;
      col_dirs[k] = eev;
      keys2onto[k] = ev2v[e * 2 + (1 - eev)];
      return;
    }
  }
}

// mark_dead_ents (src/Omega_h_coarsen_topology.cpp:15-60): the rails die,
// and so do the cells around a rail and each such cell's side opposite the
// target vertex; rail_col[e] is the collapsing end of rail e (-1 elsewhere)
__global__ void mark_rails(LO nkeys, LO const* rails2edges, LO const* col_dirs, I8* dead_edges,
    I8* rail_col) {
  LO k = thread_index();
  if (k >= nkeys) return;
  LO e = rails2edges[k];
  dead_edges[e] = 1;
  rail_col[e] = I8(col_dirs[k]);
}

template <int CD>
__global__ void mark_dead_cells(LO ncells, LO const* c2e, LO const* cv2v, LO const* cs2s,
    LO const* ev2v, I8 const* rail_col, I8* dead_cells, I8* dead_sides) {
  LO c = thread_index();
  if (c >= ncells) return;
  constexpr int nepc = (CD == 3) ? 6 : 3;
  for (int j = 0; j < nepc; ++j) {
    LO e = c2e[c * nepc + j];
    int col = rail_col[e];
    if (col < 0) continue;
    int ccv_onto = local_vertex<CD>(cv2v + c * (CD + 1), ev2v[e * 2 + (1 - col)]);
    dead_cells[c] = 1;
    dead_sides[cs2s[c * (CD + 1) + opposite_template(CD, 0, ccv_onto)]] = 1;
  }
}

// a key's products are the live entries of its vertex's row, in row order;
// kLanes threads per key share the row (the counts of all dimensions in one
// launch, added into zeroed counts)
struct LiveRows {
  LO const* a2ab[3];
  LO const* ab2b[3];
  I8 const* dead[3];
  LO* counts[3];
};

__global__ void live_counts(LO nt, LO nkeys, LO const* keys2verts, LiveRows r) {
  LO t = thread_index();
  if (t >= nt) return;
  LO per = nkeys * kLanes;
  int d = t / per;
  LO k = (t % per) / kLanes;
  int lane = t % kLanes;
  LO v = keys2verts[k];
  LO n = 0;
  for (LO j = r.a2ab[d][v] + lane; j < r.a2ab[d][v + 1]; j += kLanes) n += !r.dead[d][r.ab2b[d][j]];
  if (n) atomicAdd(&r.counts[d][k], n);
}

__global__ void coarsen_products(LO nt, int nvd, LO const* keys2verts, LO const* keys2onto,
    LO const* a2ab, LO const* ab2b, I8 const* codes, I8 const* dead, LO const* dv2v,
    LO const* o2n, LO const* keys2prods, LO* pv, LO* psrc) {
  LO t = thread_index();
  if (t >= nt) return;
  LO k = t / kLanes;
  int lane = t % kLanes;
  LO v = keys2verts[k];
  LO b = a2ab[v], e = a2ab[v + 1];
  for (LO j = b + lane; j < e; j += kLanes) {
    LO cell = ab2b[j];
    if (dead[cell]) continue;
    LO prod = keys2prods[k];
    for (LO i = b; i < j; ++i) prod += !dead[ab2b[i]];
    int ccv_col = code_which_down(codes[j]);
    for (int ccv = 0; ccv < nvd; ++ccv) {
      LO ov = (ccv == ccv_col) ? keys2onto[k] : dv2v[cell * nvd + ccv];
      pv[prod * nvd + ccv] = o2n[ov];
    }
    psrc[prod] = cell;
  }
}

// ---- prevent_coarsen_flip (src/Omega_h_coarsen_flip.cpp:68-150), taken
// only when AdaptOpts::should_prevent_coarsen_flip is set.  Exposed sides
// (fewer than two elements) around a vertex of a candidate edge "matter", and
// so do their vertices; each such vertex gets a normal, and a collapse
// direction is vetoed when a surviving side that matters around the
// collapsing vertex would turn so that its normal has a dot product below
// 1e-10 with the normal of one of its vertices.

// get_side_vector (src/Omega_h_shape.hpp:276-298): the triangle's
// (b - a) x (c - a) in 3D, -perp(b - a) in 2D
template <int D>
__device__ __forceinline__ void side_vector(double const* const* p, double* out) {
  if (D == 3) {
    double u[3], w[3];
    for (int i = 0; i < 3; ++i) {
      u[i] = p[1][i] - p[0][i];
      w[i] = p[2][i] - p[0][i];
    }
    cross3(u, w, out);
  } else {
    double v0 = p[1][0] - p[0][0], v1 = p[1][1] - p[0][1];
    out[0] = -(-v1);
    out[1] = -v0;
  }
}

template <int D>
__device__ __forceinline__ void normalize_vec(double* v) {
  double l = vnorm<D>(v);
  for (int i = 0; i < D; ++i) v[i] = v[i] / l;
}

__device__ __forceinline__ bool close_rel(double a, double b) {
  double am = math::fabs_(a), bm = math::fabs_(b);
  if (am <= 1e-10 && bm <= 1e-10) return true;
  return math::fabs_(b - a) / math::max2(am, bm) <= 1e-10;
}

// get_most_normal_normal (src/Omega_h_most_normal.hpp:15-93; Aubry and
// Lohner 2008): the centre of the smallest spherical cap holding the n unit
// vectors, from pairs (their normalized bisector) and then from triplets
// (the direction with one dot product with all three, from the inverse of
// the matrix whose rows are the three).  In 2D the normalized sum of the
// first two.  False where the reference's checks fail.
template <int D>
__device__ bool most_normal(int n, double const* N, double* c) {
  if (D == 2) {
    c[0] = N[0] + N[2];
    c[1] = N[1] + N[3];
    normalize_vec<2>(c);
    return true;
  }
  double const tol = 1e-10;
  double scalmin = -1.0;
  c[0] = c[1] = c[2] = 0.0;
  for (int i = 0; i < n; ++i) {
    for (int j = i + 1; j < n; ++j) {
      double b[3];
      for (int k = 0; k < 3; ++k) b[k] = N[i * 3 + k] + N[j * 3 + k];
      normalize_vec<3>(b);
      double scal = dot<3>(b, N + i * 3);
      if (scal < scalmin) continue;
      int l;
      for (l = 0; l < n; ++l)
        if (dot<3>(N + l * 3, b) + tol < scal) break;
      if (l < n) continue;
      for (int k = 0; k < 3; ++k) c[k] = b[k];
      scalmin = scal;
    }
  }
  if (scalmin != -1.0) return true;
  for (int i = 0; i < n; ++i) {
    for (int j = i + 1; j < n; ++j) {
      for (int k = j + 1; k < n; ++k) {
        Mat M;  // columns N_i, N_j, N_k
        for (int r = 0; r < 3; ++r) {
          M.m[0][r] = N[i * 3 + r];
          M.m[1][r] = N[j * 3 + r];
          M.m[2][r] = N[k * 3 + r];
        }
        if (math::fabs_(det<3>(M)) < 1e-6) continue;
        // invert(transpose(M)) * (1, 1, 1) (src/Omega_h_matrix.hpp:89-94,
        // 344-350): rows of the cofactor matrix from the cross products of
        // the transpose's columns, divided by its determinant, summed
        // column by column
        Mat T = transpose<3>(M);
        double bm[3][3];
        for (int a = 0; a < 3; ++a) cross3(T.m[(a + 1) % 3], T.m[(a + 2) % 3], bm[a]);
        double dt = det<3>(T);
        double nc[3];
        for (int r = 0; r < 3; ++r) {
          double s = (bm[r][0] / dt) * 1.0;
          s = s + (bm[r][1] / dt) * 1.0;
          s = s + (bm[r][2] / dt) * 1.0;
          nc[r] = s;
        }
        normalize_vec<3>(nc);
        double scal = dot<3>(nc, N + i * 3);
        if (!close_rel(scal, dot<3>(nc, N + j * 3)) || !close_rel(scal, dot<3>(nc, N + k * 3)))
          return false;
        if (scal < scalmin) continue;
        int l;
        for (l = 0; l < n; ++l)
          if (dot<3>(N + l * 3, nc) + tol < scal) break;
        if (l < n) continue;
        for (int q = 0; q < 3; ++q) c[q] = nc[q];
        scalmin = scal;
      }
    }
  }
  return scalmin > -1.0;
}

__global__ void flip_vert_cands(LO ncands, LO const* cands, LO const* ev2v, I8 const* codes,
    I8* vcand) {
  LO i = thread_index();
  if (i >= ncands) return;
  LO e = cands[i];
  if (codes[e] == kDont) return;
  vcand[ev2v[e * 2]] = 1;
  vcand[ev2v[e * 2 + 1]] = 1;
}

// mark_exposed_sides (src/Omega_h_mark.cpp:12-19) and the sides and
// vertices that matter (src/Omega_h_coarsen_flip.cpp:128-134)
__global__ void flip_sides(LO ns, int nsv, LO const* sv2v, unsigned long long const* side_elems,
    I8 const* vcand, I8* exposed, I8* smatter, I8* vmatter) {
  LO s = thread_index();
  if (s >= ns) return;
  I8 ex = (side_elems[s] >> 32) < 2 ? 1 : 0;
  bool adj = false;
  for (int k = 0; k < nsv; ++k) adj = adj || vcand[sv2v[s * nsv + k]] != 0;
  I8 mt = (adj && ex) ? 1 : 0;
  exposed[s] = ex;
  smatter[s] = mt;
  if (mt)
    for (int k = 0; k < nsv; ++k) vmatter[sv2v[s * nsv + k]] = 1;
}

// compute_flip_normals_dim (src/Omega_h_coarsen_flip.cpp:14-64): zero for
// vertices that do not matter
template <int D>
__global__ void flip_normals(LO nv, LO const* a2ab, LO const* ab2b, LO const* sv2v,
    Real const* coords, I8 const* exposed, I8 const* vmatter, Real* normals, int* error) {
  constexpr int kMaxAdj = (D == 3) ? 24 : 8;
  LO v = thread_index();
  if (v >= nv) return;
  double Nc[D];
  for (int k = 0; k < D; ++k) Nc[k] = 0.0;
  if (vmatter[v]) {
    double N[kMaxAdj * D];
    int n = 0;
    for (LO vs = a2ab[v]; vs < a2ab[v + 1]; ++vs) {
      LO s = ab2b[vs];
      if (!exposed[s]) continue;
      if (n >= kMaxAdj) {
        *error = kFlipNormals;
        return;
      }
      double const* p[D];
      for (int k = 0; k < D; ++k) p[k] = coords + size_t(sv2v[s * D + k]) * D;
      double sv[3];
      side_vector<D>(p, sv);
      for (int k = 0; k < D; ++k) Nc[k] = Nc[k] + sv[k];
      double l = vnorm<D>(sv);
      for (int k = 0; k < D; ++k) N[n * D + k] = sv[k] / l;
      ++n;
    }
    normalize_vec<D>(Nc);
    int i;
    for (i = 0; i < n; ++i)
      if (dot<D>(Nc, N + i * D) < 0.95) break;
    if (i < n && !most_normal<D>(n, N, Nc)) {
      *error = kFlipNormals;
      return;
    }
  }
  for (int k = 0; k < D; ++k) normals[size_t(v) * D + k] = Nc[k];
}

// prevent_coarsen_flip2_dim (src/Omega_h_coarsen_flip.cpp:70-121)
template <int D>
__global__ void flip_veto(LO ncands, LO const* cands, LO const* ev2v, LO const* a2ab,
    LO const* ab2b, I8 const* vs_codes, LO const* sv2v, Real const* coords,
    Real const* normals, I8 const* smatter, I8 const* vmatter, I8* codes) {
  LO i = thread_index();
  if (i >= ncands) return;
  LO e = cands[i];
  I8 code = codes[e];
  if (code == kDont) return;
  for (int col = 0; col < 2; ++col) {
    if (!collapses(code, col)) continue;
    LO v_col = ev2v[e * 2 + col];
    if (!vmatter[v_col]) continue;
    LO v_onto = ev2v[e * 2 + (1 - col)];
    for (LO vs = a2ab[v_col]; vs < a2ab[v_col + 1]; ++vs) {
      LO s = ab2b[vs];
      if (!smatter[s]) continue;
      int ssv_col = vs_codes[vs] >> 3;
      LO sv[D];
      int k;
      for (k = 0; k < D; ++k) sv[k] = sv2v[s * D + k];
      for (k = 0; k < D; ++k)
        if (sv[k] == v_onto) break;
      if (k != D) continue;
      sv[ssv_col] = v_onto;
      double const* p[D];
      for (k = 0; k < D; ++k) p[k] = coords + size_t(sv[k]) * D;
      double sn[3];
      side_vector<D>(p, sn);
      for (k = 0; k < D; ++k)
        if (dot<D>(normals + size_t(sv[k]) * D, sn) < 1e-10) break;
      if (k < D) code = dont_collapse(code, col);
    }
  }
  codes[e] = code;
}

// the whole check on the candidates' codes, in place
void prevent_coarsen_flip(Ctx& c, DMesh& m, LOs const& cands, Bytes& codes) {
  Range range(c, "prevent_coarsen_flip");
  int dim = m.dim;
  int sd = dim - 1;
  LO nv = m.nents[0], ns = m.nents[sd];
  LO ncands = cands.size();
  LO const* ev2v = m.down[1].data();
  LOs const& sv2v = verts_of(c, m, sd);
  Adj const& v2s = ask_up(c, m, 0, sd);
  auto const& se = ask_side_elems(c, m);
  Bytes vcand = fill_i8(c, nv, 0);
  OSH_LAUNCH(flip_vert_cands, ncands, c, cands.data(), ev2v, codes.data(), vcand.data());
  Bytes exposed(c, ns), smatter(c, ns);
  Bytes vmatter = fill_i8(c, nv, 0);
  OSH_LAUNCH(flip_sides, ns, c, sd + 1, sv2v.data(), se.data(), vcand.data(), exposed.data(),
      smatter.data(), vmatter.data());
  Reals normals(c, LO(nv * dim));
  if (dim == 3) {
    OSH_LAUNCH(flip_normals<3>, nv, c, v2s.a2ab.data(), v2s.ab2b.data(), sv2v.data(),
        m.coords.data(), exposed.data(), vmatter.data(), normals.data(), c.error_word());
    OSH_LAUNCH(flip_veto<3>, ncands, c, cands.data(), ev2v, v2s.a2ab.data(), v2s.ab2b.data(),
        v2s.codes.data(), sv2v.data(), m.coords.data(), normals.data(), smatter.data(),
        vmatter.data(), codes.data());
  } else {
    OSH_LAUNCH(flip_normals<2>, nv, c, v2s.a2ab.data(), v2s.ab2b.data(), sv2v.data(),
        m.coords.data(), exposed.data(), vmatter.data(), normals.data(), c.error_word());
    OSH_LAUNCH(flip_veto<2>, ncands, c, cands.data(), ev2v, v2s.a2ab.data(), v2s.ab2b.data(),
        v2s.codes.data(), sv2v.data(), m.coords.data(), normals.data(), smatter.data(),
        vmatter.data(), codes.data());
  }
}

bool coarsen(Ctx& c, DMesh& m, AdaptOptions const& o, Bytes& codes, bool overshoot_allowed,
    bool improve) {
  int dim = m.dim;
  LO ne = m.nents[1];
  LO nv = m.nents[0];
  int nc = m.metric_ncomps;
  int md = (nc == 1) ? 1 : dim;
  LO const* ev2v = m.down[1].data();
  OSH_LAUNCH(check_class_kernel, ne, c, ev2v, m.class_dim[0].data(), m.class_dim[1].data(),
      codes.data());
  // the candidate edges: every later sweep runs over this list
  LOs cands = collect_marked(c, codes);
  LO ncands = cands.size();
  if (ncands == 0) return false;
  // upward rows only where the candidates look: the candidate edges' cells
  // (read by order-free tests), the collapsing vertices' edges (ordered:
  // choose_rails keeps the first best rail) and cells (order-free minimum)
  Bytes vmarks = fill_i8(c, nv, 0);
  OSH_LAUNCH(mark_collapsing_verts, ncands, c, cands.data(), ev2v, codes.data(), vmarks.data());
  Bytes eveto = fill_i8(c, 2 * ne, 0);
  for (int cell_dim = 2; cell_dim <= dim; ++cell_dim) {
    LO const* c2e = ask_down(c, m, cell_dim, 1, nullptr).data();
    LO const* cv2v = verts_of(c, m, cell_dim).data();
    if (cell_dim == 3)
      OSH_LAUNCH(check_exposure_cells<3>, m.nents[3], c, c2e, cv2v, m.down[3].data(),
          m.class_dim[3].data(), m.class_dim[2].data(), ev2v, codes.data(), eveto.data());
    else
      OSH_LAUNCH(check_exposure_cells<2>, m.nents[2], c, c2e, cv2v, m.down[2].data(),
          m.class_dim[2].data(), m.class_dim[1].data(), ev2v, codes.data(), eveto.data());
  }
  OSH_LAUNCH(apply_edge_vetoes, ncands, c, cands.data(), eveto.data(), codes.data());
  Real max_length = overshoot_allowed ? o.max_length_allowed : o.max_length_desired;
  Adj v2e = marked_up(c, m, 0, 1, vmarks);
  int lanes = fit_lanes(ncands * 2, c.tuning.coarsen_lanes);
  int lshift = lanes_shift(lanes);
  LO nlanes = ncands * 2 * lanes;
  Range overshoot_range(c, "prevent_overshoot");
  Bytes veto = fill_i8(c, ncands * 2, 0);
  // mixed precision: an FP32 screen first, then FP64 only for the sides it
  // left open
  bool mixed = o.precision == 1;
  Bytes need;
  if (mixed) need = fill_i8(c, ncands * 2, 0);
#define OSH_OV(D, M, F32)                                                                \
  OSH_LAUNCH((overshoot_lanes<D, M, F32>), nlanes, c, lshift, cands.data(), ev2v,          \
      v2e.a2ab.data(), v2e.ab2b.data(), v2e.codes.data(), m.coords.data(), m.metric.data(), \
      max_length, codes.data(), veto.data(), mixed ? need.data() : (I8*)nullptr)
#define OSH_OV_ALL(F32)      \
  if (dim == 3 && md == 3)   \
    OSH_OV(3, 3, F32);       \
  else if (dim == 3)         \
    OSH_OV(3, 1, F32);       \
  else if (md == 2)          \
    OSH_OV(2, 2, F32);       \
  else                       \
    OSH_OV(2, 1, F32)
  if (mixed) {
    OSH_OV_ALL(true);
  }
  OSH_OV_ALL(false);
#undef OSH_OV_ALL
#undef OSH_OV
  OSH_LAUNCH(apply_vetoes, ncands, c, cands.data(), veto.data(), codes.data());
  overshoot_range.end();
  if (o.should_prevent_coarsen_flip) prevent_coarsen_flip(c, m, cands, codes);
  // sides without a candidate left are skipped by every sweep below, and the
  // check after the quality filter returns
  Range quality_range(c, "coarsen_qualities");
  Reals quals(c, ne * 2);
  LOs const& cv2v = verts_of(c, m, dim);
  Adj v2c = marked_up(c, m, 0, dim, vmarks, nullptr, false);
  Reals partial(c, nlanes);
  Bytes unsure;
  if (mixed) unsure = fill_i8(c, ncands * 2, 0);
#define OSH_CQ(D, M, F32)                                                                \
  OSH_LAUNCH((coarsen_quality_lanes<D, M, F32>), nlanes, c, lshift, cands.data(), ev2v,     \
      cv2v.data(), v2c.a2ab.data(), v2c.ab2b.data(), v2c.codes.data(), m.coords.data(),     \
      m.metric.data(), codes.data(), mixed ? need.data() : (I8 const*)nullptr,              \
      mixed ? unsure.data() : (I8*)nullptr, partial.data())
#define OSH_CQ_ALL(F32)      \
  if (dim == 3 && md == 3)   \
    OSH_CQ(3, 3, F32);       \
  else if (dim == 3)         \
    OSH_CQ(3, 1, F32);       \
  else if (md == 2)          \
    OSH_CQ(2, 2, F32);       \
  else                       \
    OSH_CQ(2, 1, F32)
  if (mixed) {
    OSH_CQ_ALL(true);
    OSH_LAUNCH(screen_quality_sides, ncands * 2, c, lanes, cands.data(), codes.data(),
        unsure.data(), partial.data(), o.min_quality_allowed, need.data());
  }
  OSH_CQ_ALL(false);
#undef OSH_CQ_ALL
#undef OSH_CQ
  OSH_LAUNCH(combine_quality_lanes, ncands, c, lanes, cands.data(), codes.data(),
      mixed ? need.data() : (I8 const*)nullptr, partial.data(), quals.data());
  Reals vold;
  if (improve) {
    // read only at collapsing vertices, whose rows are built
    vold = Reals(c, nv);
    OSH_LAUNCH(vert_min_quality, nv, c, v2c.a2ab.data(), v2c.ab2b.data(),
        ask_qualities(c, m).data(), vold.data());
  }
  OSH_LAUNCH(filter_quals_kernel, ncands, c, cands.data(), ev2v, quals.data(),
      o.min_quality_allowed, improve ? vold.data() : (Real const*)nullptr, codes.data());
  quality_range.end();
  if (max_i8(c, codes) <= kDont) return false;
  Reals const& lengths = ask_lengths(c, m);
  Bytes vcand(c, nv);
  Reals vqual(c, nv);
  LOs vrail(c, nv);
  OSH_LAUNCH(choose_rails_kernel, nv, c, v2e.a2ab.data(), v2e.ab2b.data(), v2e.codes.data(),
      codes.data(), lengths.data(), quals.data(), vcand.data(), vqual.data(), vrail.data());
  // the independent set reads the star rows of candidate vertices only
  Adj star;
  star.a2ab = v2e.a2ab;
  star.ab2b = LOs(c, v2e.ab2b.size());
  OSH_LAUNCH(star_from_v2e, nv, c, v2e.a2ab.data(), v2e.ab2b.data(), v2e.codes.data(), ev2v,
      star.ab2b.data());
  Bytes keys = find_indset(c, m, 0, vqual, vcand, nullptr, nullptr, &star);
  LOs keys2verts = collect_marked(c, keys);
  LO nkeys = keys2verts.size();
  note_changed_parts(c, m, 0, keys2verts);
  if (o.verbosity >= 2) std::printf("coarsening %d vertices\n", int(nkeys));
  Range modify_range(c, "modify_ents");
  LOs rails2edges(c, nkeys), col_dirs(c, nkeys), keys2onto(c, nkeys);
  OSH_LAUNCH(find_rails_kernel, nkeys, c, keys2verts.data(), vrail.data(), v2e.a2ab.data(),
      v2e.ab2b.data(), v2e.codes.data(), ev2v, rails2edges.data(), col_dirs.data(),
      keys2onto.data());
  Bytes dead[4];
  dead[1] = fill_i8(c, ne, 0);
  Bytes rail_col = fill_i8(c, ne, -1);
  OSH_LAUNCH(mark_rails, nkeys, c, rails2edges.data(), col_dirs.data(), dead[1].data(),
      rail_col.data());
  for (int d = 2; d <= dim; ++d) dead[d] = fill_i8(c, m.nents[d], 0);
  for (int cell_dim = dim; cell_dim > 1; --cell_dim) {
    LO const* c2e = ask_down(c, m, cell_dim, 1, nullptr).data();
    LO const* cv2v = verts_of(c, m, cell_dim).data();
    if (cell_dim == 3)
      OSH_LAUNCH(mark_dead_cells<3>, m.nents[3], c, c2e, cv2v, m.down[3].data(), ev2v,
          rail_col.data(), dead[3].data(), dead[2].data());
    else
      OSH_LAUNCH(mark_dead_cells<2>, m.nents[2], c, c2e, cv2v, m.down[2].data(), ev2v,
          rail_col.data(), dead[2].data(), dead[1].data());
  }
  // the keys' rows of every dimension, in order: products follow them
  Adj v2d[4];
  v2d[1] = v2e;
  for (int d = 2; d <= dim; ++d) v2d[d] = marked_up(c, m, 0, d, keys);
  DMesh nw;
  nw.dim = dim;
  nw.metric_ncomps = nc;
  LOs o2n, old_lows2new_lows;
  // products per key of every dimension, numbered by one scan launch
  LOs k2p[4];
  LO nprods[4] = {0, 0, 0, 0};
  k2p[0] = fill_lo(c, nkeys + 1, 0);
  {
    LOs counts = fill_lo(c, dim * nkeys, 0);
    LiveRows rows{};
    ScanJob jobs[3];
    for (int d = 1; d <= dim; ++d) {
      rows.a2ab[d - 1] = v2d[d].a2ab.data();
      rows.ab2b[d - 1] = v2d[d].ab2b.data();
      rows.dead[d - 1] = dead[d].data();
      rows.counts[d - 1] = counts.data() + (d - 1) * nkeys;
      k2p[d] = LOs(c, nkeys + 1);
      jobs[d - 1] = {rows.counts[d - 1], kScanLO, nkeys, k2p[d].data(), slot_ptr<LO>(c, d)};
    }
    OSH_LAUNCH(live_counts, dim * nkeys * kLanes, c, nkeys, keys2verts.data(), rows);
    offset_scans(c, jobs, dim);
    c.sync();
    for (int d = 1; d <= dim; ++d) nprods[d] = read_slot<LO>(c, d);
  }
  // every dimension's new numbering by one more scan launch
  ModPlan plans[4];
  modify_plans(c, m, 0, 0, dim, keys2verts, k2p, plans);
  modify_scan(c, plans, dim + 1);
  for (int d = 0; d <= dim; ++d) {
    LOs pv, psrc;
    if (d > 0) {
      pv = LOs(c, nprods[d] * (d + 1));
      psrc = LOs(c, nprods[d]);
      LOs const& dv2v = verts_of(c, m, d);
      OSH_LAUNCH(coarsen_products, nkeys * kLanes, c, d + 1, keys2verts.data(), keys2onto.data(),
          v2d[d].a2ab.data(), v2d[d].ab2b.data(), v2d[d].codes.data(), dead[d].data(),
          dv2v.data(), o2n.data(), k2p[d].data(), pv.data(), psrc.data());
    }
    ModOut mo;
    modify_build(c, m, nw, plans[d], keys2verts, k2p[d], nprods[d], pv, old_lows2new_lows, mo,
        d > 0 ? &o2n : nullptr);
    if (d == 0) {
      o2n = mo.old2new;
      nw.coords = transfer<Real>(c, mo, dim, m.coords, nullptr);
      nw.metric = transfer<Real>(c, mo, nc, m.metric, nullptr);
      for (auto const& t : m.vtags)
        nw.vtags.push_back(t.with_data(transfer<Real>(c, mo, t.ncomps, t.data, nullptr)));
      nw.itags = m.itags;
      for (std::size_t t = 0; t < m.itags.size(); ++t)
        nw.itags[t].data[0] = transfer<Real>(c, mo, m.itags[t].ncomps, m.itags[t].data[0], nullptr);
      nw.class_id[0] = transfer<LO>(c, mo, 1, m.class_id[0], nullptr);
      nw.class_dim[0] = transfer<I8>(c, mo, 1, m.class_dim[0], nullptr);
    } else {
      transfer_class(c, m, nw, d, mo, psrc, d, nullptr);
      if (d == dim) {
        transfer_pointwise(c, m, nw, mo, 0, keys2verts, k2p[d], pv);
        conserve_pass(c, m, nw, 1, 0, keys2verts, k2p[d], mo);
      }
      transfer_measures(c, m, nw, d, mo);
    }
    old_lows2new_lows = mo.old2new;
  }
  m = nw;
  return true;
}

bool coarsen_verts(Ctx& c, DMesh& m, AdaptOptions const& o, Bytes const& vmarks,
    bool overshoot_allowed, bool improve) {
  LO ne = m.nents[1];
  Bytes codes(c, ne);
  OSH_LAUNCH(edge_codes_from_verts, ne, c, m.down[1].data(), vmarks.data(), codes.data());
  return coarsen(c, m, o, codes, overshoot_allowed, improve);
}

}  // namespace

bool coarsen_by_size(Ctx& c, DMesh& m, AdaptOptions const& o) {
  LO ne = m.nents[1];
  Reals const& lengths = ask_lengths(c, m);
  ask_qualities(c, m);
  Bytes cand(c, ne);
  OSH_LAUNCH(lt_kernel, ne, c, lengths.data(), o.min_length_desired, cand.data());
  if (max_i8(c, cand) != 1) return false;
  Bytes vm = mark_down(c, m, 1, 0, cand);
  return coarsen_verts(c, m, o, vm, false, false);
}

bool coarsen_slivers(Ctx& c, DMesh& m, AdaptOptions const& o) {
  ask_lengths(c, m);
  Bytes marks = mark_sliver_layers(c, m, o.min_quality_desired, o.nsliver_layers);
  Bytes vm = mark_down(c, m, m.dim, 0, marks);
  return coarsen_verts(c, m, o, vm, true, true);
}

void coarsen_qualities(Ctx& c, DMesh& m, LO ncands, LO const* cands, I8 const* cand_codes,
    Real* out, int precision, Real qmin) {
  int dim = m.dim;
  int md = (m.metric_ncomps == 1) ? 1 : dim;
  if (ncands == 0) return;
  LO const* ev2v = m.down[1].data();
  LOs const& cv2v = verts_of(c, m, dim);
  Adj const& v2c = ask_up(c, m, 0, dim);
  Bytes codes = fill_i8(c, m.nents[1], kDont);
  OSH_LAUNCH(scatter_codes, ncands, c, cands, cand_codes, codes.data());
  int lanes = fit_lanes(ncands * 2, c.tuning.coarsen_lanes);
  int lshift = lanes_shift(lanes);
  LO nlanes = ncands * 2 * lanes;
  Reals partial(c, nlanes);
  bool mixed = precision == 1;
  Bytes need, unsure;
  if (mixed) {
    need = Bytes(c, ncands * 2);
    unsure = fill_i8(c, ncands * 2, 0);
  }
#define OSH_CQ(D, M, F32)                                                                \
  OSH_LAUNCH((coarsen_quality_lanes<D, M, F32>), nlanes, c, lshift, cands, ev2v,            \
      cv2v.data(), v2c.a2ab.data(), v2c.ab2b.data(), v2c.codes.data(), m.coords.data(),     \
      m.metric.data(), codes.data(), mixed ? need.data() : (I8 const*)nullptr,              \
      mixed ? unsure.data() : (I8*)nullptr, partial.data())
#define OSH_CQ_ALL(F32)      \
  if (dim == 3 && md == 3)   \
    OSH_CQ(3, 3, F32);       \
  else if (dim == 3)         \
    OSH_CQ(3, 1, F32);       \
  else if (md == 2)          \
    OSH_CQ(2, 2, F32);       \
  else                       \
    OSH_CQ(2, 1, F32)
  if (mixed) {
    OSH_CQ_ALL(true);
    OSH_LAUNCH(screen_quality_sides, ncands * 2, c, lanes, cands, codes.data(), unsure.data(),
        partial.data(), qmin, need.data());
  }
  OSH_CQ_ALL(false);
#undef OSH_CQ_ALL
#undef OSH_CQ
  OSH_LAUNCH(combine_to_cands, ncands, c, lanes, cands, codes.data(),
      mixed ? need.data() : (I8 const*)nullptr, partial.data(), out);
}

}  // namespace osh

// Edge-swap pass on the device (3D edge swap over edge loops of up to 7
// tets, 2D edge flip).
//
//   swap_edges / swap_part1 / filter_swap_improve src/Omega_h_swap.cpp:14-55
//   swap3d_ghosted / swap3d_element_based        src/Omega_h_swap3d.cpp:14-84
//   swap3d_qualities                             src/Omega_h_swap3d_qualities.cpp:11-75
//   find_loop                                    src/Omega_h_swap3d_loop.hpp:24-78
//   choose                                       src/Omega_h_swap3d_choice.hpp:16-100
//   swap3d_keys_to_prods / swap3d_topology       src/Omega_h_swap3d_topology.cpp:12-119
//   swap2d_qualities / swap2d_topology           src/Omega_h_swap2d_qualities.cpp:16-72,
//                                                src/Omega_h_swap2d_topology.cpp:11-79
//   transfer_swap / transfer_copy                src/Omega_h_transfer.cpp:631-730
#include <cstdio>

#include "osh_math.hpp"
#include "osh_passes.hpp"
#include "osh_swap_tables.hpp"

namespace osh {

namespace {

using swap::Tables;

constexpr int kLoopLanes = 8;  // at least swap::kMaxLoop

struct Loop {
  int size;
  LO eev2v[2];
  LO verts[swap::kMaxLoop];
};

// a ring of tets that does not close is a broken mesh: the reference aborts
// (src/Omega_h_swap3d_loop.hpp:71); here the loop is marked broken (size
// -1) and the caller raises kBrokenSwapLoop
__device__ void find_loop(LO edge, LO const* a2ab, LO const* ab2b, I8 const* codes,
    LO const* ev2v, LO const* tv2v, Loop& loop) {
  LO begin = a2ab[edge], end = a2ab[edge + 1];
  loop.size = end - begin;
  if (loop.size > swap::kMaxLoop) return;
  loop.eev2v[0] = ev2v[edge * 2];
  loop.eev2v[1] = ev2v[edge * 2 + 1];
  LO tmp[swap::kMaxLoop][2];
  for (int le = 0; le < loop.size; ++le) {
    LO tet = ab2b[begin + le];
    int rre = code_which_down(codes[begin + le]);
    int rot = code_rotation(codes[begin + le]);
    int rre_opp = opposite_template(3, 1, rre);
    for (int eev = 0; eev < 2; ++eev)
      tmp[le][eev] = tv2v[tet * 4 + down_template(3, 1, rre_opp, rot ^ eev)];
  }
  for (int i = 0; i < loop.size - 1; ++i) {
    int j;
    for (j = i + 1; j < loop.size; ++j)
      if (tmp[j][0] == tmp[i][1]) break;
    if (j >= loop.size) {
      loop.size = -1;
      return;
    }
    LO t0 = tmp[i + 1][0], t1 = tmp[i + 1][1];
    tmp[i + 1][0] = tmp[j][0];
    tmp[i + 1][1] = tmp[j][1];
    tmp[j][0] = t0;
    tmp[j][1] = t1;
  }
  for (int lv = 0; lv < loop.size; ++lv) loop.verts[lv] = tmp[lv][0];
}

template <int M>
__device__ double tet_quality(LO const* v, Real const* coords, Real const* metric) {
  constexpr int nc = (M == 1) ? 1 : 6;
  double p[4][3];
  double m[4][nc];
  double const* pp[4];
  double const* ms[4];
  for (int k = 0; k < 4; ++k) {
    for (int j = 0; j < 3; ++j) p[k][j] = coords[v[k] * 3 + j];
    for (int j = 0; j < nc; ++j) m[k][j] = metric[v[k] * nc + j];
    pp[k] = p[k];
    ms[k] = m[k];
  }
  return element_quality<3, M>(pp, ms);
}

// the FP32 screen of tet_quality (mixed precision); *unsure as in
// f32::element_quality
template <int M>
__device__ float tet_quality32(LO const* v, Real const* coords, Real const* metric,
    bool* unsure) {
  constexpr int nc = (M == 1) ? 1 : 6;
  double p[4][3];
  double m[4][nc];
  double const* pp[4];
  double const* ms[4];
  for (int k = 0; k < 4; ++k) {
    for (int j = 0; j < 3; ++j) p[k][j] = coords[v[k] * 3 + j];
    for (int j = 0; j < nc; ++j) m[k][j] = metric[v[k] * nc + j];
    pp[k] = p[k];
    ms[k] = m[k];
  }
  return f32::element_quality<3, M>(pp, ms, unsure);
}

template <int S, int M>
__device__ float length_between32(LO a, LO b, Real const* coords, Real const* metric) {
  constexpr int nc = (M == 1) ? 1 : (M == 2 ? 3 : 6);
  double p0[S], p1[S], m0[nc], m1[nc];
  for (int k = 0; k < S; ++k) {
    p0[k] = coords[a * S + k];
    p1[k] = coords[b * S + k];
  }
  for (int k = 0; k < nc; ++k) {
    m0[k] = metric[a * nc + k];
    m1[k] = metric[b * nc + k];
  }
  return f32::edge_length<S, M>(p0, p1, m0, m1);
}

template <int S, int M>
__device__ double length_between(LO a, LO b, Real const* coords, Real const* metric) {
  constexpr int nc = (M == 1) ? 1 : (M == 2 ? 3 : 6);
  double p0[S], p1[S], m0[nc], m1[nc];
  for (int k = 0; k < S; ++k) {
    p0[k] = coords[a * S + k];
    p1[k] = coords[b * S + k];
  }
  for (int k = 0; k < nc; ++k) {
    m0[k] = metric[a * nc + k];
    m1[k] = metric[b * nc + k];
  }
  return edge_length<S, M>(p0, p1, m0, m1);
}

// swap3d_qualities (src/Omega_h_swap3d_qualities.cpp:11-75) in four flat
// sweeps instead of one thread per candidate edge:
//   swap3d_loops:     the loop of each candidate and its counts of unique
//                     triangles and unique interior edges
//   swap3d_loop_measures: one thread per (candidate, unique triangle): the
//                     min quality of the two tets on that triangle, and one
//                     per (candidate, unique edge): its length
//   swap3d_mesh_quals: one thread per (candidate, triangulation): its value
//                     for choose() of src/Omega_h_swap3d_choice.hpp
//   swap3d_choose:    per candidate, the choice over those values in the
//                     reference's order
// Every value is the one the reference computes (each unique triangle and
// edge evaluated once), so the choice is the same.
__global__ void swap3d_loops(LO nt, LO const* cands, LO const* e2t_a2ab, LO const* e2t_ab2b,
    I8 const* e2t_codes, LO const* ev2v, LO const* tv2v, Tables const* tb, I8 const* cached,
    LO* lsize, LO* leev, LO* lverts, LO* ntris, LO* nedges, int* error) {
  // kLoopLanes threads per candidate: lane le reads the le-th tet around
  // the edge; the group then orders the loop as find_loop does
  LO t = thread_index();
  if (t >= nt) return;
  LO cand = t / kLoopLanes;
  int lane = t % kLoopLanes;
  LO edge = cands[cand];
  LO begin = e2t_a2ab[edge];
  int size = e2t_a2ab[edge + 1] - begin;
  LO x0 = -1, x1 = -1;
  if (size <= swap::kMaxLoop && lane < size) {
    LO tet = e2t_ab2b[begin + lane];
    I8 code = e2t_codes[begin + lane];
    int rre_opp = opposite_template(3, 1, code_which_down(code));
    int rot = code_rotation(code);
    x0 = tv2v[tet * 4 + down_template(3, 1, rre_opp, rot)];
    x1 = tv2v[tet * 4 + down_template(3, 1, rre_opp, rot ^ 1)];
  }
  int base = int(threadIdx.x % 64) & ~(kLoopLanes - 1);
  LO tmp[swap::kMaxLoop][2];
#pragma unroll
  for (int k = 0; k < swap::kMaxLoop; ++k) {
    tmp[k][0] = __shfl(x0, base + k, 64);
    tmp[k][1] = __shfl(x1, base + k, 64);
  }
  if (lane != 0) return;
  lsize[cand] = size;
  if (size > swap::kMaxLoop) {
    ntris[cand] = 0;
    nedges[cand] = 0;
    return;
  }
  for (int i = 0; i < size - 1; ++i) {
    int j;
    for (j = i + 1; j < size; ++j)
      if (tmp[j][0] == tmp[i][1]) break;
    if (j >= size) {
      // the tets around the edge do not close a ring (the reference aborts,
      // src/Omega_h_swap3d_loop.hpp:71): no configuration, and an error
      *error = kBrokenSwapLoop;
      ntris[cand] = 0;
      nedges[cand] = 0;
      lsize[cand] = swap::kMaxLoop + 1;
      return;
    }
    LO t0 = tmp[i + 1][0], t1 = tmp[i + 1][1];
    tmp[i + 1][0] = tmp[j][0];
    tmp[i + 1][1] = tmp[j][1];
    tmp[j][0] = t0;
    tmp[j][1] = t1;
  }
  leev[cand * 2 + 0] = ev2v[edge * 2];
  leev[cand * 2 + 1] = ev2v[edge * 2 + 1];
  for (int k = 0; k < swap::kMaxLoop; ++k)
    lverts[cand * swap::kMaxLoop + k] = (k < size) ? tmp[k][0] : -1;
  // a carried-over choice needs the loop (for products) but no qualities
  bool carried = cached && cached[edge];
  ntris[cand] = carried ? 0 : tb->ntris[size];
  nedges[cand] = carried ? 0 : tb->nuedges[size];
}

// F32 (mixed precision): FP32 estimates, and a candidate with an element
// whose max-determinant metric is not clear is marked in `unsure`; with
// `need` given, the FP64 sweep evaluates only the marked candidates
// the loop measures of the candidates in one launch: threads [0, ntp) take
// the (candidate, unique triangle) slots, the rest the (candidate, unique
// edge) slots
struct LoopMeasures {
  LO ntp;
  LO const* ntris;
  LO const* nedges;
  LO const* lsize;
  LO const* leev;
  LO const* lverts;
  Real const* coords;
  Real const* metric;
  Tables const* tb;
  I8 const* need;
  I8* unsure;
  Real* tq;
  Real* el;
};

template <int M, bool F32>
__device__ __forceinline__ void swap3d_tri_qual(LO p, LO const* ntris, LO const* lsize,
    LO const* leev, LO const* lverts, Real const* coords, Real const* metric, Tables const* tb,
    I8 const* need, I8* unsure, Real* tq) {
  // a fixed stride of kMaxUniqueTris slots per candidate
  LO cand = p / swap::kMaxUniqueTris;
  int ut = p % swap::kMaxUniqueTris;
  if (ut >= ntris[cand]) return;
  if (!F32 && need && !need[cand]) return;
  int n = lsize[cand];
  LO tv[4];
  for (int k = 0; k < 3; ++k) tv[k] = lverts[cand * swap::kMaxLoop + tb->tris[n][ut][k]];
  double tri_minqual = 1.0;
  float minq32 = 1.0f;
  bool unclear = false;
  for (int tri_tet = 0; tri_tet < 2; ++tri_tet) {
    tv[3] = leev[cand * 2 + 1 - tri_tet];
    if constexpr (F32) {
      float q = tet_quality32<M>(tv, coords, metric, &unclear);
      minq32 = (q < minq32) ? q : minq32;
    } else {
      tri_minqual = math::min2(tri_minqual, tet_quality<M>(tv, coords, metric));
    }
    LO t = tv[1];
    tv[1] = tv[2];
    tv[2] = t;
  }
  if constexpr (F32) {
    if (unclear) unsure[cand] = 1;
    tq[p] = double(minq32);
  } else {
    tq[p] = tri_minqual;
  }
}

template <int M, bool F32>
__device__ __forceinline__ void swap3d_edge_len(LO q, LO const* nedges, LO const* lsize,
    LO const* lverts, Real const* coords, Real const* metric, Tables const* tb, I8 const* need,
    Real* el) {
  // a fixed stride of kMaxUniqueEdges slots per candidate
  LO cand = q / swap::kMaxUniqueEdges;
  int ue = q % swap::kMaxUniqueEdges;
  if (ue >= nedges[cand]) return;
  if (!F32 && need && !need[cand]) return;
  int n = lsize[cand];
  LO const* lv = lverts + cand * swap::kMaxLoop;
  LO a = lv[tb->uedges[n][ue][0]], b = lv[tb->uedges[n][ue][1]];
  if constexpr (F32)
    el[q] = double(length_between32<3, M>(a, b, coords, metric));
  else
    el[q] = length_between<3, M>(a, b, coords, metric);
}

template <int M, bool F32>
__global__ void swap3d_loop_measures(LO nt, LoopMeasures a) {
  LO t = thread_index();
  if (t >= nt) return;
  if (t < a.ntp)
    swap3d_tri_qual<M, F32>(t, a.ntris, a.lsize, a.leev, a.lverts, a.coords, a.metric, a.tb,
        a.need, a.unsure, a.tq);
  else
    swap3d_edge_len<M, F32>(t - a.ntp, a.nedges, a.lsize, a.lverts, a.coords, a.metric, a.tb,
        a.need, a.el);
}

// one thread per (candidate, triangulation) on a fixed stride of
// kMaxMeshes: the triangulation's minimum triangle quality, or -1 when one of
// its interior edges is longer than max_length.  choose() keeps the first
// triangulation whose quality exceeds every earlier kept one (starting from
// 0) and whose edges pass, so it equals the first triangulation attaining
// the largest of these values, when that is positive; a minimum that went
// negative (where choose() stops early) is never kept either way.
constexpr int kMaxMeshes = 42;

__global__ void swap3d_mesh_quals(LO nt, LO const* lsize, LO const* ntris, Real const* tq,
    Real const* el, Real max_length, Tables const* tb, I8 const* need, Real* mq) {
  LO t = thread_index();
  if (t >= nt) return;
  LO cand = t / kMaxMeshes;
  int mesh = t % kMaxMeshes;
  int n = lsize[cand];
  if (n > swap::kMaxLoop || mesh >= tb->mesh_counts[n]) return;
  if (ntris[cand] == 0) return;  // a carried-over choice
  if (need && !need[cand]) return;  // rejected by the mixed-precision screen
  int nmesh_tris = tb->mesh_sizes[n];
  int nmesh_edges = tb->nedges[n];
  Real const* ctq = tq + LO(cand) * swap::kMaxUniqueTris;
  Real const* cel = el + LO(cand) * swap::kMaxUniqueEdges;
  double q = 1.0;
  for (int mt = 0; mt < nmesh_tris; ++mt)
    q = math::min2(q, ctq[tb->meshes[n][mesh * nmesh_tris + mt]]);
  for (int me = 0; me < nmesh_edges; ++me)
    if (cel[tb->medges[n][mesh * nmesh_edges + me]] > max_length) q = -1.0;
  mq[t] = q;
}

// the mixed-precision screen of one candidate (Tuning::choose_lanes threads): the
// best FP32 triangulation value, where an interior edge counts as too long
// only when its FP32 length exceeds max_length by more than the band, is an
// upper estimate of the FP64 choice's quality; the candidate needs FP64
// unless that estimate is below the edge's current minimum quality by more
// than the band (with every element's metric clear and that minimum
// positive)
__global__ void swap3d_screen(LO nt, int lanes, LO const* cands, LO const* lsize,
    LO const* ntris, Real const* tq, Real const* el, Real max_length, Tables const* tb,
    LO const* e2t_a2ab, LO const* e2t_ab2b, Real const* elem_quals, I8 const* unsure,
    I8* need) {
  LO t = thread_index();
  if (t >= nt) return;
  LO cand = t / lanes;
  int lane = t % lanes;
  LO edge = cands[cand];
  int n = lsize[cand];
  bool open = ntris[cand] != 0 && n <= swap::kMaxLoop;  // not carried over, a loop
  double best = -1.0;
  if (open) {
    int nmesh_tris = tb->mesh_sizes[n];
    int nmesh_edges = tb->nedges[n];
    Real const* ctq = tq + LO(cand) * swap::kMaxUniqueTris;
    Real const* cel = el + LO(cand) * swap::kMaxUniqueEdges;
    double lmax = max_length * (1.0 + double(f32::kScreenBand));
    for (int mesh = lane; mesh < tb->mesh_counts[n]; mesh += lanes) {
      double q = 1.0;
      for (int mt = 0; mt < nmesh_tris; ++mt)
        q = math::min2(q, ctq[tb->meshes[n][mesh * nmesh_tris + mt]]);
      for (int me = 0; me < nmesh_edges; ++me)
        if (cel[tb->medges[n][mesh * nmesh_edges + me]] > lmax) q = -1.0;
      best = math::max2(best, q);
    }
  }
  double old = 1.79769313486231570815e+308;
  for (LO j = e2t_a2ab[edge] + lane; j < e2t_a2ab[edge + 1]; j += lanes)
    old = math::min2(old, elem_quals[e2t_ab2b[j]]);
  for (int d = lanes / 2; d > 0; d >>= 1) {
    best = math::max2(best, __shfl_xor(best, d, 64));
    old = math::min2(old, __shfl_xor(old, d, 64));
  }
  if (lane != 0) return;
  bool reject = open && !unsure[cand] && old > 0.0 &&
                best < old * (1.0 - double(f32::kScreenBand));
  need[cand] = reject ? 0 : 1;
}

__global__ void swap3d_choose(LO nt, int lanes, LO const* cands, LO const* lsize,
    Real const* mq, LO const* e2t_a2ab, LO const* e2t_ab2b, Real const* elem_quals,
    Tables const* tb, I8 const* need, I8* configs, Real* quals, I8* keep, I8* c_valid,
    I8* c_cfg, I8* c_keep, Real* c_q) {
  // `lanes` threads (a power of two, Tuning::choose_lanes) per candidate
  // share its triangulations and tets
  LO t = thread_index();
  if (t >= nt) return;
  LO cand = t / lanes;
  int lane = t % lanes;
  LO edge = cands[cand];
  if (c_valid && c_valid[edge]) {
    if (lane == 0) {
      configs[edge] = c_cfg[edge];
      if (c_keep[edge]) {
        keep[edge] = 1;
        quals[edge] = c_q[edge];
      }
    }
    return;
  }
  if (need && !need[cand]) {
    // the mixed-precision screen showed that no triangulation improves on
    // the edge's elements: not kept, as the FP64 choice would decide
    if (lane == 0) {
      configs[edge] = -1;
      if (c_valid) {
        c_cfg[edge] = -1;
        c_keep[edge] = 0;
        c_q[edge] = -1.0;
        c_valid[edge] = 1;
      }
    }
    return;
  }
  int n = lsize[cand];
  // each lane keeps the first of its triangulations with the largest value
  // above 0; the group then takes the largest value, the lowest index on
  // ties: the first triangulation attaining the maximum, as the sequential
  // choice finds it
  int choice_mesh = -1;
  double choice_quality = -1.0;
  if (n <= swap::kMaxLoop) {
    choice_quality = 0.0;
    int nmeshes = tb->mesh_counts[n];
    Real const* cmq = mq + LO(cand) * kMaxMeshes;
    for (int mesh = lane; mesh < nmeshes; mesh += lanes) {
      double q = cmq[mesh];
      if (q > choice_quality) {
        choice_mesh = mesh;
        choice_quality = q;
      }
    }
  }
  double old = 1.79769313486231570815e+308;
  for (LO j = e2t_a2ab[edge] + lane; j < e2t_a2ab[edge + 1]; j += lanes)
    old = math::min2(old, elem_quals[e2t_ab2b[j]]);
  for (int d = lanes / 2; d > 0; d >>= 1) {
    double oq = __shfl_xor(choice_quality, d, 64);
    int om = __shfl_xor(choice_mesh, d, 64);
    if (oq > choice_quality || (oq == choice_quality && om >= 0 &&
                                   (choice_mesh < 0 || om < choice_mesh))) {
      choice_quality = oq;
      choice_mesh = om;
    }
    old = math::min2(old, __shfl_xor(old, d, 64));
  }
  if (lane != 0) return;
  configs[edge] = I8(choice_mesh);
  // a key needs a triangulation; with inverted elements the reference can
  // keep an edge without one and then index its tables out of bounds
  bool kept = choice_mesh >= 0 && choice_quality > old;
  if (kept) {
    keep[edge] = 1;
    quals[edge] = choice_quality;
  }
  if (c_valid) {
    c_cfg[edge] = I8(choice_mesh);
    c_keep[edge] = I8(kept);
    c_q[edge] = choice_quality;
    c_valid[edge] = 1;
  }
}

// reflect_down for the products of one swap key, searched locally: a side
// of a product (a face of a new tet, an edge of a new face) is a product of
// the same key one dimension down or a surviving side of one of the old tets
// around the key edge, so those are the only candidates (vertices keep their
// numbers in a swap).  `lanes` threads (a power of two up to 64,
// Tuning::side_lanes) share one side's candidates; the one that matches
// writes the side and the code match_uses would compute.
template <int HIGH>
__global__ void swap3d_local_down(LO nt, int lanes, LO const* prod_key, LO const* keys2edges,
    LO const* k2p_low, LO const* pv_high, LO const* pv_low, LO const* low_prods2new,
    LO const* e2t_a2ab, LO const* e2t_ab2b, LO const* t2low, LO const* old_low_verts,
    LO const* low_old2new, LO* out, I8* out_codes, int* error) {
  LO t = thread_index();
  if (t >= nt) return;
  LO u = t / lanes;
  int lane = t % lanes;
  constexpr int LOW = HIGH - 1;
  constexpr int nvh = HIGH + 1;
  constexpr int nvl = LOW + 1;
  constexpr int nlpt = (HIGH == 3) ? 4 : 6;  // faces or edges of an old tet
  LO p = u / nvh;
  int which = u % nvh;
  LO key = prod_key[p];
  LO av[3];
#pragma unroll
  for (int k = 0; k < nvl; ++k) av[k] = pv_high[p * nvh + down_template(HIGH, LOW, which, k)];
  LO q0 = k2p_low[key];
  LO nq = k2p_low[key + 1] - q0;
  LO e = keys2edges[key];
  LO r0 = e2t_a2ab[e];
  LO ncand = nq + (e2t_a2ab[e + 1] - r0) * nlpt;
  bool found = false;
  for (LO cnd = lane; cnd < ncand; cnd += lanes) {
    LO id;
    LO bv[3];
    if (cnd < nq) {
      id = low_prods2new[q0 + cnd];
#pragma unroll
      for (int k = 0; k < nvl; ++k) bv[k] = pv_low[(q0 + cnd) * nvl + k];
    } else {
      LO r = cnd - nq;
      LO o = t2low[e2t_ab2b[r0 + r / nlpt] * nlpt + r % nlpt];
      id = low_old2new[o];
      if (id < 0) continue;
#pragma unroll
      for (int k = 0; k < nvl; ++k) bv[k] = old_low_verts[o * nvl + k];
    }
    int wd = (bv[0] == av[0]) ? 0 : (bv[1] == av[0]) ? 1 : (nvl == 3 && bv[2] == av[0]) ? 2 : -1;
    if (wd < 0) continue;
    I8 code;
    if (nvl == 2) {
      if (av[1] != bv[1 - wd]) continue;
      code = make_code(0, wd, 0);
    } else if (av[1] == bv[(wd + 1) % 3] && av[2] == bv[(wd + 2) % 3]) {
      code = make_code(0, (3 - wd) % 3, 0);
    } else if (av[1] == bv[(wd + 2) % 3] && av[2] == bv[(wd + 1) % 3]) {
      code = make_code(1, (3 - wd) % 3, 0);
    } else {
      continue;
    }
    out[u] = id;
    out_codes[u] = code;
    found = true;
  }
  // the side's lanes are `lanes` consecutive lanes of one wave
  unsigned long long any = __ballot(found);
  int base = int(threadIdx.x % 64) & ~(lanes - 1);
  unsigned long long group = (lanes >= 64) ? ~0ull : ((1ull << lanes) - 1);
  if (lane == 0 && ((any >> base) & group) == 0ull) {
    *error = kReflectDownNoMatch;
    out[u] = 0;
    out_codes[u] = 0;
  }
}

// the edges of the tets the pass replaces: their tet rings change
__global__ void swap3d_mark_dirty(LO ntets, I8 const* tet_same, LO const* t2e, I8* dirty) {
  LO t = thread_index();
  if (t >= ntets || tet_same[t]) return;
  for (int i = 0; i < 6; ++i) dirty[t2e[t * 6 + i]] = 1;
}

// the carried-over choices of the same edges whose tets did not change
// threads [0, nsame) carry the same edges' choices; the rest mark the new
// edges (products) as having none
__global__ void swap3d_cache_transfer(LO n, LO nsame, LO const* same2old, LO const* same2new,
    LO const* prods2new, I8 const* dirty, I8 const* valid, I8 const* cfg, I8 const* keep,
    Real const* q, I8* n_valid, I8* n_cfg, I8* n_keep, Real* n_q) {
  LO i = thread_index();
  if (i >= n) return;
  if (i >= nsame) {
    n_valid[prods2new[i - nsame]] = 0;
    return;
  }
  LO o = same2old[i], w = same2new[i];
  bool ok = valid[o] && !dirty[o];
  n_valid[w] = I8(ok);
  if (ok) {
    n_cfg[w] = cfg[o];
    n_keep[w] = keep[o];
    n_q[w] = q[o];
  }
}

// the interior edges among the marked ones (mark_by_class_dim), and the
// swap choice arrays set to "none"
__global__ void and_interior(LO ne, int dim, I8 const* edim, I8* marks, I8* configs, I8* keep,
    Real* quals) {
  LO e = thread_index();
  if (e >= ne) return;
  marks[e] = I8(marks[e] && (edim[e] == dim));
  if (configs) configs[e] = -1;
  keep[e] = 0;
  quals[e] = -1.0;
}

// per edge, the edges, faces and tets a key edge's swap produces (0 for
// other edges), so that one scan launch numbers the keys and their products
__global__ void swap3d_counts(LO nedges, I8 const* keys, LO const* e2t_a2ab, Tables const* tb,
    LO* ne, LO* nf, LO* nt) {
  LO e = thread_index();
  if (e >= nedges) return;
  int npt = 0, npe = 0;
  if (keys[e]) {
    int ls = e2t_a2ab[e + 1] - e2t_a2ab[e];
    npt = tb->mesh_sizes[ls];
    npe = tb->nedges[ls];
  }
  ne[e] = npe;
  nf[e] = npt + 2 * npe;
  nt[e] = 2 * npt;
}

// the key list, each key's product offsets and its representatives in the
// rebuild of faces and tets, from the per-edge scans and sweeps
__global__ void swap3d_key_offsets(LO nedges, I8 const* keys, LO const* key_off,
    LO const* eoff, LO const* foff, LO const* toff, LO const* frep_of, LO const* trep_of,
    LO* keys2edges, LO* k2e, LO* k2f, LO* k2t, LO* ereps, LO* freps, LO* treps) {
  LO e = thread_index();
  if (e >= nedges) return;
  if (e == 0) {
    LO nk = key_off[nedges];
    k2e[nk] = eoff[nedges];
    k2f[nk] = foff[nedges];
    k2t[nk] = toff[nedges];
  }
  if (!keys[e]) return;
  LO k = key_off[e];
  keys2edges[k] = e;
  k2e[k] = eoff[e];
  k2f[k] = foff[e];
  k2t[k] = toff[e];
  ereps[k] = e;
  freps[k] = frep_of[e];
  treps[k] = trep_of[e];
}

// swap3d_topology (src/Omega_h_swap3d_topology.cpp:12-119): each key's new
// edges, faces and tets in template order, from the loop its candidate sweep
// already found; kLanes threads per key share the products
constexpr int kLanes = 8;

__global__ void swap3d_products(LO nt, LO const* keys2edges, I8 const* configs,
    LO const* cand_of_edge, LO const* lsize, LO const* leev, LO const* lverts, Tables const* tb,
    LO const* k2e, LO const* k2f, LO const* k2t, LO* pe, LO* pf, LO* pt, LO* fkey, LO* tkey) {
  LO t = thread_index();
  if (t >= nt) return;
  LO k = t / kLanes;
  int lane = t % kLanes;
  LO edge = keys2edges[k];
  int config = configs[edge];
  LO cand = cand_of_edge[edge];
  int n = lsize[cand];
  LO const* lv = lverts + cand * swap::kMaxLoop;
  LO eev2v[2] = {leev[cand * 2], leev[cand * 2 + 1]};
  int npt = tb->mesh_sizes[n], npe = tb->nedges[n];
  for (int pe_i = lane; pe_i < npe; pe_i += kLanes) {
    int ue = tb->medges[n][config * npe + pe_i];
    LO a = lv[tb->uedges[n][ue][0]];
    LO b = lv[tb->uedges[n][ue][1]];
    LO pedge = k2e[k] + pe_i;
    pe[pedge * 2] = a;
    pe[pedge * 2 + 1] = b;
    for (int eev = 0; eev < 2; ++eev) {
      LO ptri = k2f[k] + 2 * pe_i + eev;
      pf[ptri * 3 + 0] = a;
      pf[ptri * 3 + 1] = b;
      pf[ptri * 3 + 2] = eev2v[eev];
      fkey[ptri] = k;
    }
  }
  for (int pt_i = lane; pt_i < npt; pt_i += kLanes) {
    int ut = tb->meshes[n][config * npt + pt_i];
    LO tv[3];
    for (int j = 0; j < 3; ++j) tv[j] = lv[tb->tris[n][ut][j]];
    LO ptri = k2f[k] + 2 * npe + pt_i;
    for (int j = 0; j < 3; ++j) pf[ptri * 3 + j] = tv[j];
    fkey[ptri] = k;
    for (int eev = 0; eev < 2; ++eev) {
      LO ptet = k2t[k] + 2 * pt_i + eev;
      tkey[ptet] = k;
      LO nt4[4] = {tv[0], tv[1], tv[2], eev2v[eev]};
      if (eev == 0) {
        LO tmp = nt4[1];
        nt4[1] = nt4[2];
        nt4[2] = tmp;
      }
      for (int j = 0; j < 4; ++j) pt[ptet * 4 + j] = nt4[j];
    }
  }
}

__global__ void loops_kernel(LO n, LO const* edges, LO const* a2ab, LO const* ab2b,
    I8 const* codes, LO const* ev2v, LO const* tv2v, LO* eev2v, LO* verts, LO* sizes,
    int* error) {
  LO i = thread_index();
  if (i >= n) return;
  Loop loop;
  find_loop(edges[i], a2ab, ab2b, codes, ev2v, tv2v, loop);
  if (loop.size < 0) {
    *error = kBrokenSwapLoop;
    sizes[i] = 0;
    return;
  }
  sizes[i] = loop.size;
  if (loop.size > swap::kMaxLoop) return;
  eev2v[i * 2] = loop.eev2v[0];
  eev2v[i * 2 + 1] = loop.eev2v[1];
  for (int k = 0; k < loop.size; ++k) verts[i * swap::kMaxLoop + k] = loop.verts[k];
}

// 2D: the quadrilateral around an interior edge and its flipped diagonal
__device__ void opposite_verts(LO e, LO const* e2t_a2ab, LO const* e2t_ab2b,
    I8 const* e2t_codes, LO const* tv2v, LO* ov) {
  for (int i = 0; i < 2; ++i) {
    LO et = e2t_a2ab[e] + i;
    int tte = code_which_down(e2t_codes[et]);
    int rot = code_rotation(e2t_codes[et]);
    LO t = e2t_ab2b[et];
    ov[rot] = tv2v[t * 3 + opposite_template(2, 1, tte)];
  }
}

template <int M>
__global__ void swap2d_qualities_kernel(LO ncands, LO const* cands, LO const* e2t_a2ab,
    LO const* e2t_ab2b, I8 const* e2t_codes, LO const* ev2v, LO const* tv2v,
    Real const* coords, Real const* metric, Real const* elem_quals, Real max_length,
    I8* keep, Real* quals) {
  LO cand = thread_index();
  if (cand >= ncands) return;
  constexpr int nc = (M == 1) ? 1 : 3;
  LO e = cands[cand];
  LO ov[2];
  opposite_verts(e, e2t_a2ab, e2t_ab2b, e2t_codes, tv2v, ov);
  double q;
  if (length_between<2, M>(ov[0], ov[1], coords, metric) > max_length) {
    q = -1.0;
  } else {
    double minqual = 1.0;
    for (int i = 0; i < 2; ++i) {
      LO ntv[3] = {ev2v[e * 2 + (1 - i)], ov[i], ov[1 - i]};
      double p[3][2];
      double m[3][nc];
      double const* pp[3];
      double const* ms[3];
      for (int k = 0; k < 3; ++k) {
        for (int j = 0; j < 2; ++j) p[k][j] = coords[ntv[k] * 2 + j];
        for (int j = 0; j < nc; ++j) m[k][j] = metric[ntv[k] * nc + j];
        pp[k] = p[k];
        ms[k] = m[k];
      }
      minqual = math::min2(minqual, element_quality<2, M>(pp, ms));
    }
    q = minqual;
  }
  double old = 1.79769313486231570815e+308;
  for (LO j = e2t_a2ab[e]; j < e2t_a2ab[e + 1]; ++j) old = math::min2(old, elem_quals[e2t_ab2b[j]]);
  if (q > old) {
    keep[e] = 1;
    quals[e] = q;
  }
}

__global__ void swap2d_products(LO nkeys, LO const* keys2edges, LO const* e2t_a2ab,
    LO const* e2t_ab2b, I8 const* e2t_codes, LO const* ev2v, LO const* tv2v, LO* pe, LO* pf) {
  LO k = thread_index();
  if (k >= nkeys) return;
  LO e = keys2edges[k];
  LO ov[2];
  opposite_verts(e, e2t_a2ab, e2t_ab2b, e2t_codes, tv2v, ov);
  pe[k * 2] = ov[0];
  pe[k * 2 + 1] = ov[1];
  for (int i = 0; i < 2; ++i) {
    LO* ntv = pf + (k * 2 + i) * 3;
    ntv[0] = ev2v[e * 2 + (1 - i)];
    ntv[1] = ov[i];
    ntv[2] = ov[1 - i];
  }
}

// the context keeps one device copy of the tables
Tables const* device_tables(Ctx& c) {
  if (!c.swap_tables) {
    Tables t = swap::make_tables();
    auto d = std::make_shared<DArray<unsigned char>>(c, LO(sizeof(Tables)));
    upload(c, *d, reinterpret_cast<unsigned char const*>(&t), LO(sizeof(Tables)));
    c.sync();
    c.swap_tables = d;
  }
  return reinterpret_cast<Tables const*>(
      static_cast<DArray<unsigned char>*>(c.swap_tables.get())->data());
}

// `plans` (when given) are the rebuild plans of dimensions 1..dim, scanned
void swap_rebuild(Ctx& c, DMesh& m, LOs const& keys2edges, LOs const* k2p, LO const* nprods,
    LOs const* pv, ModPlan* plans = nullptr, SwapCache* cache = nullptr,
    Bytes const* dirty = nullptr, Adj const* key_tets = nullptr,
    LOs const* prod_keys = nullptr) {
  int dim = m.dim;
  LO nv = m.nents[0];
  DMesh nw;
  nw.dim = dim;
  nw.metric_ncomps = m.metric_ncomps;
  nw.nents[0] = nv;
  nw.coords = m.coords;
  nw.metric = m.metric;
  nw.vtags = m.vtags;
  nw.itags = m.itags;
  nw.class_id[0] = m.class_id[0];
  nw.class_dim[0] = m.class_dim[0];
  LOs old_lows2new_lows;  // vertices keep their numbers
  ModPlan own[3];
  if (!plans) {
    plans = own;
    modify_plans(c, m, 1, 1, dim, keys2edges, k2p, plans);
    modify_scan(c, plans, dim);
  }
  ModOut low_mo;  // the previous dimension's numbering
  for (int d = 1; d <= dim; ++d) {
    ModOut mo;
    LOs src;  // each product's key edge
    // in 3D the products' sides come from a search around each key
    LOs gdown;
    Bytes gcodes;
    if (key_tets && prod_keys && d >= 2) {
      LO nuses = nprods[d] * (d + 1);
      int side_lanes = fit_lanes(nuses, c.tuning.side_lanes);
      gdown = LOs(c, nuses);
      gcodes = Bytes(c, nuses);
      LO const* t2low = (d == 3) ? m.down[3].data() : ask_down(c, m, 3, 1, nullptr).data();
      if (d == 3)
        OSH_LAUNCH(swap3d_local_down<3>, nuses * side_lanes, c, side_lanes, prod_keys[3].data(),
            keys2edges.data(), k2p[2].data(), pv[3].data(), pv[2].data(),
            low_mo.prods2new.data(), key_tets->a2ab.data(), key_tets->ab2b.data(), t2low,
            verts_of(c, m, 2).data(), low_mo.old2new.data(), gdown.data(), gcodes.data(),
            c.error_word());
      else
        OSH_LAUNCH(swap3d_local_down<2>, nuses * side_lanes, c, side_lanes, prod_keys[2].data(),
            keys2edges.data(), k2p[1].data(), pv[2].data(), pv[1].data(),
            low_mo.prods2new.data(), key_tets->a2ab.data(), key_tets->ab2b.data(), t2low,
            verts_of(c, m, 1).data(), low_mo.old2new.data(), gdown.data(), gcodes.data(),
            c.error_word());
    }
    modify_build(c, m, nw, plans[d - 1], keys2edges, k2p[d], nprods[d], pv[d],
        old_lows2new_lows, mo, nullptr, &src, gdown.exists() ? &gdown : nullptr,
        gcodes.exists() ? &gcodes : nullptr);
    transfer_class(c, m, nw, d, mo, src, 1, nullptr);
    if (d == 1 && cache) {
      SwapCache next;
      next.valid = Bytes(c, mo.nnew);
      next.cfg = Bytes(c, mo.nnew);
      next.keep = Bytes(c, mo.nnew);
      next.q = Reals(c, mo.nnew);
      LO nsame = mo.same2old.size();
      OSH_CHECK(nsame + mo.prods2new.size() == mo.nnew, "swap cache: same + new edges");
      OSH_LAUNCH(swap3d_cache_transfer, mo.nnew, c, nsame, mo.same2old.data(),
          mo.same2new.data(), mo.prods2new.data(), dirty->data(), cache->valid.data(),
          cache->cfg.data(), cache->keep.data(), cache->q.data(), next.valid.data(),
          next.cfg.data(), next.keep.data(), next.q.data());
      *cache = next;
    }
    if (d == dim) {
      transfer_pointwise(c, m, nw, mo, 1, keys2edges, k2p[d], pv[d]);
      conserve_pass(c, m, nw, 2, 1, keys2edges, k2p[d], mo);
    }
    transfer_measures(c, m, nw, d, mo);
    old_lows2new_lows = mo.old2new;
    low_mo = mo;
  }
  m = nw;
}

// sliver-layer elements' interior edges (src/Omega_h_swap.cpp:21-30); the
// same launch sets the choice arrays
Bytes swap_candidates(Ctx& c, DMesh& m, AdaptOptions const& o, Bytes* configs, Bytes& keep,
    Reals& quals) {
  int dim = m.dim;
  LO ne = m.nents[1];
  Bytes elems = mark_sliver_layers(c, m, o.min_quality_desired, o.nsliver_layers);
  Bytes edges = mark_down(c, m, dim, 1, elems);
  if (configs) *configs = Bytes(c, ne);
  keep = Bytes(c, ne);
  quals = Reals(c, ne);
  OSH_LAUNCH(and_interior, ne, c, dim, m.class_dim[1].data(), edges.data(),
      configs ? configs->data() : (I8*)nullptr, keep.data(), quals.data());
  return edges;
}

// swap3d_qualities (src/Omega_h_swap3d_qualities.cpp:11-75) with the choice
// of each candidate's triangulation (src/Omega_h_swap3d_choice.hpp): for the
// candidate edges `cands`, whose edge -> tet rows e2t holds, configs[e]
// receives the chosen triangulation (-1 for none) and, when it improves on
// the elements around the edge, keep[e] = 1 and quals[e] its quality
// (entries of other edges are left as they are).  The candidates' loops stay
// in `loops` for the products.
struct SwapLoops {
  LOs lsize, leev, lverts;
};

void swap3d_evaluate(Ctx& c, DMesh& m, AdaptOptions const& o, LOs const& cands, Adj const& e2t,
    SwapCache* cache, Bytes& configs, Reals& quals, Bytes& keep, SwapLoops& loops) {
  LO ne = m.nents[1];
  int md = (m.metric_ncomps == 1) ? 1 : 3;
  Tables const* tb = device_tables(c);
  LOs const& tv2v = verts_of(c, m, 3);
  Reals const& eq = ask_qualities(c, m);
  LO ncands = cands.size();
  int choose_lanes = fit_lanes(ncands, c.tuning.choose_lanes);
  if (cache && !cache->valid.exists()) {
    cache->valid = fill_i8(c, ne, 0);
    cache->cfg = Bytes(c, ne);
    cache->keep = Bytes(c, ne);
    cache->q = Reals(c, ne);
  }
  Range quality_range(c, "swap3d_qualities");
  loops.lsize = LOs(c, ncands);
  loops.leev = LOs(c, ncands * 2);
  loops.lverts = LOs(c, ncands * swap::kMaxLoop);
  LOs const& lsize = loops.lsize;
  LOs const& leev = loops.leev;
  LOs const& lverts = loops.lverts;
  LOs ntris(c, ncands), nedges(c, ncands);
  OSH_LAUNCH(swap3d_loops, ncands * kLoopLanes, c, cands.data(), e2t.a2ab.data(), e2t.ab2b.data(),
      e2t.codes.data(), m.down[1].data(), tv2v.data(), tb,
      cache ? cache->valid.data() : (I8 const*)nullptr, lsize.data(), leev.data(),
      lverts.data(), ntris.data(), nedges.data(), c.error_word());
  // the unique triangles and edges of all loops at fixed strides per
  // candidate, so that neither a scan nor a sync sizes them
  LO ntp = ncands * swap::kMaxUniqueTris, nep = ncands * swap::kMaxUniqueEdges;
  Reals tq(c, ntp), el(c, nep);
  // mixed precision: an FP32 screen of every candidate first, then the FP64
  // sweeps, choice included, only for the candidates it left open
  bool mixed = o.precision == 1;
  Bytes need, unsure;
  if (mixed) {
    need = Bytes(c, ncands);
    unsure = fill_i8(c, ncands, 0);
  }
  I8 const* needp = mixed ? need.data() : nullptr;
  LoopMeasures lm{ntp, ntris.data(), nedges.data(), lsize.data(), leev.data(), lverts.data(),
      m.coords.data(), m.metric.data(), tb, needp, mixed ? unsure.data() : (I8*)nullptr,
      tq.data(), el.data()};
#define OSH_SWQ(M, F32) OSH_LAUNCH((swap3d_loop_measures<M, F32>), ntp + nep, c, lm)
  if (mixed) {
    if (md == 3) {
      OSH_SWQ(3, true);
    } else {
      OSH_SWQ(1, true);
    }
    OSH_LAUNCH(swap3d_screen, ncands * choose_lanes, c, choose_lanes, cands.data(),
        lsize.data(),
        ntris.data(), tq.data(), el.data(), o.max_length_allowed, tb, e2t.a2ab.data(),
        e2t.ab2b.data(), eq.data(), unsure.data(), need.data());
  }
  if (md == 3) {
    OSH_SWQ(3, false);
  } else {
    OSH_SWQ(1, false);
  }
#undef OSH_SWQ
  Reals mq(c, ncands * kMaxMeshes);
  OSH_LAUNCH(swap3d_mesh_quals, ncands * kMaxMeshes, c, lsize.data(), ntris.data(),
      tq.data(), el.data(), o.max_length_allowed, tb, needp, mq.data());
  OSH_LAUNCH(swap3d_choose, ncands * choose_lanes, c, choose_lanes, cands.data(),
      lsize.data(), mq.data(),
      e2t.a2ab.data(),
      e2t.ab2b.data(), eq.data(), tb, needp, configs.data(), quals.data(), keep.data(),
      cache ? cache->valid.data() : (I8*)nullptr, cache ? cache->cfg.data() : (I8*)nullptr,
      cache ? cache->keep.data() : (I8*)nullptr, cache ? cache->q.data() : (Real*)nullptr);
}

bool swap_edges_3d(Ctx& c, DMesh& m, AdaptOptions const& o, SwapCache* cache) {
  ask_lengths(c, m);
  Bytes configs, keep;
  Reals quals;
  Bytes cand = swap_candidates(c, m, o, &configs, keep, quals);
  // every edge -> tet row this pass reads is a candidate's
  LOs cands, cand_of_edge;
  Adj e2t = marked_up(c, m, 1, 3, cand, &cands, true, &cand_of_edge);
  if (cands.size() == 0) return false;
  LO ne = m.nents[1];
  Tables const* tb = device_tables(c);
  SwapLoops loops;
  swap3d_evaluate(c, m, o, cands, e2t, cache, configs, quals, keep, loops);
  LOs const& lsize = loops.lsize;
  LOs const& leev = loops.leev;
  LOs const& lverts = loops.lverts;
  bool any;
  Bytes keys = find_indset(c, m, 1, quals, keep, &any, &e2t);
  if (!any) return false;
  Range modify_range(c, "modify_ents");
  // per-edge product counts and the rebuild plans from the key mask, then
  // one scan launch and one sync for the keys, their products and the new
  // numbering of edges, faces and tets
  LOs counts[4], offs[4], rep_of[4];
  ScanJob jobs[kMaxScanJobs];
  offs[0] = LOs(c, ne + 1);
  jobs[0] = {keys.data(), kScanNonzero, ne, offs[0].data(), slot_ptr<LO>(c, 0)};
  for (int d = 1; d <= 3; ++d) {
    counts[d] = LOs(c, ne);
    offs[d] = LOs(c, ne + 1);
    jobs[d] = {counts[d].data(), kScanLO, ne, offs[d].data(), slot_ptr<LO>(c, d)};
  }
  OSH_LAUNCH(swap3d_counts, ne, c, keys.data(), e2t.a2ab.data(), tb, counts[1].data(),
      counts[2].data(), counts[3].data());
  ModPlan plans[3];
  LO const* pc[4] = {nullptr, counts[1].data(), counts[2].data(), counts[3].data()};
  modify_plans_masked(c, m, 1, 3, keys, pc, plans, rep_of);
  int njobs = 4 + modify_scan_jobs(c, plans, 3, jobs + 4, 4);
  offset_scans(c, jobs, njobs);
  c.sync();
  modify_scan_totals(c, plans, 3, 4);
  LO nkeys = read_slot<LO>(c, 0);
  if (o.verbosity >= 2) std::printf("swapping %d 3D edges\n", int(nkeys));
  LOs keys2edges(c, nkeys);
  LOs k2p[4], pv[4];
  LO nprods[4] = {0, 0, 0, 0};
  for (int d = 1; d <= 3; ++d) {
    nprods[d] = read_slot<LO>(c, d);
    pv[d] = LOs(c, nprods[d] * (d + 1));
    k2p[d] = LOs(c, nkeys + 1);
    plans[d - 1].reps = LOs(c, nkeys);
  }
  OSH_LAUNCH(swap3d_key_offsets, ne, c, keys.data(), offs[0].data(), offs[1].data(),
      offs[2].data(), offs[3].data(), rep_of[2].data(), rep_of[3].data(), keys2edges.data(),
      k2p[1].data(), k2p[2].data(), k2p[3].data(), plans[0].reps.data(), plans[1].reps.data(),
      plans[2].reps.data());
  note_changed_parts(c, m, 1, keys2edges);
  LOs prod_keys[4];  // the key of each product face and tet
  prod_keys[2] = LOs(c, nprods[2]);
  prod_keys[3] = LOs(c, nprods[3]);
  OSH_LAUNCH(swap3d_products, nkeys * kLanes, c, keys2edges.data(), configs.data(),
      cand_of_edge.data(), lsize.data(), leev.data(), lverts.data(), tb, k2p[1].data(),
      k2p[2].data(), k2p[3].data(), pv[1].data(), pv[2].data(), pv[3].data(),
      prod_keys[2].data(), prod_keys[3].data());
  Bytes dirty;
  if (cache) {
    dirty = fill_i8(c, ne, 0);
    OSH_LAUNCH(swap3d_mark_dirty, m.nents[3], c, plans[2].same.data(),
        ask_down(c, m, 3, 1, nullptr).data(), dirty.data());
  }
  swap_rebuild(c, m, keys2edges, k2p, nprods, pv, plans, cache, cache ? &dirty : nullptr, &e2t,
      prod_keys);
  return true;
}

bool swap_edges_2d(Ctx& c, DMesh& m, AdaptOptions const& o) {
  ask_lengths(c, m);
  Bytes keep;
  Reals quals;
  Bytes cand = swap_candidates(c, m, o, nullptr, keep, quals);
  LOs cands = collect_marked(c, cand);
  if (cands.size() == 0) return false;
  int md = (m.metric_ncomps == 1) ? 1 : 2;
  Adj const& e2t = ask_up(c, m, 1, 2);
  LOs const& tv2v = verts_of(c, m, 2);
  Reals const& eq = ask_qualities(c, m);
  Range quality_range(c, "swap2d_qualities");
  if (md == 2)
    OSH_LAUNCH(swap2d_qualities_kernel<2>, cands.size(), c, cands.data(), e2t.a2ab.data(),
        e2t.ab2b.data(), e2t.codes.data(), m.down[1].data(), tv2v.data(), m.coords.data(),
        m.metric.data(), eq.data(), o.max_length_allowed, keep.data(), quals.data());
  else
    OSH_LAUNCH(swap2d_qualities_kernel<1>, cands.size(), c, cands.data(), e2t.a2ab.data(),
        e2t.ab2b.data(), e2t.codes.data(), m.down[1].data(), tv2v.data(), m.coords.data(),
        m.metric.data(), eq.data(), o.max_length_allowed, keep.data(), quals.data());
  quality_range.end();
  bool any;
  Bytes keys = find_indset(c, m, 1, quals, keep, &any);
  if (!any) return false;
  Range modify_range(c, "modify_ents");
  LOs keys2edges = collect_marked(c, keys);
  LO nkeys = keys2edges.size();
  note_changed_parts(c, m, 1, keys2edges);
  if (o.verbosity >= 2) std::printf("swapping %d 2D edges\n", int(nkeys));
  LOs k2p[4], pv[4];
  k2p[1] = iota(c, nkeys + 1);
  k2p[2] = iota(c, nkeys + 1, 0, 2);
  LO nprods[3] = {0, nkeys, 2 * nkeys};
  pv[1] = LOs(c, nkeys * 2);
  pv[2] = LOs(c, nkeys * 6);
  OSH_LAUNCH(swap2d_products, nkeys, c, keys2edges.data(), e2t.a2ab.data(), e2t.ab2b.data(),
      e2t.codes.data(), m.down[1].data(), tv2v.data(), pv[1].data(), pv[2].data());
  swap_rebuild(c, m, keys2edges, k2p, nprods, pv);
  return true;
}

}  // namespace

bool swap_edges(Ctx& c, DMesh& m, AdaptOptions const& o, SwapCache* cache) {
  if (m.dim == 3) return swap_edges_3d(c, m, o, cache);
  return swap_edges_2d(c, m, o);
}

namespace {

__global__ void mark_listed(LO n, LO const* list, I8* marks) {
  LO i = thread_index();
  if (i < n) marks[list[i]] = 1;
}

__global__ void gather_choices(LO n, LO const* edges, I8 const* configs, Real const* quals,
    I8 const* keep, LO* out_cfg, Real* out_q) {
  LO i = thread_index();
  if (i >= n) return;
  LO e = edges[i];
  out_cfg[i] = configs[e];
  out_q[i] = keep[e] ? quals[e] : -1.0;
}

}  // namespace

void swap3d_choices(Ctx& c, DMesh& m, AdaptOptions const& o, LOs const& edges,
    LOs& configs_out, Reals& quals_out) {
  OSH_CHECK(m.dim == 3, "swap3d choices need a 3D mesh");
  LO ne = m.nents[1];
  LO n = edges.size();
  Bytes marks = fill_i8(c, ne, 0);
  OSH_LAUNCH(mark_listed, n, c, edges.data(), marks.data());
  LOs cands, cand_of_edge;
  Adj e2t = marked_up(c, m, 1, 3, marks, &cands, true, &cand_of_edge);
  Bytes configs = fill_i8(c, ne, -1);
  Bytes keep = fill_i8(c, ne, 0);
  Reals quals = fill_real(c, ne, -1.0);
  if (cands.size() > 0) {
    SwapLoops loops;
    swap3d_evaluate(c, m, o, cands, e2t, nullptr, configs, quals, keep, loops);
  }
  configs_out = LOs(c, n);
  quals_out = Reals(c, n);
  OSH_LAUNCH(gather_choices, n, c, edges.data(), configs.data(), quals.data(), keep.data(),
      configs_out.data(), quals_out.data());
}

void find_loops(Ctx& c, DMesh& m, LOs const& edges, LOs& eev2v, LOs& loop_verts, LOs& sizes) {
  LO n = edges.size();
  eev2v = fill_lo(c, n * 2, -1);
  loop_verts = fill_lo(c, n * swap::kMaxLoop, -1);
  sizes = LOs(c, n);
  Adj const& e2t = ask_up(c, m, 1, 3);
  LOs const& tv2v = verts_of(c, m, 3);
  OSH_LAUNCH(loops_kernel, n, c, edges.data(), e2t.a2ab.data(), e2t.ab2b.data(),
      e2t.codes.data(), m.down[1].data(), tv2v.data(), eev2v.data(), loop_verts.data(),
      sizes.data(), c.error_word());
}

}  // namespace osh

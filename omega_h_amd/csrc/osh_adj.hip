// Adjacency derivation on the device.
//
//   transit                 src/Omega_h_adj.cpp:443-510
//   invert_adj              src/Omega_h_adj.cpp:178-263 (count, scan, fill,
//                           then each row ordered by high index so the result
//                           does not depend on atomic order)
//   reflect_down / find_matches src/Omega_h_adj.cpp:297-441
//   verts_across_edges, edges_across_tris/tets, add_edges
//                           src/Omega_h_adj.cpp:512-589, src/Omega_h_graph.cpp:14-40
//   side_elems (the two elements of each side, for elements_across_sides)
//                           src/Omega_h_adj.cpp:591-657
//   find_unique             src/Omega_h_adj.cpp:55-153
#include "osh_mesh.hpp"

namespace osh {

int simplex_degree(int from_dim, int to_dim) {
  static int const deg[4][4] = {{1, -1, -1, -1}, {2, 1, -1, -1}, {3, 3, 1, -1}, {4, 6, 4, 1}};
  return deg[from_dim][to_dim];
}

void DMesh::clear_caches() {
  for (int d = 0; d < 4; ++d) {
    verts[d] = LOs();
    for (int e = 0; e < 4; ++e) up[d][e] = Adj();
  }
  star[0] = star[1] = Adj();
  side_elems = DArray<unsigned long long>();
  dual = Adj();
  size = Reals();
  t2e = LOs();
  t2e_codes = Bytes();
}

namespace {

__device__ __forceinline__ int rotate_index(int n, int index, int rotation) {
  return (index + rotation) % n;
}
__device__ __forceinline__ int invert_rotation(int n, int rotation) {
  return (n - rotation) % n;
}
__device__ __forceinline__ I8 invert_alignment(int n, I8 code) {
  if (code_is_flipped(code)) return code;
  return make_code(0, invert_rotation(n, code_rotation(code)), 0);
}
__device__ __forceinline__ int align_vert_index(int n, int index, I8 code) {
  int r = rotate_index(n, index, code_rotation(code));
  return code_is_flipped(code) ? (r == 0 ? 0 : (1 + ((n - 1) - 1 - (r - 1)))) : r;
}
__device__ __forceinline__ int align_edge_index(int n, int index, I8 code) {
  int r = rotate_index(n, index, code_rotation(code));
  return code_is_flipped(code) ? (n - 1 - r) : r;
}

// first entry of simplex_up_template, src/Omega_h_simplex.hpp:91-229
__device__ __forceinline__ void up_template0(int elem_dim, int bdry_dim, int which,
    int& up, int& which_down, int& flipped) {
  if (elem_dim == 3 && bdry_dim == 0) {
    up = (which == 3) ? 5 : which;
    which_down = (which == 3) ? 1 : 0;
    flipped = 0;
  } else if (elem_dim == 3 && bdry_dim == 1) {
    int const t[6][2] = {{0, 2}, {0, 1}, {0, 0}, {1, 2}, {2, 2}, {3, 2}};
    up = t[which][0];
    which_down = t[which][1];
    flipped = 1;
  } else {
    up = which;
    which_down = 0;
    flipped = 0;
  }
}

// one thread per (high entity, local low entity), so that the gathers of a
// high entity's lows run side by side
template <int HIGH, int LOW>
__global__ void transit_kernel(LO nuses, LO const* hm2m, I8 const* m2hm_codes,
    LO const* ml2l, I8 const* ml_codes, LO* hl2l, I8* codes) {
  LO t = thread_index();
  if (t >= nuses) return;
  constexpr int MID = LOW + 1;
  constexpr int nmph = (HIGH == 3) ? (MID == 2 ? 4 : 6) : 3;
  constexpr int nlpm = MID + 1;
  constexpr int nlph = (HIGH == 3) ? (LOW == 0 ? 4 : 6) : 3;
  LO h = t / nlph;
  int hl = t % nlph;
  {
    int hm, hml, canon_flipped;
    up_template0(HIGH, LOW, hl, hm, hml, canon_flipped);
    LO m = hm2m[h * nmph + hm];
    I8 hm2m_code = invert_alignment(nlpm, m2hm_codes[h * nmph + hm]);
    int ml = (LOW == 0) ? align_vert_index(nlpm, hml, hm2m_code)
                        : align_edge_index(nlpm, hml, hm2m_code);
    hl2l[h * nlph + hl] = ml2l[m * nlpm + ml];
    if (LOW == 1) {
      int region_face_flipped = code_is_flipped(hm2m_code);
      int face_edge_flipped = (code_rotation(ml_codes[m * nlpm + ml]) == 1);
      codes[h * nlph + hl] =
          make_code(0, region_face_flipped ^ face_edge_flipped ^ canon_flipped, 0);
    }
  }
}

// invert_adj in three sweeps over the uses hl = high * deg + which:
//   count_lows: per-low use counts by atomics; each use keeps its slot
//   fill_uses:  ab2hl[a2ab[l] + slot] = hl (slots are in atomic order)
//   rank_uses:  each use counts the uses of its row with a smaller hl and
//               writes itself at that rank, which sorts every row by high
//               index (sort_by_high_index, src/Omega_h_adj.cpp:178-200)
//               and forms the codes of separate_upward_with_codes (:202-217)
// With `marks`, only the rows of marked lows are built (the rest are empty).
__global__ void count_lows(LO nhl, LO const* hl2l, I8 const* marks, LO* counts, LO* slot) {
  LO i = thread_index();
  if (i >= nhl) return;
  LO l = hl2l[i];
  if (!marks || marks[l]) slot[i] = atomicAdd(&counts[l], 1);
}

__global__ void fill_uses(LO nhl, LO const* hl2l, I8 const* marks, LO const* a2ab,
    LO const* slot, LO* ab2hl) {
  LO hl = thread_index();
  if (hl >= nhl) return;
  LO l = hl2l[hl];
  if (!marks || marks[l]) ab2hl[a2ab[l] + slot[hl]] = hl;
}

__global__ void rank_uses(LO nhl, LO const* hl2l, I8 const* marks, LO const* a2ab,
    LO const* ab2hl, int deg, I8 const* down_codes, LO* ab2b, I8* codes) {
  LO hl = thread_index();
  if (hl >= nhl) return;
  LO l = hl2l[hl];
  if (marks && !marks[l]) return;
  LO b = a2ab[l], e = a2ab[l + 1];
  LO rank = 0;
  for (LO j = b; j < e; ++j) rank += (ab2hl[j] < hl);
  int which_down = hl % deg;
  I8 c = down_codes ? make_code(code_is_flipped(down_codes[hl]), code_rotation(down_codes[hl]),
                          which_down)
                    : make_code(0, 0, which_down);
  ab2b[b + rank] = hl / deg;
  codes[b + rank] = c;
}

// rows in atomic order, for uses that do not depend on the order of a row
__global__ void fill_uses_unsorted(LO nhl, LO const* hl2l, I8 const* marks, LO const* a2ab,
    LO const* slot, int deg, I8 const* down_codes, LO* ab2b, I8* codes) {
  LO hl = thread_index();
  if (hl >= nhl) return;
  LO l = hl2l[hl];
  if (marks && !marks[l]) return;
  LO pos = a2ab[l] + slot[hl];
  int which_down = hl % deg;
  ab2b[pos] = hl / deg;
  codes[pos] = down_codes ? make_code(code_is_flipped(down_codes[hl]),
                                code_rotation(down_codes[hl]), which_down)
                          : make_code(0, 0, which_down);
}

// With `marks`, only marked lows get rows; with `marked`, the list of marked
// lows comes from the same scan launch; without `sorted`, rows keep the
// order of the atomics (one sweep less)
Adj invert_adj(Ctx& c, LO nhigh, int deg, LO nlow, LOs const& hl2l, Bytes const* down_codes,
    Bytes const* marks = nullptr, LOs* marked = nullptr, bool sorted = true,
    LOs* marked_index = nullptr) {
  LO nhl = nhigh * deg;
  I8 const* mk = marks ? marks->data() : nullptr;
  LOs counts = fill_lo(c, nlow, 0);
  LOs slot(c, nhl);
  OSH_LAUNCH(count_lows, nhl, c, hl2l.data(), mk, counts.data(), slot.data());
  Adj a;
  a.a2ab = LOs(c, nlow + 1);
  if (marked) {
    LOs marked_offsets(c, nlow + 1);
    ScanJob jobs[2] = {{counts.data(), kScanLO, nlow, a.a2ab.data(), nullptr},
        {mk, kScanNonzero, nlow, marked_offsets.data(), slot_ptr<LO>(c, 1)}};
    offset_scans(c, jobs, 2);
    c.sync();
    *marked = LOs(c, read_slot<LO>(c, 1));
    scatter_marked_list(c, *marks, marked_offsets, *marked);
    if (marked_index) *marked_index = marked_offsets;
  } else {
    ScanJob job{counts.data(), kScanLO, nlow, a.a2ab.data(), nullptr};
    offset_scans(c, &job, 1);
  }
  // with marks the rows hold at most nhl uses; the size stays an upper bound
  a.ab2b = LOs(c, nhl);
  a.codes = Bytes(c, nhl);
  I8 const* dc = down_codes ? down_codes->data() : nullptr;
  if (sorted) {
    LOs ab2hl(c, nhl);
    OSH_LAUNCH(fill_uses, nhl, c, hl2l.data(), mk, a.a2ab.data(), slot.data(), ab2hl.data());
    OSH_LAUNCH(rank_uses, nhl, c, hl2l.data(), mk, a.a2ab.data(), ab2hl.data(), deg, dc,
        a.ab2b.data(), a.codes.data());
  } else {
    OSH_LAUNCH(fill_uses_unsorted, nhl, c, hl2l.data(), mk, a.a2ab.data(), slot.data(), deg, dc,
        a.ab2b.data(), a.codes.data());
  }
  return a;
}

// reflect_down: each existing lower entity is listed once, in the row of its
// smallest vertex, and rows are built only for vertices that are the
// smallest vertex of some use.  A use's match has the same vertex set, so it
// sits in the row of the use's smallest vertex; the match is unique, hence it
// is the entity the walk over the full vertex -> lower-entity adjacency finds
// (find_matches, src/Omega_h_adj.cpp:297-422).
template <int NV>
__device__ __forceinline__ LO min_vertex(LO const* v) {
  LO m = v[0];
#pragma unroll
  for (int k = 1; k < NV; ++k) m = (v[k] < m) ? v[k] : m;
  return m;
}

// counts[v] = 1 marks v; a marked row's final slot stays unused; each use's
// match count starts at zero
template <int HIGH>
__global__ void mark_use_min_verts(LO nuses, LO const* prod_verts, LO* counts, LO* found) {
  LO u = thread_index();
  if (u >= nuses) return;
  found[u] = 0;
  constexpr int LOW = HIGH - 1;
  LO h = u / (HIGH + 1);
  int which = u % (HIGH + 1);
  LO av[3];
#pragma unroll
  for (int k = 0; k < LOW + 1; ++k)
    av[k] = prod_verts[h * (HIGH + 1) + down_template(HIGH, LOW, which, k)];
  counts[min_vertex<LOW + 1>(av)] = 1;
}

template <int NVL>
__global__ void count_min_rows(LO nl, LO const* lv2v, LO* counts, LO* slot) {
  LO l = thread_index();
  if (l >= nl) return;
  LO vm = min_vertex<NVL>(lv2v + l * NVL);
  // unmarked counts stay 0 throughout this kernel, marked ones are >= 1
  slot[l] = counts[vm] ? atomicAdd(&counts[vm], 1) - 1 : -1;
}

// a row entry is {entity, its vertices}, so the match reads no other array
template <int NVL>
__global__ void fill_min_rows(LO nl, LO const* lv2v, LO const* a2ab, LO const* slot, int4* rows) {
  LO l = thread_index();
  if (l >= nl) return;
  LO s = slot[l];
  if (s < 0) return;
  LO v[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < NVL; ++k) v[k] = lv2v[l * NVL + k];
  rows[a2ab[min_vertex<NVL>(v)] + s] = make_int4(l, v[0], v[1], v[2]);
}

// kMatchLanes threads per use share its row; a lane that finds a match
// writes it and counts it in found[u], which the consumer checks (a use needs
// exactly one match)
constexpr int kMatchLanes = 4;

template <int HIGH>
__global__ void match_uses(LO nt, LO const* __restrict__ prod_verts,
    LO const* __restrict__ a2ab, int4 const* __restrict__ rows, LO* __restrict__ out,
    I8* __restrict__ out_codes, LO* __restrict__ found) {
  LO t = thread_index();
  if (t >= nt) return;
  LO u = t / kMatchLanes;
  int lane = t % kMatchLanes;
  constexpr int LOW = HIGH - 1;
  constexpr int nvh = HIGH + 1;
  constexpr int nvl = LOW + 1;
  LO h = u / (HIGH + 1);
  int which = u % (HIGH + 1);
  LO av[3];
#pragma unroll
  for (int k = 0; k < nvl; ++k) av[k] = prod_verts[h * nvh + down_template(HIGH, LOW, which, k)];
  LO vm = min_vertex<nvl>(av);
  LO k0 = a2ab[vm], k1 = a2ab[vm + 1] - 1;
  for (LO k = k0 + lane; k < k1; k += kMatchLanes) {
    int4 r = rows[k];
    LO bv[3] = {r.y, r.z, r.w};
    // wd: the position of the use's first vertex in the candidate
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
    out[u] = r.x;
    out_codes[u] = code;
    atomicAdd(&found[u], 1);
  }
}

// for callers that use the matches directly: a use without exactly one match
// sets the error word and gets entity 0, which keeps every index in range
__global__ void check_matches(LO nuses, LO const* found, LO* out, I8* out_codes, int* error) {
  LO u = thread_index();
  if (u >= nuses || found[u] == 1) return;
  out[u] = 0;
  out_codes[u] = 0;
  *error = kReflectDownNoMatch;
}

// per side, the number of elements on it (high 32 bits) and the sum of their
// indices (low 32 bits) by one 64-bit atomic add per element side: with two
// elements, either one finds the other as the sum minus itself
__global__ void side_elems_init(LO ns, unsigned long long* se) {
  LO s = thread_index();
  if (s < ns) se[s] = 0ull;
}

__global__ void side_elems_kernel(LO nuses, int nspe, LO const* es2s, unsigned long long* se) {
  LO i = thread_index();
  if (i >= nuses) return;
  LO e = i / nspe;
  atomicAdd(&se[es2s[i]], (1ull << 32) + static_cast<unsigned long long>(e));
}

// elements_across_sides (src/Omega_h_adj.cpp:591-657) from the per-side
// words: each element lists the other element across each of its sides that
// has two, in side order
__global__ void dual_degrees(LO ne, int nspe, LO const* es2s, unsigned long long const* se,
    LO* deg, int* error) {
  LO e = thread_index();
  if (e >= ne) return;
  LO d = 0;
  for (int j = 0; j < nspe; ++j) {
    unsigned count = unsigned(se[es2s[e * nspe + j]] >> 32);
    if (count == 2) ++d;
    if (count > 2) *error = kNonManifoldSide;
  }
  deg[e] = d;
}

__global__ void dual_fill(LO ne, int nspe, LO const* es2s, unsigned long long const* se,
    LO const* a2ab, LO* ab2b) {
  LO e = thread_index();
  if (e >= ne) return;
  LO k = a2ab[e];
  for (int j = 0; j < nspe; ++j) {
    unsigned long long w = se[es2s[e * nspe + j]];
    if (unsigned(w >> 32) == 2) ab2b[k++] = LO(unsigned(w)) - e;
  }
}

__global__ void verts_across_edges_kernel(LO nve, LO const* ve2e, I8 const* codes,
    LO const* ev2v, LO* vv2v) {
  LO ve = thread_index();
  if (ve >= nve) return;
  LO e = ve2e[ve];
  int eev = code_which_down(codes[ve]);
  vv2v[ve] = ev2v[e * 2 + (1 - eev)];
}

__global__ void edge_star_degrees(LO ne, LO const* e2f_a2ab, LO const* e2t_a2ab, LO* deg) {
  LO e = thread_index();
  if (e >= ne) return;
  LO d = 2 * (e2f_a2ab[e + 1] - e2f_a2ab[e]);
  if (e2t_a2ab) d += e2t_a2ab[e + 1] - e2t_a2ab[e];
  deg[e] = d;
}

__global__ void edge_star_fill(LO ne, LO const* e2f_a2ab, LO const* e2f_ab2b,
    I8 const* e2f_codes, LO const* fe2e, LO const* e2t_a2ab, LO const* e2t_ab2b,
    I8 const* e2t_codes, LO const* te2e, LO const* a2ab, LO* ab2b) {
  LO e = thread_index();
  if (e >= ne) return;
  LO k = a2ab[e];
  for (LO ef = e2f_a2ab[e]; ef < e2f_a2ab[e + 1]; ++ef) {
    LO f = e2f_ab2b[ef];
    int ffe = code_which_down(e2f_codes[ef]);
    ab2b[k++] = fe2e[f * 3 + ((ffe + 1) % 3)];
    ab2b[k++] = fe2e[f * 3 + ((ffe + 2) % 3)];
  }
  if (e2t_a2ab) {
    for (LO er = e2t_a2ab[e]; er < e2t_a2ab[e + 1]; ++er) {
      LO r = e2t_ab2b[er];
      int rre = code_which_down(e2t_codes[er]);
      ab2b[k++] = te2e[r * 6 + opposite_template(3, 1, rre)];
    }
  }
}

}  // namespace

Adj invert_adj_rows(Ctx& c, LO nhigh, int deg, LO nlow, LOs const& hl2l,
    Bytes const* down_codes) {
  return invert_adj(c, nhigh, deg, nlow, hl2l, down_codes);
}

LOs const& verts_of(Ctx& c, DMesh& m, int d) {
  if (d == 1) return m.down[1];
  if (!m.verts[d].exists()) {
    LO n = m.nents[d];
    LOs out(c, n * (d + 1));
    if (d == 2) {
      OSH_LAUNCH((transit_kernel<2, 0>), n * 3, c, m.down[2].data(), m.down_codes[2].data(),
          m.down[1].data(), (I8 const*)nullptr, out.data(), (I8*)nullptr);
    } else {
      Bytes const* codes;
      LOs const& t2e = ask_down(c, m, 3, 1, &codes);
      OSH_LAUNCH((transit_kernel<3, 0>), n * 4, c, t2e.data(), codes->data(), m.down[1].data(),
          (I8 const*)nullptr, out.data(), (I8*)nullptr);
    }
    m.verts[d] = out;
  }
  return m.verts[d];
}

LOs const& ask_down(Ctx& c, DMesh& m, int high, int low, Bytes const** codes) {
  if (low == 0) {
    if (codes) *codes = nullptr;
    return verts_of(c, m, high);
  }
  if (low == high - 1) {
    if (codes) *codes = &m.down_codes[high];
    return m.down[high];
  }
  if (!m.t2e.exists()) {
    LO n = m.nents[3];
    m.t2e = LOs(c, n * 6);
    m.t2e_codes = Bytes(c, n * 6);
    OSH_LAUNCH((transit_kernel<3, 1>), n * 6, c, m.down[3].data(), m.down_codes[3].data(),
        m.down[2].data(), m.down_codes[2].data(), m.t2e.data(), m.t2e_codes.data());
  }
  if (codes) *codes = &m.t2e_codes;
  return m.t2e;
}

Adj marked_up(Ctx& c, DMesh& m, int low, int high, Bytes const& marks, LOs* marked,
    bool sorted, LOs* marked_index) {
  Bytes const* codes = nullptr;
  LOs const& down = ask_down(c, m, high, low, &codes);
  return invert_adj(c, m.nents[high], simplex_degree(high, low), m.nents[low], down, codes,
      &marks, marked, sorted, marked_index);
}

Adj const& ask_up(Ctx& c, DMesh& m, int low, int high) {
  if (!m.up[low][high].exists()) {
    Bytes const* codes = nullptr;
    LOs const& down = ask_down(c, m, high, low, &codes);
    m.up[low][high] = invert_adj(c, m.nents[high], simplex_degree(high, low), m.nents[low],
        down, codes);
  }
  return m.up[low][high];
}

void reflect_down(Ctx& c, DMesh& m, int high_dim, LO nprods, LOs const& prod_verts,
    LOs& down, Bytes& codes, LOs* found_out) {
  int low_dim = high_dim - 1;
  LO nuses = nprods * (high_dim + 1);
  down = LOs(c, nuses);
  codes = Bytes(c, nuses);
  LOs found(c, nuses);
  if (found_out) *found_out = found;
  if (nuses == 0) return;
  LOs const& lv2v = verts_of(c, m, low_dim);
  LO nv = m.nents[0];
  LO nl = m.nents[low_dim];
  LOs counts = fill_lo(c, nv, 0);
  LOs slot(c, nl);
  if (high_dim == 2) {
    OSH_LAUNCH(mark_use_min_verts<2>, nuses, c, prod_verts.data(), counts.data(), found.data());
    OSH_LAUNCH(count_min_rows<2>, nl, c, lv2v.data(), counts.data(), slot.data());
  } else {
    OSH_LAUNCH(mark_use_min_verts<3>, nuses, c, prod_verts.data(), counts.data(), found.data());
    OSH_LAUNCH(count_min_rows<3>, nl, c, lv2v.data(), counts.data(), slot.data());
  }
  LOs a2ab = offset_scan(c, counts);
  // every lower entity in at most one row plus one spare slot per marked row
  DArray<int4> rows(c, nl + nv);
  if (high_dim == 2) {
    OSH_LAUNCH(fill_min_rows<2>, nl, c, lv2v.data(), a2ab.data(), slot.data(), rows.data());
    OSH_LAUNCH(match_uses<2>, nuses * kMatchLanes, c, prod_verts.data(), a2ab.data(),
        rows.data(), down.data(), codes.data(), found.data());
  } else {
    OSH_LAUNCH(fill_min_rows<3>, nl, c, lv2v.data(), a2ab.data(), slot.data(), rows.data());
    OSH_LAUNCH(match_uses<3>, nuses * kMatchLanes, c, prod_verts.data(), a2ab.data(),
        rows.data(), down.data(), codes.data(), found.data());
  }
  // a use without a unique match sets the context's error word, which the
  // next stream sync raises; a caller that takes the match counts checks
  // them where it reads the matches
  if (!found_out)
    OSH_LAUNCH(check_matches, nuses, c, found.data(), down.data(), codes.data(),
        c.error_word());
}

Adj const& ask_star(Ctx& c, DMesh& m, int d) {
  if (!m.star[d].exists()) {
    Adj g;
    if (d == 0) {
      Adj const& v2e = ask_up(c, m, 0, 1);
      g.a2ab = v2e.a2ab;
      g.ab2b = LOs(c, v2e.ab2b.size());
      OSH_LAUNCH(verts_across_edges_kernel, v2e.ab2b.size(), c, v2e.ab2b.data(),
          v2e.codes.data(), m.down[1].data(), g.ab2b.data());
    } else {
      LO ne = m.nents[1];
      Adj const& e2f = ask_up(c, m, 1, 2);
      Adj const* e2t = nullptr;
      LO const* te2e = nullptr;
      if (m.dim == 3) {
        e2t = &ask_up(c, m, 1, 3);
        te2e = ask_down(c, m, 3, 1, nullptr).data();
      }
      LOs deg(c, ne);
      OSH_LAUNCH(edge_star_degrees, ne, c, e2f.a2ab.data(),
          e2t ? e2t->a2ab.data() : (LO const*)nullptr, deg.data());
      LO total;
      g.a2ab = offset_scan(c, deg, &total);
      g.ab2b = LOs(c, total);
      OSH_LAUNCH(edge_star_fill, ne, c, e2f.a2ab.data(), e2f.ab2b.data(), e2f.codes.data(),
          m.down[2].data(), e2t ? e2t->a2ab.data() : (LO const*)nullptr,
          e2t ? e2t->ab2b.data() : (LO const*)nullptr,
          e2t ? e2t->codes.data() : (I8 const*)nullptr, te2e, g.a2ab.data(), g.ab2b.data());
    }
    m.star[d] = g;
  }
  return m.star[d];
}

Adj const& ask_dual(Ctx& c, DMesh& m) {
  if (!m.dual.exists()) {
    int dim = m.dim;
    LO ne = m.nents[dim];
    auto const& se = ask_side_elems(c, m);
    LOs deg(c, ne);
    OSH_LAUNCH(dual_degrees, ne, c, dim + 1, m.down[dim].data(), se.data(), deg.data(),
        c.error_word());
    Adj g;
    LO total;
    g.a2ab = offset_scan(c, deg, &total);
    g.ab2b = LOs(c, total);
    OSH_LAUNCH(dual_fill, ne, c, dim + 1, m.down[dim].data(), se.data(), g.a2ab.data(),
        g.ab2b.data());
    m.dual = g;
  }
  return m.dual;
}

DArray<unsigned long long> const& ask_side_elems(Ctx& c, DMesh& m) {
  if (!m.side_elems.exists()) {
    int dim = m.dim;
    LO ns = m.nents[dim - 1];
    DArray<unsigned long long> se(c, ns);
    OSH_LAUNCH(side_elems_init, ns, c, se.data());
    OSH_LAUNCH(side_elems_kernel, m.nents[dim] * (dim + 1), c, dim + 1, m.down[dim].data(),
        se.data());
    m.side_elems = se;
  }
  return m.side_elems;
}

// ---------------------------------------------------------------- find_unique
namespace {

template <int HIGH, int LOW>
__global__ void form_canonical_uses(LO nuses, LO const* hv2v, LO* uv2v, I64* canon) {
  LO u = thread_index();
  if (u >= nuses) return;
  constexpr int nvh = HIGH + 1;
  constexpr int nlh = (HIGH == 3) ? (LOW == 1 ? 6 : 4) : 3;
  constexpr int deg = LOW + 1;
  LO h = u / nlh;
  int which = u % nlh;
  LO ev[3];
  for (int k = 0; k < deg; ++k) {
    ev[k] = hv2v[h * nvh + down_template(HIGH, LOW, which, k)];
    uv2v[u * deg + k] = ev[k];
  }
  int min_j = 0;
  LO min_v = ev[0];
  for (int j = 1; j < deg; ++j)
    if (ev[j] < min_v) {
      min_j = j;
      min_v = ev[j];
    }
  int rotation = (deg - min_j) % deg;
  LO tmp[3];
  for (int j = 0; j < deg; ++j) tmp[(j + rotation) % deg] = ev[j];
  if (deg == 3 && tmp[2] < tmp[1]) {
    LO t = tmp[1];
    tmp[1] = tmp[2];
    tmp[2] = t;
  }
  for (int j = 0; j < deg; ++j) canon[I64(u) * deg + j] = tmp[j];
}

__global__ void mark_group_ends(LO n, int deg, I64 const* canon, LO const* sorted2u, I8* last) {
  LO s = thread_index();
  if (s >= n) return;
  I8 r = 1;
  if (s < n - 1) {
    LO a = sorted2u[s], b = sorted2u[s + 1];
    r = 0;
    for (int j = 0; j < deg; ++j)
      if (canon[I64(a) * deg + j] != canon[I64(b) * deg + j]) r = 1;
  }
  last[s] = r;
}

__global__ void gather_unique(LO nlow, int deg, LO const* ends, LO const* sorted2u,
    LO const* uv2v, LO* out) {
  LO i = thread_index();
  if (i >= nlow) return;
  LO u = sorted2u[ends[i]];
  for (int j = 0; j < deg; ++j) out[i * deg + j] = uv2v[u * deg + j];
}

}  // namespace

LOs find_unique(Ctx& c, LO nhigh, int high_dim, int low_dim, LOs const& hv2v) {
  int nlh = simplex_degree(high_dim, low_dim);
  int deg = low_dim + 1;
  LO nuses = nhigh * nlh;
  LOs uv2v(c, nuses * deg);
  DArray<I64> canon(c, nuses * deg);
  if (high_dim == 3 && low_dim == 1)
    OSH_LAUNCH((form_canonical_uses<3, 1>), nuses, c, hv2v.data(), uv2v.data(), canon.data());
  else if (high_dim == 3 && low_dim == 2)
    OSH_LAUNCH((form_canonical_uses<3, 2>), nuses, c, hv2v.data(), uv2v.data(), canon.data());
  else
    OSH_LAUNCH((form_canonical_uses<2, 1>), nuses, c, hv2v.data(), uv2v.data(), canon.data());
  LOs sorted2u = sort_by_keys(c, canon, deg);
  Bytes ends(c, nuses);
  OSH_LAUNCH(mark_group_ends, nuses, c, deg, canon.data(), sorted2u.data(), ends.data());
  LOs e2sorted = collect_marked(c, ends);
  LO nlow = e2sorted.size();
  LOs out(c, nlow * deg);
  OSH_LAUNCH(gather_unique, nlow, c, deg, e2sorted.data(), sorted2u.data(), uv2v.data(),
      out.data());
  return out;
}

}  // namespace osh

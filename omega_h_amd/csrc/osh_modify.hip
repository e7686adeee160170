// Topology rebuild, field transfer, independent sets and marking on the
// device.
//
//   modify_ents_adapt / get_mods2reps / get_rep_counts / get_rep2md_order /
//   assign_new_numbering / modify_conn   src/Omega_h_modify.cpp:20-70,141-517
//   (on one rank the global numbers equal the new local numbers, so the
//   modify_globals exchange reduces to the same scan)
//   transfer_common / transfer_inherit_* src/Omega_h_transfer.cpp:149-180
//   indset::find / QualityCompare         src/Omega_h_indset_inline.hpp:12-72,
//                                         src/Omega_h_indset.cpp:5-35
//   mark_down / mark_dual_layers / mark_sliver_layers src/Omega_h_mark.cpp:12-47,190-212
#include "osh_math.hpp"
#include "osh_passes.hpp"

namespace osh {

namespace {

// the same entities of dimension > key dimension: those with no key among
// their lower entities, each counting one new entity; each key also receives
// the lowest index among the entities above it, which is the first entry of
// its row of the upward adjacency (rows are sorted by index), without
// building that adjacency
// the dimensions above the keys in one launch: job w owns the threads
// [start[w], start[w + 1])
struct SameUpJobs {
  LO start[4];
  int deg[3];
  LO const* hl2l[3];
  I8* same[3];
  LO* counts[3];
  LO* first_up[3];
  int n;
};

__global__ void mark_same_up_kernel(LO nt, I8 const* low_marks, SameUpJobs j) {
  LO t = thread_index();
  if (t >= nt) return;
  int w = 0;
  while (w + 1 < j.n && t >= j.start[w + 1]) ++w;
  LO h = t - j.start[w];
  int deg = j.deg[w];
  LO const* hl2l = j.hl2l[w];
  I8 r = 1;
  for (int k = 0; k < deg; ++k) {
    LO l = hl2l[h * deg + k];
    if (low_marks[l]) {
      r = 0;
      atomicMin(&j.first_up[w][l], h);
    }
  }
  j.same[w][h] = r;
  j.counts[w][h] = r;
}

// new numbers of the old entities in one sweep: same entities take their
// offset, the rest -1 (src/Omega_h_modify.cpp:141-161); the same entities
// are also listed in order (collect_same) and, when `down` is given, write
// their connectivity renumbered (modify_conn, :446-466; a null
// old_lows2new_lows keeps the lower numbers) and, when `verts` is given,
// their vertex list renumbered
struct SameNumbering {
  I8 const* same;
  LO const* same_offsets;
  LO const* offsets;
  LO* same2old;
  LO* same2new;
  LO* old2new;
  int deg;
  LO const* old_down;
  I8 const* old_codes;
  LO const* old_lows2new_lows;
  LO* down;
  I8* codes;
  int nvpe;
  LO const* old_verts;
  LO const* verts_old2new;
  LO* verts;
};

__device__ __forceinline__ void number_same(LO o, SameNumbering const& a) {
  if (a.same && !a.same[o]) {
    a.old2new[o] = -1;
    return;
  }
  LO s = a.same ? a.same_offsets[o] : o;
  LO w = a.offsets[o];
  a.same2old[s] = o;
  a.same2new[s] = w;
  a.old2new[o] = w;
  int deg = a.deg;
  if (a.down)
    for (int j = 0; j < deg; ++j) {
      LO l = a.old_down[o * deg + j];
      a.down[w * deg + j] = a.old_lows2new_lows ? a.old_lows2new_lows[l] : l;
      if (a.codes) a.codes[w * deg + j] = a.old_codes[o * deg + j];
    }
  if (a.verts)
    for (int k = 0; k < a.nvpe; ++k) {
      LO v = a.old_verts[o * a.nvpe + k];
      a.verts[w * a.nvpe + k] = a.verts_old2new ? a.verts_old2new[v] : v;
    }
}

// One sweep over the key-dimension entities starts the plans of every
// dimension of a pass: no key marked, no entity above found yet, and (when
// the key dimension itself is rebuilt) every entity the same, counting one.
struct KeyDimInit {
  I8* key_marks;
  LO* first_up[4];  // per dimension above the keys (null elsewhere)
  I8* same;         // when the key dimension is rebuilt
  LO* counts;
  // with a key mask, the keys are marked in the same sweep: no longer the
  // same, each counting its products
  I8 const* key_mask;
  LO const* key_prod_counts;
};

__global__ void init_key_dim(LO n, KeyDimInit a) {
  LO i = thread_index();
  if (i >= n) return;
  if (a.key_marks) a.key_marks[i] = 0;
  for (int d = 0; d < 4; ++d)
    if (a.first_up[d]) a.first_up[d][i] = 0x7fffffff;
  if (a.same) {
    bool key = a.key_mask && a.key_mask[i];
    a.same[i] = key ? 0 : 1;
    a.counts[i] = key ? a.key_prod_counts[i] : 1;
  }
}

// the keys: marked; when the key dimension is rebuilt, a key is its own
// representative (get_mods2reps, src/Omega_h_modify.cpp:20-45), is no longer
// the same and counts its products (get_rep_counts, :47-70)
__global__ void mark_keys(LO nkeys, LO const* keys2kds, LO const* keys2prods, I8* key_marks,
    I8* same, LO* counts, LO* reps) {
  LO k = thread_index();
  if (k >= nkeys) return;
  LO kd = keys2kds[k];
  if (key_marks) key_marks[kd] = 1;
  if (same) {
    same[kd] = 0;
    counts[kd] = keys2prods[k + 1] - keys2prods[k];
    reps[k] = kd;
  }
}

// representatives of the keys for up to three dimensions in one launch: the
// lowest entity above the key (first_up) or the key's first lower entity
// (down); the representative counts the key's products
struct RepsJob {
  LO const* first_up[3];
  LO const* down[3];
  int deg[3];
  LO const* keys2prods[3];
  LO* reps[3];
  LO* counts[3];
};

__global__ void reps_kernel(LO nt, LO nkeys, LO const* keys2kds, RepsJob j, int* error) {
  LO t = thread_index();
  if (t >= nt) return;
  int w = t / nkeys;
  LO k = t % nkeys;
  LO kd = keys2kds[k];
  LO r;
  if (j.first_up[w]) {
    r = j.first_up[w][kd];
    if (r == 0x7fffffff) {  // keeps every index in range until the host sees it
      r = 0;
      *error = kKeyWithoutUp;
    }
  } else {
    r = j.down[w][kd * j.deg[w]];
  }
  j.reps[w][k] = r;
  atomicAdd(&j.counts[w][r], j.keys2prods[w][k + 1] - j.keys2prods[w][k]);
}

// the same starts with the keys given as a mask over the key dimension and
// their product counts per key-dimension entity: a key of the rebuilt key
// dimension is no longer the same and counts its products; for the
// dimensions above, each key's representative (kept per key-dimension
// entity) counts the key's products
struct MaskedReps {
  LO const* first_up[3];
  LO const* prod_counts[3];
  LO* rep_of[3];
  LO* counts[3];
};

__global__ void reps_masked(LO nt, LO nkd, I8 const* key_mask, MaskedReps j, int* error) {
  LO t = thread_index();
  if (t >= nt) return;
  int w = t / nkd;
  LO kd = t % nkd;
  if (!key_mask[kd]) return;
  LO r = j.first_up[w][kd];
  if (r == 0x7fffffff) {  // keeps every index in range until the host sees it
    r = 0;
    *error = kKeyWithoutUp;
  }
  j.rep_of[w][kd] = r;
  atomicAdd(&j.counts[w][r], j.prod_counts[w][kd]);
}

__global__ void scatter_index(LO n, LO const* a2b, LO* b2a) {
  LO i = thread_index();
  if (i < n) b2a[a2b[i]] = i;
}

__global__ void rep2md_order_kernel(LO nreps, LO const* a2ab, LO const* ab2b, I8 const* codes,
    LO const* md2mod, LO const* keys2prods, LO* order) {
  LO rep = thread_index();
  if (rep >= nreps) return;
  LO offset = 0;
  for (LO j = a2ab[rep]; j < a2ab[rep + 1]; ++j) {
    LO md = ab2b[j];
    LO mod = md2mod[md];
    if (mod >= 0 && code_which_down(codes[j]) == 0) {
      order[md] = offset;
      offset += keys2prods[mod + 1] - keys2prods[mod];
    }
  }
}

// the products' new numbers, following their representative
// (src/Omega_h_modify.cpp:141-161), with their connectivity (modify_conn),
// their vertex lists and the entity each came from when those are asked for.
// A key's products take consecutive numbers, so kProdLanes threads per key
// share its products without talking to each other.
constexpr int kProdLanes = 8;

struct ProdNumbering {
  LO const* reps;
  LO const* keys2kds;
  LO const* offsets;
  LO const* order;
  LO const* keys2prods;
  LO* prods2new;
  int deg;
  LO const* pdown;
  I8 const* pcodes;
  LO const* found;
  LO* down;
  I8* codes;
  int nvpe;
  LO const* pverts;
  LO* verts;
  LO const* key_src;
  LO* src;
  int* error;
};

__device__ __forceinline__ void number_prods(LO t, ProdNumbering const& a) {
  LO k = t / kProdLanes;
  int lane = t % kProdLanes;
  LO first = a.keys2prods[k];
  LO base = a.offsets[a.reps[k]];
  if (a.order) base += a.order[a.keys2kds[k]] + 1;
  int deg = a.deg;
  for (LO p = first + lane; p < a.keys2prods[k + 1]; p += kProdLanes) {
    LO w = base + (p - first);
    a.prods2new[p] = w;
    if (a.down)
      for (int j = 0; j < deg; ++j) {
        // a side matched by reflect_down must have matched exactly once; an
        // error keeps the index in range until the host sees it
        bool ok = !a.found || a.found[p * deg + j] == 1;
        if (!ok) *a.error = kReflectDownNoMatch;
        a.down[w * deg + j] = ok ? a.pdown[p * deg + j] : 0;
        if (a.codes) a.codes[w * deg + j] = ok ? a.pcodes[p * deg + j] : I8(0);
      }
    if (a.verts)
      for (int j = 0; j < a.nvpe; ++j) a.verts[w * a.nvpe + j] = a.pverts[p * a.nvpe + j];
    if (a.src) a.src[p] = a.key_src[k];
  }
}

// the new numbers of one dimension in one launch: threads [0, nold) number
// the old entities (same ones keep their order), the rest (kProdLanes per
// key) the products; the two halves write disjoint entries
__global__ void number_entities(LO n, LO nold, SameNumbering same, ProdNumbering prods) {
  LO t = thread_index();
  if (t >= n) return;
  if (t < nold)
    number_same(t, same);
  else
    number_prods(t - nold, prods);
}

// one launch for a transferred tag: threads [0, nsame) copy same entities,
// the rest place the products' values (src/Omega_h_transfer.cpp:631-653)
template <class T>
__global__ void scatter_data(LO n, LO nsame, int nc, LO const* same2old, LO const* same2new,
    T const* old_data, LO const* prods2new, T const* pdata, T* data) {
  LO i = thread_index();
  if (i >= n) return;
  if (i < nsame) {
    for (int k = 0; k < nc; ++k) data[same2new[i] * nc + k] = old_data[same2old[i] * nc + k];
  } else {
    LO p = i - nsame;
    for (int k = 0; k < nc; ++k) data[prods2new[p] * nc + k] = pdata[p * nc + k];
  }
}

// class_id and class_dim of the new entities in one launch: same entities
// copy theirs, products inherit from their source entity of dimension
// src_dims[p] (or src_dim), src/Omega_h_transfer.cpp:20-34,212-267
__global__ void transfer_class_kernel(LO n, LO nsame, LO const* same2old, LO const* same2new,
    LO const* old_ids, I8 const* old_dims, LO const* prods2new, LO const* src, int src_dim,
    I8 const* src_dims, LO const* ids_a, I8 const* dims_a, LO const* ids_b, I8 const* dims_b,
    LO* ids, I8* dims) {
  LO i = thread_index();
  if (i >= n) return;
  if (i < nsame) {
    LO o = same2old[i], w = same2new[i];
    ids[w] = old_ids[o];
    dims[w] = old_dims[o];
    return;
  }
  LO p = i - nsame;
  int sd = src_dims ? src_dims[p] : src_dim;
  bool second = (sd != src_dim);
  LO s = src[p];
  LO w = prods2new[p];
  ids[w] = second ? ids_b[s] : ids_a[s];
  dims[w] = second ? dims_b[s] : dims_a[s];
}

// row p of out = row src[p] of in
__global__ void gather_rows(LO n, int nc, LO const* src, Real const* in, Real* out) {
  LO p = thread_index();
  if (p >= n) return;
  for (int k = 0; k < nc; ++k) out[std::size_t(p) * nc + k] = in[std::size_t(src[p]) * nc + k];
}

// transfer_pointwise_tmpl (src/Omega_h_transfer.cpp:515-563), one thread per
// key: each product takes the old element around the key (rows in the order
// of the upward adjacency) in which its centroid has the largest smallest
// barycentric coordinate, the first one on ties
template <int D>
__global__ void pointwise_kernel(LO nkeys, LO const* keys2kds, LO const* keys2prods,
    LO const* up_a2ab, LO const* up_ab2b, LO const* old_ev2v, Real const* old_coords,
    LO const* prod_verts, Real const* new_coords, LO* best) {
  LO key = thread_index();
  if (key >= nkeys) return;
  LO kd = keys2kds[key];
  for (LO p = keys2prods[key]; p < keys2prods[key + 1]; ++p) {
    double x[3];
    LO v0 = prod_verts[std::size_t(p) * (D + 1)];
    for (int j = 0; j < D; ++j) x[j] = new_coords[std::size_t(v0) * D + j];
    for (int i = 1; i <= D; ++i) {
      LO vi = prod_verts[std::size_t(p) * (D + 1) + i];
      for (int j = 0; j < D; ++j) x[j] = x[j] + new_coords[std::size_t(vi) * D + j];
    }
    for (int j = 0; j < D; ++j) x[j] = x[j] / double(D + 1);
    LO b = -1;
    double bd = 0.0;
    for (LO ke = up_a2ab[kd]; ke < up_a2ab[kd + 1]; ++ke) {
      LO e = up_ab2b[ke];
      double pp[D + 1][3];
      for (int i = 0; i <= D; ++i) {
        LO v = old_ev2v[std::size_t(e) * (D + 1) + i];
        for (int j = 0; j < D; ++j) pp[i][j] = old_coords[std::size_t(v) * D + j];
      }
      double xi[D + 1];
      barycentric<D>(x, pp, xi);
      double mn = xi[0];
      for (int i = 1; i <= D; ++i) mn = (mn < xi[i]) ? mn : xi[i];
      double dist = -mn;
      if (b == -1 || dist < bd) {
        b = e;
        bd = dist;
      }
    }
    best[p] = b;
  }
}

// an inheritance tag of the new entities, by the same sources as
// transfer_class_kernel
__global__ void inherit_kernel(LO n, LO nsame, int nc, LO const* same2old, LO const* same2new,
    Real const* old_vals, LO const* prods2new, LO const* src, int src_dim, I8 const* src_dims,
    Real const* vals_a, Real const* vals_b, Real* out) {
  LO i = thread_index();
  if (i >= n) return;
  if (i < nsame) {
    for (int k = 0; k < nc; ++k) out[same2new[i] * nc + k] = old_vals[same2old[i] * nc + k];
    return;
  }
  LO p = i - nsame;
  int sd = src_dims ? src_dims[p] : src_dim;
  Real const* from = (sd != src_dim) ? vals_b : vals_a;
  LO s = src[p];
  LO w = prods2new[p];
  for (int k = 0; k < nc; ++k) out[w * nc + k] = from[s * nc + k];
}

// one round of indset::find (src/Omega_h_indset.cpp:5-35).  In the first
// round the states come straight from the candidate marks (2 for a
// candidate, 0 otherwise), so no node is in the set yet.
template <bool FIRST>
__device__ __forceinline__ I8 indset_state(I8 const* old_state, LO x) {
  I8 s = old_state[x];
  return FIRST ? I8(s ? 2 : 0) : s;
}

// A node joins the set once every neighbor that beats it (higher quality,
// or the same quality and a higher index) is out, and leaves once one of
// them is in: a neighbor it beats can be in only after the node is out, so
// this is the round above restated over the beating neighbors alone.  The
// first round records, per candidate, which row positions hold candidates
// that beat it (bits of `dom`); later rounds read the states at those
// positions only, not the qualities.  A row longer than 64 positions is
// read whole every round.
constexpr int kDomBits = 64;

template <bool FIRST>
__global__ void indset_iteration(LO n, LO const* a2ab, LO const* ab2b, Real const* q,
    I8 const* old_state, I8* new_state, unsigned long long* dom) {
  LO v = thread_index();
  if (v >= n) return;
  I8 s = indset_state<FIRST>(old_state, v);
  new_state[v] = s;
  if (s != 2) return;
  LO b = a2ab[v], e = a2ab[v + 1];
  bool undecided = false;
  if (!FIRST && e - b <= kDomBits) {
    unsigned long long mask = dom[v];
    while (mask) {
      int j = __ffsll(static_cast<long long>(mask)) - 1;
      mask &= mask - 1;
      I8 su = old_state[ab2b[b + j]];
      if (su == 1) {
        new_state[v] = 0;
        return;
      }
      if (su == 2) undecided = true;
    }
  } else {
    unsigned long long mask = 0;
    Real qv = q[v];
    for (LO j = b; j < e; ++j) {
      LO u = ab2b[j];
      I8 su = indset_state<FIRST>(old_state, u);
      if (su == 0) continue;
      Real qu = q[u];
      bool beaten = (qu != qv) ? (qu < qv) : (u < v);
      if (beaten) continue;
      if (su == 1) {
        new_state[v] = 0;
        return;
      }
      undecided = true;
      if (FIRST && j - b < kDomBits) mask |= 1ull << (j - b);
    }
    if (FIRST) dom[v] = mask;
  }
  if (!undecided) new_state[v] = 1;
}

// the first round for edges with `lanes` threads per edge (a power of two,
// Tuning::first_lanes): each takes every lanes-th row position, and the
// group ORs its findings, so the row's gathers run side by side instead of
// one after another
template <int NEPC>
__global__ void indset_first_edges(LO nt, int lanes, LO const* e2c_a2ab, LO const* e2c_ab2b,
    LO const* c2e, Real const* q, I8 const* cands, I8* new_state, unsigned long long* dom) {
  LO t = thread_index();
  if (t >= nt) return;
  LO v = t / lanes;
  int lane = t % lanes;
  I8 s = cands[v] ? I8(2) : I8(0);
  unsigned long long mask = 0;
  int undecided = 0;
  if (s == 2) {
    LO b = e2c_a2ab[v], e = e2c_a2ab[v + 1];
    Real qv = q[v];
    LO np = (e - b) * NEPC;
    for (LO p = lane; p < np; p += lanes) {
      LO u = c2e[e2c_ab2b[b + p / NEPC] * NEPC + p % NEPC];
      if (u == v || !cands[u]) continue;
      Real qu = q[u];
      bool beaten = (qu != qv) ? (qu < qv) : (u < v);
      if (beaten) continue;
      undecided = 1;
      if (p < kDomBits) mask |= 1ull << p;
    }
  }
  for (int d = lanes / 2; d > 0; d >>= 1) {
    mask |= __shfl_xor(mask, d, 64);
    undecided |= __shfl_xor(undecided, d, 64);
  }
  if (lane != 0) return;
  new_state[v] = (s == 2 && !undecided) ? I8(1) : s;
  if (s == 2) dom[v] = mask;
}

// the same round for edges, whose star (edges_across_tris/tets,
// src/Omega_h_adj.cpp:532-589) is the set of other edges of the elements
// around the edge: read through the edge -> element adjacency instead of a
// built star (an edge met twice is checked twice, which changes nothing);
// row position (j, k) is bit j * NEPC + k of `dom`
template <int NEPC, bool FIRST>
__global__ void indset_iteration_edges(LO n, LO const* e2c_a2ab, LO const* e2c_ab2b,
    LO const* c2e, Real const* q, I8 const* old_state, I8* new_state, unsigned long long* dom) {
  LO v = thread_index();
  if (v >= n) return;
  I8 s = indset_state<FIRST>(old_state, v);
  new_state[v] = s;
  if (s != 2) return;
  LO b = e2c_a2ab[v], e = e2c_a2ab[v + 1];
  bool undecided = false;
  if (!FIRST && (e - b) * NEPC <= kDomBits) {
    unsigned long long mask = dom[v];
    while (mask) {
      int p = __ffsll(static_cast<long long>(mask)) - 1;
      mask &= mask - 1;
      I8 su = old_state[c2e[e2c_ab2b[b + p / NEPC] * NEPC + p % NEPC]];
      if (su == 1) {
        new_state[v] = 0;
        return;
      }
      if (su == 2) undecided = true;
    }
  } else {
    unsigned long long mask = 0;
    Real qv = q[v];
    for (LO j = b; j < e; ++j) {
      LO const* ce = c2e + e2c_ab2b[j] * NEPC;
#pragma unroll
      for (int k = 0; k < NEPC; ++k) {
        LO u = ce[k];
        if (u == v) continue;
        I8 su = indset_state<FIRST>(old_state, u);
        if (su == 0) continue;
        Real qu = q[u];
        bool beaten = (qu != qv) ? (qu < qv) : (u < v);
        if (beaten) continue;
        if (su == 1) {
          new_state[v] = 0;
          return;
        }
        undecided = true;
        int p = (j - b) * NEPC + k;
        if (FIRST && p < kDomBits) mask |= 1ull << p;
      }
    }
    if (FIRST) dom[v] = mask;
  }
  if (!undecided) new_state[v] = 1;
}

__global__ void mark_down_kernel(LO nl, LO const* a2ab, LO const* ab2b, I8 const* marks, I8* out) {
  LO l = thread_index();
  if (l >= nl) return;
  I8 r = 0;
  for (LO j = a2ab[l]; j < a2ab[l + 1]; ++j)
    if (marks[ab2b[j]]) r = 1;
  out[l] = r;
}

// the same marks by scattering from the marked high entities, for when the
// upward adjacency is not built
__global__ void mark_down_scatter(LO nh, int deg, LO const* hl2l, I8 const* marks, I8* out) {
  LO h = thread_index();
  if (h >= nh || !marks[h]) return;
  for (int j = 0; j < deg; ++j) out[hl2l[h * deg + j]] = 1;
}

__global__ void below_kernel(LO n, Real const* q, Real ceil, I8* out) {
  LO i = thread_index();
  if (i < n) out[i] = q[i] < ceil;
}

// graph_reduce over the dual graph with max (src/Omega_h_graph.cpp:62-80 on
// elements_across_sides, src/Omega_h_adj.cpp:591-657): the max over the
// elements across each side that has two, -128 for an element with none; a
// side with more than two elements (not a manifold mesh) is an error.  The
// input is `in`, or, when `q` is given, the sliver test q < ceil.
template <int NSPE>
__device__ __forceinline__ I8 dual_max_at(LO e, LO const* es2s,
    unsigned long long const* side_elems, I8 const* in, Real const* q, Real ceil, int* error) {
  I8 r = -128;
#pragma unroll
  for (int j = 0; j < NSPE; ++j) {
    unsigned long long w = side_elems[es2s[e * NSPE + j]];
    unsigned count = unsigned(w >> 32);
    if (count == 2) {
      LO o = LO(unsigned(w)) - e;
      I8 x = q ? I8(q[o] < ceil) : in[o];
      r = (r < x) ? x : r;
    } else if (count > 2) {
      *error = kNonManifoldSide;
    }
  }
  return r;
}

__global__ void dual_max_kernel(LO n, int nspe, LO const* es2s,
    unsigned long long const* side_elems, I8 const* in, Real const* q, Real ceil, I8* out,
    int* error) {
  LO e = thread_index();
  if (e >= n) return;
  out[e] = nspe == 4 ? dual_max_at<4>(e, es2s, side_elems, in, q, ceil, error)
                     : dual_max_at<3>(e, es2s, side_elems, in, q, ceil, error);
}

__global__ void degrees_kernel(LO n, LO const* a2ab, LO const* rows, LO* out) {
  LO i = thread_index();
  if (i >= n) return;
  LO r = rows ? rows[i] : i;
  out[i] = a2ab[r + 1] - a2ab[r];
}

}  // namespace

LOs degrees_of(Ctx& c, LOs const& a2ab, LOs const& rows) {
  LO n = rows.exists() ? rows.size() : a2ab.size() - 1;
  LOs out(c, n);
  OSH_LAUNCH(degrees_kernel, n, c, a2ab.data(), rows.exists() ? rows.data() : (LO const*)nullptr,
      out.data());
  return out;
}

void modify_plans(Ctx& c, DMesh& old, int key_dim, int d0, int d1, LOs const& keys2kds,
    LOs const* keys2prods, ModPlan* plans) {
  LO nkeys = keys2kds.size();
  LO nkd = old.nents[key_dim];
  KeyDimInit init{};
  Bytes key_marks;
  LOs first_up[4];
  bool above = false;
  for (int d = d0; d <= d1; ++d) {
    ModPlan& plan = plans[d - d0];
    plan.ent_dim = d;
    plan.key_dim = key_dim;
    plan.nold = old.nents[d];
    plan.reps = LOs(c, nkeys);
    // same entities: those not adjacent to a key (collect_same); without a
    // `same` array every old entity is the same
    if (d >= key_dim) {
      plan.same = Bytes(c, plan.nold);
      plan.counts = LOs(c, plan.nold);
    }
    if (d > key_dim) {
      first_up[d] = LOs(c, nkd);
      init.first_up[d] = first_up[d].data();
      above = true;
    }
    if (d == key_dim) {
      init.same = plan.same.data();
      init.counts = plan.counts.data();
    }
  }
  if (above) {
    key_marks = Bytes(c, nkd);
    init.key_marks = key_marks.data();
  }
  if (above || init.same) {
    OSH_LAUNCH(init_key_dim, nkd, c, init);
    ModPlan* kp = (key_dim >= d0 && key_dim <= d1) ? &plans[key_dim - d0] : nullptr;
    OSH_LAUNCH(mark_keys, nkeys, c, keys2kds.data(),
        kp ? keys2prods[key_dim].data() : (LO const*)nullptr, init.key_marks,
        kp ? kp->same.data() : (I8*)nullptr, kp ? kp->counts.data() : (LO*)nullptr,
        kp ? kp->reps.data() : (LO*)nullptr);
  }
  RepsJob jobs{};
  int njobs = 0;
  SameUpJobs up{};
  for (int d = d0; d <= d1; ++d) {
    ModPlan& plan = plans[d - d0];
    if (d > key_dim) {
      up.deg[up.n] = simplex_degree(d, key_dim);
      up.hl2l[up.n] = ask_down(c, old, d, key_dim, nullptr).data();
      up.same[up.n] = plan.same.data();
      up.counts[up.n] = plan.counts.data();
      up.first_up[up.n] = first_up[d].data();
      up.start[up.n + 1] = up.start[up.n] + plan.nold;
      ++up.n;
      jobs.first_up[njobs] = first_up[d].data();
    } else if (d < key_dim) {
      plan.counts = fill_lo(c, plan.nold, 1);
      jobs.deg[njobs] = simplex_degree(key_dim, d);
      jobs.down[njobs] = ask_down(c, old, key_dim, d, nullptr).data();
    } else {
      continue;
    }
    jobs.keys2prods[njobs] = keys2prods[d].data();
    jobs.reps[njobs] = plan.reps.data();
    jobs.counts[njobs] = plan.counts.data();
    ++njobs;
  }
  OSH_CHECK(njobs <= 3, "modify_plans: too many dimensions");
  if (up.n) OSH_LAUNCH(mark_same_up_kernel, up.start[up.n], c, key_marks.data(), up);
  if (njobs) OSH_LAUNCH(reps_kernel, njobs * nkeys, c, nkeys, keys2kds.data(), jobs,
      c.error_word());
}

void modify_plans_masked(Ctx& c, DMesh& old, int key_dim, int d1, Bytes const& key_mask,
    LO const* const* prod_counts, ModPlan* plans, LOs* rep_of) {
  LO nkd = old.nents[key_dim];
  KeyDimInit init{};
  LOs first_up[4];
  for (int d = key_dim; d <= d1; ++d) {
    ModPlan& plan = plans[d - key_dim];
    plan.ent_dim = d;
    plan.key_dim = key_dim;
    plan.nold = old.nents[d];
    plan.same = Bytes(c, plan.nold);
    plan.counts = LOs(c, plan.nold);
    if (d > key_dim) {
      first_up[d] = LOs(c, nkd);
      init.first_up[d] = first_up[d].data();
      rep_of[d] = LOs(c, nkd);
    }
  }
  init.same = plans[0].same.data();
  init.counts = plans[0].counts.data();
  init.key_mask = key_mask.data();
  init.key_prod_counts = prod_counts[key_dim];
  OSH_LAUNCH(init_key_dim, nkd, c, init);
  MaskedReps j{};
  int nj = 0;
  SameUpJobs up{};
  for (int d = key_dim + 1; d <= d1; ++d) {
    ModPlan& plan = plans[d - key_dim];
    up.deg[up.n] = simplex_degree(d, key_dim);
    up.hl2l[up.n] = ask_down(c, old, d, key_dim, nullptr).data();
    up.same[up.n] = plan.same.data();
    up.counts[up.n] = plan.counts.data();
    up.first_up[up.n] = first_up[d].data();
    up.start[up.n + 1] = up.start[up.n] + plan.nold;
    ++up.n;
    j.first_up[nj] = first_up[d].data();
    j.prod_counts[nj] = prod_counts[d];
    j.rep_of[nj] = rep_of[d].data();
    j.counts[nj] = plan.counts.data();
    ++nj;
  }
  OSH_CHECK(nj <= 3, "modify_plans_masked: too many dimensions");
  if (up.n) OSH_LAUNCH(mark_same_up_kernel, up.start[up.n], c, key_mask.data(), up);
  if (nj) OSH_LAUNCH(reps_masked, nj * nkd, c, nkd, key_mask.data(), j, c.error_word());
}

int modify_scan_jobs(Ctx& c, ModPlan* plans, int nplans, ScanJob* jobs, int first_slot) {
  int njobs = 0;
  for (int i = 0; i < nplans; ++i) {
    ModPlan& p = plans[i];
    p.offsets = LOs(c, p.nold + 1);
    jobs[njobs++] = {p.counts.data(), kScanLO, p.nold, p.offsets.data(),
        slot_ptr<LO>(c, first_slot + 2 * i)};
    if (p.same.exists()) {
      p.same_offsets = LOs(c, p.nold + 1);
      jobs[njobs++] = {p.same.data(), kScanNonzero, p.nold, p.same_offsets.data(),
          slot_ptr<LO>(c, first_slot + 2 * i + 1)};
    }
  }
  return njobs;
}

void modify_scan_totals(Ctx& c, ModPlan* plans, int nplans, int first_slot) {
  for (int i = 0; i < nplans; ++i) {
    plans[i].nnew = read_slot<LO>(c, first_slot + 2 * i);
    plans[i].nsame = plans[i].same.exists() ? read_slot<LO>(c, first_slot + 2 * i + 1)
                                             : plans[i].nold;
  }
}

void modify_scan(Ctx& c, ModPlan* plans, int nplans) {
  OSH_CHECK(nplans >= 1 && 2 * nplans <= kMaxScanJobs && 2 * nplans < Ctx::kSlots,
      "modify_scan: too many dimensions");
  ScanJob jobs[kMaxScanJobs];
  int njobs = modify_scan_jobs(c, plans, nplans, jobs, 1);
  offset_scans(c, jobs, njobs);
  c.sync();
  modify_scan_totals(c, plans, nplans, 1);
}

void modify_build(Ctx& c, DMesh& old, DMesh& nw, ModPlan const& plan, LOs const& keys2kds,
    LOs const& keys2prods, LO nprods, LOs const& prod_verts, LOs const& old_lows2new_lows,
    ModOut& out, LOs const* verts_old2new, LOs* prod_src, LOs const* given_down,
    Bytes const* given_codes) {
  int ent_dim = plan.ent_dim, key_dim = plan.key_dim;
  LO nkeys = keys2kds.size();
  LO nold = plan.nold;
  LO nsame = plan.nsame;
  Bytes const& same = plan.same;
  out.nnew = plan.nnew;
  LO nnew = out.nnew;
  out.same2old = LOs(c, nsame);
  out.same2new = LOs(c, nsame);
  out.old2new = LOs(c, nold);
  out.prods2new = LOs(c, nprods);
  LOs order;
  if (key_dim > ent_dim) {
    LO nkd = old.nents[key_dim];
    order = fill_lo(c, nkd, -1);
    LOs md2mod = fill_lo(c, nkd, -1);
    OSH_LAUNCH(scatter_index, nkeys, c, keys2kds.data(), md2mod.data());
    Adj const& up = ask_up(c, old, ent_dim, key_dim);
    OSH_LAUNCH(rep2md_order_kernel, nold, c, up.a2ab.data(), up.ab2b.data(), up.codes.data(),
        md2mod.data(), keys2prods.data(), order.data());
  }
  int deg = ent_dim + 1;
  bool with_codes = ent_dim >= 2;
  LOs down;
  Bytes codes;
  LOs pdown, found;
  Bytes pcodes;
  // the vertex lists of new faces and elements come with the numbering
  // (those of same entities renumbered, those of products as built), so the
  // new mesh does not derive them again
  LOs verts;
  LO const* old_verts = nullptr;
  if (ent_dim >= 1) {
    down = LOs(c, nnew * deg);
    if (with_codes) {
      codes = Bytes(c, nnew * deg);
      // a caller that knows where the products' sides are hands them over
      if (given_down) {
        pdown = *given_down;
        pcodes = *given_codes;
      } else {
        reflect_down(c, nw, ent_dim, nprods, prod_verts, pdown, pcodes, &found);
      }
      verts = LOs(c, nnew * deg);
      old_verts = verts_of(c, old, ent_dim).data();
    } else {
      pdown = prod_verts;
    }
  }
  if (prod_src) *prod_src = LOs(c, nprods);
  SameNumbering sn;
  sn.same = same.exists() ? same.data() : (I8 const*)nullptr;
  sn.same_offsets = same.exists() ? plan.same_offsets.data() : (LO const*)nullptr;
  sn.offsets = plan.offsets.data();
  sn.same2old = out.same2old.data();
  sn.same2new = out.same2new.data();
  sn.old2new = out.old2new.data();
  sn.deg = deg;
  sn.old_down = ent_dim >= 1 ? old.down[ent_dim].data() : (LO const*)nullptr;
  sn.old_codes = with_codes ? old.down_codes[ent_dim].data() : (I8 const*)nullptr;
  sn.old_lows2new_lows = ent_dim >= 1 ? old_lows2new_lows.data() : (LO const*)nullptr;
  sn.down = ent_dim >= 1 ? down.data() : (LO*)nullptr;
  sn.codes = with_codes ? codes.data() : (I8*)nullptr;
  sn.nvpe = deg;
  sn.old_verts = old_verts;
  sn.verts_old2new = verts_old2new ? verts_old2new->data() : (LO const*)nullptr;
  sn.verts = verts.exists() ? verts.data() : (LO*)nullptr;
  ProdNumbering pn;
  pn.reps = plan.reps.data();
  pn.keys2kds = keys2kds.data();
  pn.offsets = plan.offsets.data();
  pn.order = order.exists() ? order.data() : (LO const*)nullptr;
  pn.keys2prods = keys2prods.data();
  pn.prods2new = out.prods2new.data();
  pn.deg = deg;
  pn.pdown = ent_dim >= 1 ? pdown.data() : (LO const*)nullptr;
  pn.pcodes = with_codes ? pcodes.data() : (I8 const*)nullptr;
  pn.found = with_codes ? found.data() : (LO const*)nullptr;
  pn.down = ent_dim >= 1 ? down.data() : (LO*)nullptr;
  pn.codes = with_codes ? codes.data() : (I8*)nullptr;
  pn.nvpe = deg;
  pn.pverts = verts.exists() ? prod_verts.data() : (LO const*)nullptr;
  pn.verts = verts.exists() ? verts.data() : (LO*)nullptr;
  pn.key_src = keys2kds.data();
  pn.src = prod_src ? prod_src->data() : (LO*)nullptr;
  pn.error = c.error_word();
  OSH_LAUNCH(number_entities, nold + nkeys * kProdLanes, c, nold, sn, pn);
  nw.nents[ent_dim] = nnew;
  if (ent_dim == 0) return;
  nw.down[ent_dim] = down;
  nw.down_codes[ent_dim] = codes;
  if (verts.exists()) nw.verts[ent_dim] = verts;
}

template <class T>
DArray<T> transfer(Ctx& c, ModOut const& mo, int nc, DArray<T> const& old_data,
    DArray<T> const* prod_data) {
  DArray<T> d(c, mo.nnew * nc);
  LO nsame = mo.same2old.size();
  LO nprods = prod_data ? mo.prods2new.size() : 0;
  OSH_LAUNCH(scatter_data<T>, nsame + nprods, c, nsame, nc, mo.same2old.data(),
      mo.same2new.data(), old_data.data(), mo.prods2new.data(),
      prod_data ? prod_data->data() : (T const*)nullptr, d.data());
  return d;
}

template DArray<LO> transfer(Ctx&, ModOut const&, int, DArray<LO> const&, DArray<LO> const*);
template DArray<I8> transfer(Ctx&, ModOut const&, int, DArray<I8> const&, DArray<I8> const*);
template DArray<Real> transfer(Ctx&, ModOut const&, int, DArray<Real> const&, DArray<Real> const*);

void transfer_class(Ctx& c, DMesh& old, DMesh& nw, int d, ModOut const& mo,
    LOs const& prod_src, int src_dim, Bytes const* src_dims) {
  LO nprods = mo.prods2new.size();
  LO nsame = mo.same2old.size();
  int other = (src_dims && src_dim + 1 <= old.dim) ? src_dim + 1 : src_dim;
  LOs ids(c, mo.nnew);
  Bytes dims(c, mo.nnew);
  OSH_LAUNCH(transfer_class_kernel, nsame + nprods, c, nsame, mo.same2old.data(),
      mo.same2new.data(), old.class_id[d].data(), old.class_dim[d].data(),
      mo.prods2new.data(), prod_src.data(), src_dim,
      src_dims ? src_dims->data() : (I8 const*)nullptr, old.class_id[src_dim].data(),
      old.class_dim[src_dim].data(), old.class_id[other].data(), old.class_dim[other].data(),
      ids.data(), dims.data());
  nw.class_id[d] = ids;
  nw.class_dim[d] = dims;
  if (nw.itags.size() != old.itags.size()) nw.itags = old.itags;
  for (std::size_t t = 0; t < old.itags.size(); ++t) {
    auto const& a = old.itags[t];
    Reals out(c, mo.nnew * a.ncomps);
    OSH_LAUNCH(inherit_kernel, nsame + nprods, c, nsame, a.ncomps, mo.same2old.data(),
        mo.same2new.data(), a.data[d].data(), mo.prods2new.data(), prod_src.data(), src_dim,
        src_dims ? src_dims->data() : (I8 const*)nullptr, a.data[src_dim].data(),
        a.data[other].data(), out.data());
    nw.itags[t].data[d] = out;
  }
}

void transfer_elem_tags(Ctx& c, DMesh const& old, DMesh& nw, ModOut const& mo,
    LOs const* prod_src) {
  nw.etags.clear();
  LO nprods = mo.prods2new.size();
  for (auto const& t : old.etags) {
    Reals pd;
    if (prod_src) {
      pd = Reals(c, nprods * t.ncomps);
      OSH_LAUNCH(gather_rows, nprods, c, t.ncomps, prod_src->data(), t.data.data(), pd.data());
    }
    nw.etags.push_back(t.with_data(transfer<Real>(c, mo, t.ncomps, t.data,
        prod_src ? &pd : nullptr)));
  }
}

void transfer_pointwise(Ctx& c, DMesh& old, DMesh& nw, ModOut const& mo, int key_dim,
    LOs const& keys2kds, LOs const& keys2prods, LOs const& prod_verts) {
  if (old.etags.empty()) return;
  int dim = old.dim;
  LO nkeys = keys2kds.size();
  Adj const& up = ask_up(c, old, key_dim, dim);
  LOs const& ev2v = verts_of(c, old, dim);
  LOs best(c, mo.prods2new.size());
  if (dim == 3)
    OSH_LAUNCH(pointwise_kernel<3>, nkeys, c, keys2kds.data(), keys2prods.data(),
        up.a2ab.data(), up.ab2b.data(), ev2v.data(), old.coords.data(), prod_verts.data(),
        nw.coords.data(), best.data());
  else
    OSH_LAUNCH(pointwise_kernel<2>, nkeys, c, keys2kds.data(), keys2prods.data(),
        up.a2ab.data(), up.ab2b.data(), ev2v.data(), old.coords.data(), prod_verts.data(),
        nw.coords.data(), best.data());
  transfer_elem_tags(c, old, nw, mo, &best);
}

void transfer_measures(Ctx& c, DMesh& old, DMesh& nw, int d, ModOut const& mo) {
  LO nprods = mo.prods2new.size();
  LO nsame = mo.same2old.size();
  if (d == 1)
    nw.length = measure_transfer(c, nw, 1, nsame, mo.same2old.data(), mo.same2new.data(),
        ask_lengths(c, old).data(), nprods, mo.prods2new.data());
  if (d == old.dim)
    nw.quality = measure_transfer(c, nw, d, nsame, mo.same2old.data(), mo.same2new.data(),
        ask_qualities(c, old).data(), nprods, mo.prods2new.data());
}

Bytes find_indset(Ctx& c, DMesh& m, int d, Reals const& quality, Bytes const& cands,
    bool* any_cands, Adj const* edge_elems, Adj const* star) {
  Range range(c, "find_indset");
  LO n = m.nents[d];
  if (d == m.dim) {  // src/Omega_h_indset.cpp:29
    if (any_cands) *any_cands = max_i8(c, cands) > 0;
    return cands;
  }
  Adj const* g = nullptr;
  Adj const* e2c = nullptr;
  LO const* c2e = nullptr;
  if (d == 1) {
    e2c = edge_elems ? edge_elems : &ask_up(c, m, 1, m.dim);
    c2e = ask_down(c, m, m.dim, 1, nullptr).data();
  } else {
    g = star ? star : &ask_star(c, m, d);
  }
  Bytes a(c, n), b(c, n);
  DArray<unsigned long long> dom(c, n);
  // rounds leave decided nodes as they are and take no node from the set, so
  // extra rounds after convergence change nothing: run two rounds before each
  // of the first four checks, then 4, 8 and 16 (long chains of decreasing
  // quality, as on the first coarsening passes of a large smooth mesh, take
  // some 200 rounds and would otherwise spend a sync every two); after the
  // first check a state of 2 means undecided nodes remain, and a state of 1
  // or 2 that there was a candidate (the best candidate joins the set in the
  // first round)
  int first_lanes = fit_lanes(n, c.tuning.first_lanes);
  for (LO round = 0;;) {
    OSH_CHECK(round <= n + 1, "independent set does not converge");
    int batch = round < 8 ? 2 : (round < 16 ? 4 : (round < 32 ? 8 : 16));
    for (int k = 0; k < batch; ++k, ++round) {
      bool first = round == 0;
      I8 const* in = first ? cands.data() : a.data();
#define OSH_IS(F)                                                                         \
  do {                                                                                    \
    if (g)                                                                                \
      OSH_LAUNCH(indset_iteration<F>, n, c, g->a2ab.data(), g->ab2b.data(), quality.data(), \
          in, b.data(), dom.data());                                                      \
    else if (m.dim == 3)                                                                  \
      OSH_LAUNCH((indset_iteration_edges<6, F>), n, c, e2c->a2ab.data(), e2c->ab2b.data(),  \
          c2e, quality.data(), in, b.data(), dom.data());                                 \
    else                                                                                  \
      OSH_LAUNCH((indset_iteration_edges<3, F>), n, c, e2c->a2ab.data(), e2c->ab2b.data(),  \
          c2e, quality.data(), in, b.data(), dom.data());                                 \
  } while (0)
      if (first && !g && m.dim == 3)
        OSH_LAUNCH((indset_first_edges<6>), n * first_lanes, c, first_lanes, e2c->a2ab.data(),
            e2c->ab2b.data(), c2e, quality.data(), cands.data(), b.data(), dom.data());
      else if (first && !g)
        OSH_LAUNCH((indset_first_edges<3>), n * first_lanes, c, first_lanes, e2c->a2ab.data(),
            e2c->ab2b.data(), c2e, quality.data(), cands.data(), b.data(), dom.data());
      else if (first)
        OSH_IS(true);
      else
        OSH_IS(false);
#undef OSH_IS
      std::swap(a, b);
    }
    I8 top = max_i8(c, a);
    if (any_cands && round == 2) *any_cands = top >= 1;
    if (top != 2) break;
  }
  return a;
}

Bytes mark_down(Ctx& c, DMesh& m, int high, int low, Bytes const& marks) {
  if (high == low) return marks;
  if (!m.up[low][high].exists()) {
    Bytes out = fill_i8(c, m.nents[low], 0);
    LOs const& hl2l = ask_down(c, m, high, low, nullptr);
    OSH_LAUNCH(mark_down_scatter, m.nents[high], c, simplex_degree(high, low), hl2l.data(),
        marks.data(), out.data());
    return out;
  }
  Adj const& l2h = ask_up(c, m, low, high);
  Bytes out(c, m.nents[low]);
  OSH_LAUNCH(mark_down_kernel, m.nents[low], c, l2h.a2ab.data(), l2h.ab2b.data(), marks.data(),
      out.data());
  return out;
}

Bytes mark_sliver_layers(Ctx& c, DMesh& m, Real qual_ceil, int nlayers) {
  LO ne = m.nents[m.dim];
  Reals const& q = ask_qualities(c, m);
  Bytes marks;
  if (nlayers == 0) {
    marks = Bytes(c, ne);
    OSH_LAUNCH(below_kernel, ne, c, q.data(), qual_ceil, marks.data());
    return marks;
  }
  auto const& se = ask_side_elems(c, m);
  // the first layer reads the sliver test itself
  for (int i = 0; i < nlayers; ++i) {
    Bytes next(c, ne);
    OSH_LAUNCH(dual_max_kernel, ne, c, m.dim + 1, m.down[m.dim].data(), se.data(), marks.data(),
        i == 0 ? q.data() : (Real const*)nullptr, qual_ceil, next.data(), c.error_word());
    marks = next;
  }
  return marks;
}

void default_options(int dim, AdaptOptions& o) {
  o.min_length_desired = 1.0 / __builtin_sqrt(2.0);
  o.max_length_desired = __builtin_sqrt(2.0);
  o.max_length_allowed = o.max_length_desired * 2.0;
  if (dim == 3) {
    o.min_quality_allowed = 0.20;
    o.min_quality_desired = 0.30;
  } else {
    o.min_quality_allowed = 0.30;
    o.min_quality_desired = 0.40;
  }
  o.nsliver_layers = 4;
  o.verbosity = 2;  // EACH_REBUILD
  o.should_refine = 1;
  o.should_coarsen = 1;
  o.should_swap = 1;
  o.should_coarsen_slivers = 1;
  o.length_histogram_min = 0.0;
  o.length_histogram_max = 3.0;
  o.nlength_histogram_bins = 10;
  o.nquality_histogram_bins = 10;
  o.should_prevent_coarsen_flip = 0;
  o.precision = 0;
}

}  // namespace osh

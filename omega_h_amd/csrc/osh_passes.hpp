// Internal interfaces shared by the adapt() passes.
#pragma once

#include <functional>

#include "osh_mesh.hpp"

namespace osh {

// mirrors Omega_h::AdaptOpts (src/Omega_h_adapt.hpp:50-77)
struct AdaptOptions {
  Real min_length_desired;
  Real max_length_desired;
  Real max_length_allowed;
  Real min_quality_allowed;
  Real min_quality_desired;
  int nsliver_layers;
  int verbosity;
  int should_refine;
  int should_coarsen;
  int should_swap;
  int should_coarsen_slivers;
  Real length_histogram_min;
  Real length_histogram_max;
  int nlength_histogram_bins;
  int nquality_histogram_bins;
  int should_prevent_coarsen_flip;
  // 0 FP64; 1 mixed: cavity screens in FP32, every cavity the screen does
  // not reject by a margin re-evaluated in FP64 (same mesh, bit for bit)
  int precision;
};

void default_options(int dim, AdaptOptions& o);

// result of modify_ents for one dimension (src/Omega_h_modify.cpp:468-517)
struct ModOut {
  LO nnew = 0;
  LOs same2old;
  LOs same2new;
  LOs prods2new;
  LOs old2new;
};

// modify_ents_adapt (src/Omega_h_modify.cpp:468-517) in three steps, so
// that a pass numbers all its dimensions with one scan launch and one sync:
// modify_plans marks the same entities, counts the new entities each old one
// stands for and finds the representatives (no sync); modify_scan scans the
// plans of up to four dimensions together; modify_build numbers and connects
// one dimension
struct ModPlan {
  int ent_dim = 0, key_dim = 0;
  LO nold = 0, nnew = 0, nsame = 0;
  Bytes same;
  LOs counts, reps, offsets, same_offsets;
};
// the plans of dimensions d0..d1 (keys2prods indexed by dimension)
void modify_plans(Ctx& c, DMesh& old, int key_dim, int d0, int d1, LOs const& keys2kds,
    LOs const* keys2prods, ModPlan* plans);
void modify_scan(Ctx& c, ModPlan* plans, int nplans);
// the same plans for dimensions key_dim..d1 with the keys given as a mask
// over the key dimension and their product counts per key-dimension entity
// (prod_counts[d]); plans' reps stay empty and rep_of[d] (d > key_dim)
// receives each key's representative by key-dimension entity, so that the
// keys need not be listed before the scan
void modify_plans_masked(Ctx& c, DMesh& old, int key_dim, int d1, Bytes const& key_mask,
    LO const* const* prod_counts, ModPlan* plans, LOs* rep_of);
// modify_scan in two halves, for callers that add the plans' scans to a
// launch of their own: the jobs (totals in pinned slots first_slot...), then,
// after the caller's sync, the totals
int modify_scan_jobs(Ctx& c, ModPlan* plans, int nplans, ScanJob* jobs, int first_slot);
void modify_scan_totals(Ctx& c, ModPlan* plans, int nplans, int first_slot);
// verts_old2new: the vertex renumbering (null when vertices keep their
// numbers); prod_src, when given, receives each product's key entity
void modify_build(Ctx& c, DMesh& old, DMesh& nw, ModPlan const& plan, LOs const& keys2kds,
    LOs const& keys2prods, LO nprods, LOs const& prod_verts, LOs const& old_lows2new_lows,
    ModOut& out, LOs const* verts_old2new, LOs* prod_src = nullptr,
    LOs const* given_down = nullptr, Bytes const* given_codes = nullptr);

// new tag data: same entities copy old values, products take prod values
template <class T>
DArray<T> transfer(Ctx& c, ModOut const& mo, int ncomps, DArray<T> const& old_data,
    DArray<T> const* prod_data);
// class_id/class_dim of products from one old entity each, of dimension
// src_dims[p] (or src_dim when src_dims is null)
void transfer_class(Ctx& c, DMesh& old, DMesh& nw, int d, ModOut const& mo,
    LOs const& prod_src, int src_dim, Bytes const* src_dims);
// lengths (d == 1) and qualities (d == dim) of products on the new mesh
void transfer_measures(Ctx& c, DMesh& old, DMesh& nw, int d, ModOut const& mo);
// element tags of the new elements: products take the value of old element
// prod_src[p] (no products when prod_src is null)
void transfer_elem_tags(Ctx& c, DMesh const& old, DMesh& nw, ModOut const& mo,
    LOs const* prod_src);
// transfer_pointwise (src/Omega_h_transfer.cpp:515-583) of the element tags
// for coarsening (key_dim 0) and swap (key_dim 1): prod_verts are the
// products' vertices in the new mesh, whose coordinates are already set
void transfer_pointwise(Ctx& c, DMesh& old, DMesh& nw, ModOut const& mo, int key_dim,
    LOs const& keys2kds, LOs const& keys2prods, LOs const& prod_verts);

// conservative and momentum transfer (osh_conserve.hip): setup_conservation_tags
// (src/Omega_h_conserve.cpp:578-606) zeroes the integral errors at the start
// of adapt(); conserve_pass transfers densities by intersection and the
// errors through one pass (op 0 refine, 1 coarsen, 2 swap; keys of
// dimension key_dim, k2p their element products, after the other element
// transfers of the pass); conserve_correct is correct_integral_errors
// (src/Omega_h_conserve.cpp:799-829) at the end of adapt()
bool conserve_any(DMesh const& m);
void conserve_setup(Ctx& c, DMesh& m);
void conserve_pass(Ctx& c, DMesh& old, DMesh& nw, int op, int key_dim, LOs const& keys2kds,
    LOs const& k2p, ModOut const& mo);
constexpr int kMaxDiffusionSweeps = 100000;
void conserve_correct(Ctx& c, DMesh& m, int max_sweeps);

// *any_cands (when given) tells whether any entity was a candidate; for
// edges, `edge_elems` (when given) is an edge -> element adjacency holding at
// least the rows of the candidate edges; for vertices, `star` (when given) is
// the vertex star holding at least the rows of the candidates
Bytes find_indset(Ctx& c, DMesh& m, int d, Reals const& quality, Bytes const& cands,
    bool* any_cands = nullptr, Adj const* edge_elems = nullptr, Adj const* star = nullptr);
Bytes mark_down(Ctx& c, DMesh& m, int high, int low, Bytes const& marks);
Bytes mark_sliver_layers(Ctx& c, DMesh& m, Real qual_ceil, int nlayers);
LOs degrees_of(Ctx& c, LOs const& a2ab, LOs const& rows);

bool refine_by_size(Ctx& c, DMesh& m, AdaptOptions const& o);
bool coarsen_by_size(Ctx& c, DMesh& m, AdaptOptions const& o);
bool coarsen_slivers(Ctx& c, DMesh& m, AdaptOptions const& o);
// The 3D swap choice of each edge (its configuration, whether it improves
// on the edge's elements and the quality it reaches) is a function of the
// tets around the edge, the coordinates and metric of their vertices and
// max_length_allowed only.  Between swap passes of one adapt() nothing else
// moves, so the choices of edges whose tets a pass did not touch carry over
// to the next mesh; `valid` marks them, per edge of the current mesh.
struct SwapCache {
  Bytes valid;
  Bytes cfg;
  Bytes keep;
  Reals q;
  void clear() { *this = SwapCache(); }
};

// with a cache (only for consecutive swap passes of one adapt(), cleared by
// the caller whenever anything else changes the mesh, its metric or the
// options) the 3D pass evaluates only the candidates not carried over
bool swap_edges(Ctx& c, DMesh& m, AdaptOptions const& o, SwapCache* cache = nullptr);
// the swap choices of given interior edges of a 3D mesh (swap3d_qualities
// and the choice of a triangulation): per edge the chosen triangulation (-1
// for none) and the quality of a triangulation that improves on the edge's
// elements (-1 otherwise)
void swap3d_choices(Ctx& c, DMesh& m, AdaptOptions const& o, LOs const& edges,
    LOs& configs_out, Reals& quals_out);

// known-answer hooks used by the tests
Reals refine_qualities(Ctx& c, DMesh& m, LOs const& cands);
// coarsen_qualities (src/Omega_h_coarsen_qualities.cpp:65-88) on device
// arrays, stream-ordered: out[2 i + col] is the quality after collapsing
// candidate i's edge in direction col (-1 where cand_codes[i] does not
// collapse that way).  precision 1 screens in FP32 first: a direction whose
// estimate falls below qmin by more than the band gets -1 instead of its
// FP64 value (the filter that follows rejects it either way).
void coarsen_qualities(Ctx& c, DMesh& m, LO ncands, LO const* cands, I8 const* cand_codes,
    Real* out, int precision, Real qmin);
void find_loops(Ctx& c, DMesh& m, LOs const& edges, LOs& eev2v, LOs& loop_verts, LOs& sizes);

// the batched form (osh_batch.hip): meshes joined into one disjoint mesh
// (part k's entities of every dimension after those of parts < k; the
// vertex tag kPartTag holds each vertex's part), adapted as one, split back
extern char const* const kPartTag;
void join_meshes(Ctx& c, std::vector<DMesh const*> const& parts, DMesh& out);
// first entity and entity count of each part in every dimension; an error
// unless each part is one block and the blocks follow the part order
void part_ranges(Ctx& c, DMesh& m, int nparts, std::vector<LO> first[4],
    std::vector<LO> count[4]);
void split_mesh(Ctx& c, DMesh& m, int nparts, std::vector<DMesh>& out);
// with Ctx::part_changed set, marks the parts of the pass's keys (the part
// of a key's first vertex) as changed; called by every pass with its keys
void note_changed_parts(Ctx& c, DMesh const& m, int key_dim, LOs const& keys2kds);
// the minimum element quality of each part of a joined mesh (a part
// without elements gets +infinity)
std::vector<Real> part_min_quality(Ctx& c, DMesh& m, int nparts);

struct PassRecord {
  int kind;  // 0 refine, 1 coarsen, 2 swap, 3 coarsen slivers
  LO nelems_in;
  LO nelems_out;
  double seconds;
};

// called with the mesh after every pass that changed it (outside the pass's
// timing), for callers that record the mesh pass by pass
using PassHook = std::function<void(DMesh&)>;
bool adapt(Ctx& c, DMesh& m, AdaptOptions const& o, std::vector<PassRecord>* log,
    PassHook const* hook = nullptr);
// adapt() of many meshes at once (the batched form): joined into one
// disjoint mesh whose passes cover every mesh in one launch per kernel, with
// the parts that can no longer change set aside as they appear; each mesh
// ends where adapt() alone leaves it.  Returns adapt()'s result per mesh.
std::vector<int> adapt_batch(Ctx& c, std::vector<DMesh>& meshes, AdaptOptions const& o,
    std::vector<PassRecord>* log);
// approach_metric (src/Omega_h_approach.cpp:83-126): move the metric toward
// the "target_metric" vertex tag; false when there is no target
bool approach_metric(Ctx& c, DMesh& m, AdaptOptions const& o, Real min_step);
// Mesh::set_coords (src/Omega_h_mesh.cpp:365-368): drops the cached lengths,
// qualities and sizes
void set_coords(DMesh& m, Reals const& v);
// warp_to_limit (src/Omega_h_approach.cpp:36-81): move the vertices by the
// "warp" vertex tag as far as quality and lengths allow; false without one
bool warp_to_limit(Ctx& c, DMesh& m, AdaptOptions const& o, bool exit_on_stall, int max_niters);
// get_implied_isos (src/Omega_h_metric.cpp:316-321): nverts values
Reals implied_isos(Ctx& c, DMesh& m);
// compute_ill_metric_dim, fix_for_given_metric, fix, grade_fix_adapt
// (src/Omega_h_fix.cpp:16-176); isotropy 0 anisotropic, 1 iso length,
// 2 iso size; the fix functions return the number of adapt() calls
void ill_metric(Ctx& c, DMesh& m, AdaptOptions const& o, int isotropy);
int fix_for_given_metric(Ctx& c, DMesh& m, AdaptOptions const& o);
int fix(Ctx& c, DMesh& m, AdaptOptions const& o, int isotropy);
int grade_fix_adapt(Ctx& c, DMesh& m, AdaptOptions const& o, int ncomps, Reals const& target);

}  // namespace osh

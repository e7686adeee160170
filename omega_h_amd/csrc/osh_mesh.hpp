// Device-resident simplex mesh: the Omega_h::Mesh data model
// (src/Omega_h_mesh.hpp:36-177) with downward adjacencies stored and every
// other adjacency derived on demand and cached, as Mesh::ask_adj does
// (src/Omega_h_mesh.cpp:307-357).
#pragma once

#include <string>

#include "osh_runtime.hpp"

namespace osh {

// CSR adjacency (up adjacency or graph); codes may be absent
struct Adj {
  LOs a2ab;
  LOs ab2b;
  Bytes codes;
  bool exists() const { return ab2b.exists(); }
};

// how the integral error of a conserved quantity is diffused over the
// element graph before adapt() corrects the field (VarCompareOpts
// src/Omega_h_compare.hpp:14-21 as diffuse_integrals_weighted
// src/Omega_h_conserve.cpp:708-734 uses it): type 0 none, 1 relative,
// 2 absolute; sweeps run until every weighted error is at most `tol`
struct Diffusion {
  int type = 0;
  double tol = 0.0;
  double floor = 0.0;
};

struct DMesh {
  int dim = 0;
  LO nents[4] = {-1, -1, -1, -1};
  // down[d]: entity of dim d -> (d+1) entities of dim d-1; down[1] = edge verts
  LOs down[4];
  Bytes down_codes[4];
  Reals coords;
  Reals metric;
  int metric_ncomps = 0;
  LOs class_id[4];
  Bytes class_dim[4];
  // vertex tags carried through adapt(): kind 0 values are averaged at
  // refinement midpoints (average_field, src/Omega_h_mesh.cpp:822-844), kind 1
  // values get log-Euclidean midpoint metrics (get_mident_metrics,
  // src/Omega_h_metric.cpp:56-104); surviving vertices keep theirs
  // kind 2: a momentum velocity (OMEGA_H_MOMENTUM_VELOCITY), interpolated
  // like kind 0, whose momentum with the conserved element density `density`
  // is conserved; `err` holds the elements' momentum errors during adapt()
  struct VertexTag {
    std::string name;
    int ncomps;
    int kind;
    Reals data;
    std::string density;
    Diffusion diffuse;
    Reals err;
    // the same tag carrying other values (errors are not carried along)
    VertexTag with_data(Reals d) const {
      VertexTag t = *this;
      t.data = d;
      t.err = Reals();
      return t;
    }
  };
  std::vector<VertexTag> vtags;
  // tags on every dimension transferred by inheritance (OMEGA_H_INHERIT,
  // src/Omega_h_transfer.cpp:20-34): products take the value of the entity
  // they came from, exactly as class_id does
  struct InheritTag {
    std::string name;
    int ncomps;
    Reals data[4];
  };
  std::vector<InheritTag> itags;
  // element tags transferred pointwise (OMEGA_H_POINTWISE, should_fit
  // src/Omega_h_transfer.cpp:47-54): refinement products take their parent
  // element's value; coarsening and swap products take the value of the old
  // element of the cavity nearest to their centroid
  // (transfer_pointwise_tmpl, src/Omega_h_transfer.cpp:515-583)
  // kind 0: pointwise; kind 1: a density whose integral is conserved
  // (OMEGA_H_CONSERVE; `err` holds the elements' integral errors during
  // adapt() and `old` its values before the final correction); kind 2: a
  // density transferred by intersection (OMEGA_H_DENSITY)
  struct ElemTag {
    std::string name;
    int ncomps;
    Reals data;
    int kind = 0;
    Diffusion diffuse;
    Reals err;
    Reals old;
    ElemTag with_data(Reals d) const {
      ElemTag t = *this;
      t.data = d;
      t.err = Reals();
      t.old = Reals();
      return t;
    }
  };
  std::vector<ElemTag> etags;
  Reals length;   // cached edge lengths (may be absent)
  Reals quality;  // cached element qualities (may be absent)
  // derived caches
  LOs verts[4];
  LOs t2e;
  Bytes t2e_codes;
  Adj up[4][4];
  Adj star[2];
  // side_elems[s]: the number of elements on side s (high 32 bits) and the
  // sum of their indices (low 32 bits)
  DArray<unsigned long long> side_elems;
  Adj dual;    // element -> elements across sides (on demand, for callers)
  Reals size;  // element sizes (on demand)
  void clear_caches();
};

int simplex_degree(int from_dim, int to_dim);

LOs const& verts_of(Ctx& c, DMesh& m, int d);
LOs const& ask_down(Ctx& c, DMesh& m, int high, int low, Bytes const** codes);
Adj const& ask_up(Ctx& c, DMesh& m, int low, int high);
// the upward adjacency with only the rows of marked lows filled (the others
// are empty; entries past a2ab[nlow] are unused); not cached.  With
// `marked`, also the list of marked lows (collect_marked), from the same scan
// launch.  Without `sorted`, a row's order is arbitrary, which suits uses
// that do not depend on it (counts, any/all tests, min/max)
// `marked_index` (with `marked`) receives, for each marked low, its position
// in the list (the scan of the marks)
Adj marked_up(Ctx& c, DMesh& m, int low, int high, Bytes const& marks, LOs* marked = nullptr,
    bool sorted = true, LOs* marked_index = nullptr);
Adj const& ask_star(Ctx& c, DMesh& m, int d);
DArray<unsigned long long> const& ask_side_elems(Ctx& c, DMesh& m);
// Mesh::ask_dual (src/Omega_h_mesh.cpp:260): elements_across_sides
Adj const& ask_dual(Ctx& c, DMesh& m);
// Mesh::ask_sizes (src/Omega_h_mesh.cpp:390-396): signed element areas or
// volumes (measure_elements_real, src/Omega_h_shape.cpp:55-89)
Reals const& ask_sizes(Ctx& c, DMesh& m);
// vertex lists of new entities of dimension high_dim >= 2 matched against the
// already-built dimension high_dim-1 of m (src/Omega_h_adj.cpp:424-441)
// With found_out, the per-use match counts come back unchecked and the caller
// must treat a use whose count is not 1 as an error (and not read its match).
void reflect_down(Ctx& c, DMesh& m, int high_dim, LO nprods, LOs const& prod_verts,
    LOs& down, Bytes& codes, LOs* found_out = nullptr);
// invert_adj (src/Omega_h_adj.cpp:231-263) of any high -> low map with deg
// entries per high entity: rows sorted by high index, codes carrying the
// down codes' flip and rotation and the use's position
Adj invert_adj_rows(Ctx& c, LO nhigh, int deg, LO nlow, LOs const& hl2l,
    Bytes const* down_codes);
// unique lower entities of element vertex lists (src/Omega_h_adj.cpp:133-153)
LOs find_unique(Ctx& c, LO nhigh, int high_dim, int low_dim, LOs const& hv2v);

// metric preparation (osh_metric.hip): limit_metric_gradation on the mesh's
// metric, returning the number of sweeps; elementwise intersection of n
// metrics of dimension dim
int limit_metric_gradation(Ctx& c, DMesh& m, Real max_rate, Real tol);
void intersect_metric_arrays(Ctx& c, int dim, LO n, Reals const& a, Reals const& b,
    Reals& out);
// get_implied_metrics (src/Omega_h_metric.cpp:280-309): nverts x (3 or 6)
Reals implied_metrics(Ctx& c, DMesh& m);
// interpolate_between_metrics (src/Omega_h_metric.cpp:106-111) of n metrics
void interpolate_metric_arrays(Ctx& c, int dim, LO n, Reals const& a, Reals const& b, Real t,
    Reals& out);

// metric measures (osh_measure.hip)
void launch_edge_lengths(Ctx& c, int dim, int md, LO n, LO const* edges, LO const* ev2v,
    Real const* coords, Real const* metric, Real* out);
void launch_elem_qualities(Ctx& c, int dim, int md, LO n, LO const* elems, LO const* ev2v,
    Real const* coords, Real const* metric, Real* out);
void measure_edges(Ctx& c, DMesh& m, LO n, LO const* edges, Real* out);
// the FP32 screening estimates of the mixed-precision form for every edge
// length and element quality, with the elements whose max-determinant
// metric was not clear marked in `unsure`
void screen_estimates(Ctx& c, DMesh& m, Reals& lengths32, Reals& quals32, Bytes& unsure);
void measure_elems(Ctx& c, DMesh& m, LO n, LO const* elems, Real* out);
// the length (d = 1) or quality (d = dim) tag of a rebuilt mesh: same
// entities copy old values, products are measured on the new mesh
Reals measure_transfer(Ctx& c, DMesh& nw, int d, LO nsame, LO const* same2old,
    LO const* same2new, Real const* old_vals, LO nprods, LO const* prods2new);
Reals const& ask_lengths(Ctx& c, DMesh& m);
Reals const& ask_qualities(Ctx& c, DMesh& m);
void mident_metrics(Ctx& c, DMesh& m, LO n, LO const* edges, Real* out);
// the same for any per-vertex metric array with `ncomps` components
void mident_values(Ctx& c, DMesh& m, LO n, LO const* edges, Real const* data, int ncomps,
    Real* out);

// the .osh directory format, one rank (osh_file.cpp)
void write_osh(Ctx& c, DMesh& m, std::string const& dir);
void read_osh(Ctx& c, DMesh& m, std::string const& dir);

// build_box (osh_build.hip)
void build_box(Ctx& c, DMesh& m, int dim, Real x, Real y, Real z, LO nx, LO ny, LO nz,
    bool reorder_and_classify);
void build_from_elems(Ctx& c, DMesh& m, int dim, LO nelems, LOs const& ev2v, LO nverts);

// ---- device-side templates (src/Omega_h_simplex.hpp, src/Omega_h_align.hpp) ----
__host__ __device__ __forceinline__ int down_template(int elem_dim, int bdry_dim,
    int which_bdry, int which_vert) {
  if (bdry_dim == 0) return which_bdry;
  if (elem_dim == 3 && bdry_dim == 2) {
    // faces {0,2,1},{0,1,3},{1,2,3},{2,0,3}
    int const t[4][3] = {{0, 2, 1}, {0, 1, 3}, {1, 2, 3}, {2, 0, 3}};
    return t[which_bdry][which_vert];
  }
  if (elem_dim == 3 && bdry_dim == 1) {
    int const t[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
    return t[which_bdry][which_vert];
  }
  // triangle edges {0,1},{1,2},{2,0}
  return (which_vert == 0) ? which_bdry : ((which_bdry + 1) % 3);
}

__host__ __device__ __forceinline__ int opposite_template(int elem_dim, int bdry_dim,
    int which_bdry) {
  if (elem_dim == 3) {
    if (bdry_dim == 0) {
      int const t[4] = {2, 3, 1, 0};
      return t[which_bdry];
    }
    if (bdry_dim == 1) {
      int const t[6] = {5, 3, 4, 1, 2, 0};
      return t[which_bdry];
    }
    int const t[4] = {3, 2, 0, 1};
    return t[which_bdry];
  }
  if (elem_dim == 2) {
    if (bdry_dim == 0) return (which_bdry + 1) % 3;
    return (which_bdry + 2) % 3;
  }
  return 1 - which_bdry;
}

__host__ __device__ __forceinline__ int code_is_flipped(I8 c) { return c & 1; }
__host__ __device__ __forceinline__ int code_rotation(I8 c) { return (c >> 1) & 3; }
__host__ __device__ __forceinline__ int code_which_down(I8 c) { return c >> 3; }
__host__ __device__ __forceinline__ I8 make_code(int flipped, int rot, int which_down) {
  return I8((which_down << 3) | (rot << 1) | (flipped ? 1 : 0));
}

}  // namespace osh

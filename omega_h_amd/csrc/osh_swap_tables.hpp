// Edge-swap triangulation tables for loops of 3 to 7 vertices.
//
// For each loop size: the unique triangles (loop-vertex triples), the
// triangulations built from them, the unique interior edges and the interior
// edges of each triangulation.  The order of the triangulations decides ties
// in the swap choice and the numbering of swap products, and the vertex order
// of each triangle and edge decides floating-point bits, so the data is
// identical to src/Omega_h_swap3d_tables.hpp:8-300.  It is encoded as strings
// (digits are loop vertices, base-36 characters are table indices) and
// expanded into __constant__ memory once.
#pragma once

#include "osh_runtime.hpp"

namespace osh {
namespace swap {

constexpr int kMaxLoop = 7;
constexpr int kMaxUniqueTris = 35;
constexpr int kMaxUniqueEdges = 26;

struct Tables {
  int mesh_sizes[8];   // triangles per triangulation
  int mesh_counts[8];  // triangulations per loop size
  int nedges[8];       // interior edges per triangulation
  int ntris[8];        // unique triangles per loop size
  int nuedges[8];      // unique interior edges per loop size
  signed char tris[8][kMaxUniqueTris][3];
  signed char meshes[8][42 * 5];
  signed char uedges[8][kMaxUniqueEdges][2];
  signed char medges[8][42 * 4];
};

inline Tables make_tables() {
  struct Src {
    int n;
    char const* tris;
    char const* meshes;
    char const* edges;
    char const* medges;
  };
  static Src const srcs[5] = {
      {3, "012", "0", "", ""},
      {4, "012023013123", "0123", "2013", "01"},
      {5, "012023034014134123234024013124", "012345067258369", "20301413420331", "0123045624"},
      {6, "012023034045025245234035345024235123013015145134014135124125",
          "0123045601780369048a23bcbdef78bc3bfg8bdh6dei36gi56dj8adj",
          "203040252435044253033115144152", "01203401506703892aabc95a6ad8ab7bc67c4eb8eb"},
      {7,
          "012023034045056036356345046456035346024234026256245025246235236013123"
          "014134016156145015146135136124125126",
          "0123401567012890147a0159b034cd0defg089cd04dgh09dei07efj047hj067ek09b"
          "ek234lm567lm289lm47alm59blm34mnomopqr89mno4mors9mopt7mpqu47msu67mpv9"
          "bmpv34dnwdpqrw89dnw4drsw9dptwdgpqx4dgsxdfgpy9dipy7jpqx47jsx7fjpy67kp"
          "y9bkpy",
          "2030405036354605536404244226255263031331144116155162",
          "012301450126017801490a3b0cde0a6b07cf09cd08de078e05gd09gdh23ih45ih26i"
          "78hih49ia3jkjlmna6jk7jlo9jlm8jmn78jn5gjm9gjma3ckcmnka6ck7cko9cmkcfmn"
          "7cfncepm9cpm8fmn78fn8epm5gpm9gpm"},
  };
  auto b36 = [](char ch) { return (ch <= '9') ? (ch - '0') : (ch - 'a' + 10); };
  Tables t;
  std::memset(&t, 0, sizeof(t));
  int const sizes[8] = {0, 0, 0, 1, 2, 3, 4, 5};
  int const counts[8] = {0, 0, 0, 1, 2, 5, 14, 42};
  int const ne[8] = {0, 0, 0, 0, 1, 2, 3, 4};
  for (int i = 0; i < 8; ++i) {
    t.mesh_sizes[i] = sizes[i];
    t.mesh_counts[i] = counts[i];
    t.nedges[i] = ne[i];
  }
  for (auto const& s : srcs) {
    int n = s.n;
    t.ntris[n] = int(std::strlen(s.tris) / 3);
    t.nuedges[n] = int(std::strlen(s.edges) / 2);
    for (size_t i = 0; i < std::strlen(s.tris); ++i) t.tris[n][i / 3][i % 3] = static_cast<signed char>(s.tris[i] - '0');
    for (size_t i = 0; i < std::strlen(s.meshes); ++i) t.meshes[n][i] = static_cast<signed char>(b36(s.meshes[i]));
    for (size_t i = 0; i < std::strlen(s.edges); ++i) t.uedges[n][i / 2][i % 2] = static_cast<signed char>(s.edges[i] - '0');
    for (size_t i = 0; i < std::strlen(s.medges); ++i) t.medges[n][i] = static_cast<signed char>(b36(s.medges[i]));
  }
  return t;
}

}  // namespace swap
}  // namespace osh

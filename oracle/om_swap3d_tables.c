/* Oracle (test infrastructure only): edge-swap triangulation tables.
 *
 * For a loop of n vertices around an edge (3 <= n <= 7) these list the
 * unique triangles, the triangulations ("meshes") built from them, and the
 * interior edges of each triangulation.  The order of the triangulations
 * decides ties in the swap choice and the numbering of swap products, and
 * the vertex order of each triangle and edge decides floating-point bits, so
 * the data is identical to src/Omega_h_swap3d_tables.hpp:8-300; it is
 * encoded here as strings (digits are loop vertices, base-36 characters are
 * table indices). */
#include <stdlib.h>
#include <string.h>

#include "om_internal.h"

int const om_swap_mesh_sizes[8] = {0, 0, 0, 1, 2, 3, 4, 5};
int const om_swap_mesh_counts[8] = {0, 0, 0, 1, 2, 5, 14, 42};
int const om_swap_nedges[8] = {0, 0, 0, 0, 1, 2, 3, 4};

typedef struct {
  int n;
  char const* tris;   /* 3 digits per unique triangle */
  char const* meshes; /* (n-2) base-36 indices per triangulation */
  char const* edges;  /* 2 digits per unique interior edge */
  char const* medges; /* (n-3) base-36 indices per triangulation */
} table_src;

static table_src const srcs[5] = {
    {3, "012", "0", "", ""},
    {4, "012023013123", "0123", "2013", "01"},
    {5, "012023034014134123234024013124", "012345067258369",
        "20301413420331", "0123045624"},
    {6, "012023034045025245234035345024235123013015145134014135124125",
        "0123045601780369048a23bcbdef78bc3bfg8bdh6dei36gi56dj8adj",
        "203040252435044253033115144152",
        "01203401506703892aabc95a6ad8ab7bc67c4eb8eb"},
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

static int tris[8][OM_MAX_UNIQUE_TRIS][3];
static int meshes[8][42 * 5];
static int uedges[8][OM_MAX_UNIQUE_EDGES][2];
static int medges[8][42 * 4];
static int ready = 0;

static int b36(char c) { return (c <= '9') ? (c - '0') : (c - 'a' + 10); }

static void init_tables(void) {
  if (ready) return;
  for (int s = 0; s < 5; ++s) {
    table_src const* t = &srcs[s];
    int n = t->n;
    for (size_t i = 0; i < strlen(t->tris); ++i) tris[n][i / 3][i % 3] = t->tris[i] - '0';
    for (size_t i = 0; i < strlen(t->meshes); ++i) meshes[n][i] = b36(t->meshes[i]);
    for (size_t i = 0; i < strlen(t->edges); ++i) uedges[n][i / 2][i % 2] = t->edges[i] - '0';
    for (size_t i = 0; i < strlen(t->medges); ++i) medges[n][i] = b36(t->medges[i]);
  }
  ready = 1;
}

int const* om_swap_triangle(int loop_size, int uniq_tri) {
  init_tables();
  return tris[loop_size][uniq_tri];
}

int om_swap_mesh_tri(int loop_size, int mesh, int mesh_tri) {
  init_tables();
  return meshes[loop_size][mesh * om_swap_mesh_sizes[loop_size] + mesh_tri];
}

int const* om_swap_unique_edge(int loop_size, int uniq_edge) {
  init_tables();
  return uedges[loop_size][uniq_edge];
}

int om_swap_mesh_edge(int loop_size, int mesh, int mesh_edge) {
  init_tables();
  return medges[loop_size][mesh * om_swap_nedges[loop_size] + mesh_edge];
}

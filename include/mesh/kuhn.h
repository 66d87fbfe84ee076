/*
 * Synthetic structured meshes: the unit cube split into n^3 cubes, each cut
 * into 6 Kuhn tetrahedra along its main diagonal, every tetrahedron
 * positively oriented, with optional uniform jitter of the interior vertices.
 *
 * No counterpart in the reference (it only reads .msh files); these produce
 * the synthetic benchmark meshes either as a gmsh v2.2 ASCII file accepted by
 * mesh_read_from_msh22_file or directly as a read (not yet built) gdn_mesh.
 *
 * Vertex (i,j,k), 0 <= i,j,k <= n, has id i + (n+1) (j + (n+1) k) and
 * coordinates (i,j,k)/n. Interior vertices are moved by jitter*h*(2u-1) per
 * coordinate, u uniform in [0,1) from SplitMix64(seed), drawn x,y,z per
 * interior vertex in increasing vertex id. Cube (i,j,k) (x fastest) yields 6
 * tetrahedra, one per permutation of the axes in lexicographic order
 * (xyz, xzy, yxz, yzx, zxy, zyx); odd permutations swap their last two
 * vertices to keep det > 0.
 */
#ifndef MESH_KUHN_H
#define MESH_KUHN_H

#include <stdint.h>
#include <mesh/mesh.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Fill a fresh gdn_mesh (vertices + tetrahedra) as if it had just been read;
 * call mesh_build_connectivity afterwards. */
void mesh_kuhn_cube(gdn_mesh *m, int n, double jitter, uint64_t seed);

/* Write the same mesh as a gmsh v2.2 ASCII file; returns 0 on success. */
int mesh_kuhn_cube_write_msh22(const char *filename, int n, double jitter,
			       uint64_t seed);

/* Write the tetrahedra of a read gdn_mesh (elem2node still allocated) as a
 * gmsh v2.2 ASCII file with round-trip exact coordinates; returns 0 on
 * success. */
int mesh_write_msh22_file(const gdn_mesh *m, const char *filename);

/* SplitMix64 step and the derived double in [0,1): the project's PRNG. */
uint64_t gdn_splitmix64(uint64_t *state);
double gdn_uniform01(uint64_t *state);

#ifdef __cplusplus
}
#endif
#endif

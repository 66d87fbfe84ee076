/*
 * Tetrahedral mesh: gmsh v2.2 reader, connectivity, element geometry.
 *
 * Replaces the reference interface src/mesh/mesh.h:15-93. The struct layout
 * is field-for-field the reference's gdn_mesh (src/mesh/mesh.h:15-45) so that
 * callers compiled against either header agree on the ABI.
 *
 * Array conventions (identical to the reference, src/mesh/mesh.c:811-1175):
 *  - face2elem[4 f + {0,1,2,3}] = {ieL, ifL, ieR, ifR}; ieR = ifR = -1 on a
 *    boundary face; faces are enumerated element by element and a face is
 *    owned by the element whose neighbour id is smaller (L = higher id);
 *  - fnode2enode[12 f + a] = local T10 node of L carrying face node a,
 *    fnode2enode[12 f + 6 + a] = matching local node of R (a on a boundary);
 *  - elem2qnode[10 e + i] = global T10 node; then 6 entries per boundary face;
 *  - node[3 k + d] = coordinates of global T10 node k.
 */
#ifndef MESH_H
#define MESH_H

#include <stdbool.h>
#include <gdon3d.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gdn_mesh {
	char file_name[1024];
	int nbelems;
	int nbnodes;
	int nbqnodes_in;
	int nbqnodes_out;
	int nbfaces;
	int nboundaryfaces;
	int *boundaryface;
	int nmacrointerfaces;
	int *macrointerface;
	bool is_read;
	bool is_built;
	bool has_physical_groups;
	bool physical_group_is_alloc;
	bool phys_group_id_is_alloc;
	int nb_phys_groups;
	int *phys_group_id;
	int *physical_group;
	int *elem2node;
	int *elem2qnode;
	int *elem2elem;
	int *face2elem;
	int *fnode2enode;
	int max_node2elem;
	int *node2elem;
	gdn_real *node;
	gdn_real *qnode;
	gdn_real xmin[3];
	gdn_real xmax[3];
} gdn_mesh;

/* src/mesh/mesh.h:47-49 */
int mesh_get_nb_elem(const gdn_mesh *msh);
int mesh_get_nb_face(const gdn_mesh *msh);
int mesh_get_nb_loc_qnode(const gdn_mesh *msh);

/* src/mesh/mesh.h:51-52: ipg = 10 id_elem + id_loc_node */
int mesh_get_gauss_point_id(const gdn_mesh *msh, const int id_elem,
			    const int id_loc_node);

/* src/mesh/mesh.h:54-61: reference-element tables */
gdn_real mesh_get_mass_ref_ij(const gdn_mesh *msh, const int id_loc_node_i,
			      const int id_loc_node_j);
gdn_real mesh_get_inv_mass_ref_ij(const gdn_mesh *msh, const int id_loc_node_i,
				  const int id_loc_node_j);
gdn_real mesh_get_internal_ref_ijk(const gdn_mesh *msh, const int id_loc_node_i,
				   const int id_loc_node_j, const int dir_k);

/* src/mesh/mesh.h:63-67 */
int mesh_get_ipg_elem_left_from_face(const gdn_mesh *msh, const int id_face,
				     const int id_loc_node_on_face);
int mesh_get_ipg_elem_right_from_face(const gdn_mesh *msh, const int id_face,
				      const int id_loc_node_on_face);

/* src/mesh/mesh.h:69 */
bool mesh_get_bool_is_boundary_face(const gdn_mesh *msh, const int id_face);

/* src/mesh/mesh.h:71-78: affine geometry of element id_elem */
void mesh_get_elem_node(const gdn_mesh *msh, const int id_elem,
			gdn_real (*t4_node)[3]);
void mesh_get_elem_det(const gdn_mesh *msh, const int id_elem, gdn_real *det);
void mesh_get_elem_codtau(const gdn_mesh *msh, const int id_elem,
			  gdn_real codtau[3][3]);

/* src/mesh/mesh.h:80. Scaled outward normal of the L element of a face. Like
 * the reference (src/mesh/mesh.c:104-127) the sqrt(3) factor is applied when
 * the GLOBAL face id equals 3; the operator itself uses the local face. */
void mesh_get_face_vn(const gdn_mesh *msh, const int id_face, gdn_real *vn);

/* src/mesh/mesh.h:82: min over elements of volume / surface */
gdn_real mesh_get_t4_hmin(const gdn_mesh *msh);

/* src/mesh/mesh.h:84-93 */
void gdn_gdn_mesh_free(gdn_mesh *m);
int gdn_gdn_mesh_is_empty(gdn_mesh *m);
void mesh_read_from_msh22_file(gdn_mesh *m, const char *filename);
void mesh_build_connectivity(gdn_mesh *m);
void gdn_gdn_mesh_print(gdn_mesh *m);

#ifdef __cplusplus
}
#endif
#endif

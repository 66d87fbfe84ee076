/*
 * P2 (10-node) reference tetrahedron. Replaces src/mesh/t10.h:9-86.
 *
 * Node numbering: vertices 0..3, edge midpoints 4:(0,1) 5:(1,2) 6:(0,2)
 * 7:(0,3) 8:(1,3) 9:(2,3). Face f carries nodes
 * F0 = (0,3,2|7,9,6), F1 = (0,1,3|4,8,7), F2 = (0,2,1|6,5,4), F3 = (1,2,3|5,9,8)
 * ordered (v0, v1, v2, e01, e12, e20).
 */
#ifndef T10_H
#define T10_H

#include <gdon3d.h>

#ifdef __cplusplus
extern "C" {
#endif

#define T10_ELEM_CODE 10
#define T10_PHY_DIM 3
#define T10_FACE_PER_ELEM 4
#define T10_NODE_PER_ELEM 10
#define T10_NODE_PER_FACE 6

#define ECODE 4 /* gmsh element type of the T4 tetrahedron */
#define NFA 4 /* faces per element */
#define NNO 4 /* T4 nodes per element */
#define NVE 3 /* T4 nodes per face */
#define E2F 4 /* face2elem stride */
#define NQN 10 /* T10 nodes per element */
#define NQV 6 /* T10 nodes per face */

/* src/mesh/t10.h:72-85 */
void t10_get_ref_phi(gdn_real x, gdn_real y, gdn_real z, gdn_real *phi);
void t10_get_ref_grad_phi(gdn_real x, gdn_real y, gdn_real z,
			  gdn_real *grad_phi);
gdn_real t10_get_ref_normal_entry(const int i, const int j);
gdn_real t10_get_ref_node_entry(const int i, const int j);
void t10_get_ref_node_coord(const int i, gdn_real coord[3]);
int t10_get_ref_face2node_entry(const int i, const int j);
gdn_real t10_get_wpg_ref_entry(const int i);
gdn_real t10_get_mass_ref_entry(const int i, const int j);
gdn_real t10_get_inv_mass_ref_entry(const int i, const int j);
gdn_real t10_get_internal_ref_entry(const int i, const int j, const int id_dir);
/* int over reference face id_face of phi_i phi_j (src/mesh/t10.h:85) */
gdn_real t10_get_flux_ref_entry(const int i, const int j, const int id_face);
gdn_real t10_get_face_scale_entry(const int i, const int j);

/* P1 reference tetrahedron (src/mesh/t4.h): face -> vertex map and weights */
int t4_get_ref_face2node_entry(const int i, const int j);
gdn_real t4_get_ref_normal_entry(const int i, const int j);
gdn_real t4_get_wpg_ref_entry(const int i);

#ifdef __cplusplus
}
#endif
#endif

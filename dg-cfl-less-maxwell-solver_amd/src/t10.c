/*
 * P2 reference tetrahedron: basis, nodes, normals and integral tables.
 * Restates the reference element of src/mesh/t10.c:23-1436 and
 * src/mesh/t4.c; the integral tables come from gen_t10_tables.py.
 */
#include <mesh/t10.h>
#include "t10_tables.h"

/* Nodes of the reference element (src/mesh/t10.c:23-27). */
static const gdn_real ref_node[NQN][3] = {
	{ 0, 0, 0 },	 { 1, 0, 0 },	  { 0, 1, 0 },	 { 0, 0, 1 },
	{ 0.5, 0, 0 },	 { 0.5, 0.5, 0 }, { 0, 0.5, 0 }, { 0, 0, 0.5 },
	{ 0.5, 0, 0.5 }, { 0, 0.5, 0.5 }
};

/* Outward unit normals of faces 0..3. The last one is the binary64 value of
 * the decimal literal 0.5773502691896258420811705 used by the reference
 * (src/mesh/t10.c:29-35, src/mesh/t4.c); its product with sqrt(3) decides
 * upwind signs on face 3, so the same double is used. */
#define INV_SQRT3_LIT 0.5773502691896258420811705
static const gdn_real ref_normal[NFA][3] = { { -1, 0, 0 },
					     { 0, -1, 0 },
					     { 0, 0, -1 },
					     { INV_SQRT3_LIT, INV_SQRT3_LIT,
					       INV_SQRT3_LIT } };

static const int face2node[NFA][NQV] = { { 0, 3, 2, 7, 9, 6 },
					 { 0, 1, 3, 4, 8, 7 },
					 { 0, 2, 1, 6, 5, 4 },
					 { 1, 2, 3, 5, 9, 8 } };

static const gdn_real wpg[NQN] = T10_WPG_INIT;
static const gdn_real mass[NQN][NQN] = T10_MASS_INIT;
static const gdn_real inv_mass[NQN][NQN] = T10_INV_MASS_INIT;
static const gdn_real internal[NQN][NQN][3] = T10_INTERNAL_INIT;
static const gdn_real face_scal[NQV][NQV] = T10_FACE_SCAL_INIT;
static const gdn_real flux_ref[NQN][NQN][NFA] = T10_FLUX_REF_INIT;

/* P1 element: faces as vertex triples and the nodal weights 1/24. */
static const int t4_face2node[NFA][3] = { { 0, 3, 2 },
					  { 0, 1, 3 },
					  { 0, 2, 1 },
					  { 1, 2, 3 } };

void t10_get_ref_phi(gdn_real x, gdn_real y, gdn_real z, gdn_real *phi)
{
	const gdn_real l0 = 1 - x - y - z;
	phi[0] = l0 * (2 * l0 - 1);
	phi[1] = x * (2 * x - 1);
	phi[2] = y * (2 * y - 1);
	phi[3] = z * (2 * z - 1);
	phi[4] = 4 * x * l0;
	phi[5] = 4 * x * y;
	phi[6] = 4 * y * l0;
	phi[7] = 4 * z * l0;
	phi[8] = 4 * x * z;
	phi[9] = 4 * y * z;
}

/* grad_phi[10 k + i] = d phi_i / d x_k (direction-major, as
 * src/mesh/t10.c:1351-1384) */
void t10_get_ref_grad_phi(gdn_real x, gdn_real y, gdn_real z,
			  gdn_real *grad_phi)
{
	const gdn_real l0 = 1 - x - y - z;
	const gdn_real d0 = 1 - 4 * l0; /* d/dx_k of l0 (2 l0 - 1) */
	gdn_real g[NQN][3] = {
		{ d0, d0, d0 },
		{ 4 * x - 1, 0, 0 },
		{ 0, 4 * y - 1, 0 },
		{ 0, 0, 4 * z - 1 },
		{ 4 * (l0 - x), -4 * x, -4 * x },
		{ 4 * y, 4 * x, 0 },
		{ -4 * y, 4 * (l0 - y), -4 * y },
		{ -4 * z, -4 * z, 4 * (l0 - z) },
		{ 4 * z, 0, 4 * x },
		{ 0, 4 * z, 4 * y },
	};
	for (int i = 0; i < NQN; i++)
		for (int k = 0; k < 3; k++)
			grad_phi[NQN * k + i] = g[i][k];
}

gdn_real t10_get_ref_normal_entry(const int i, const int j)
{
	return ref_normal[i][j];
}

gdn_real t10_get_ref_node_entry(const int i, const int j)
{
	return ref_node[i][j];
}

void t10_get_ref_node_coord(const int i, gdn_real coord[3])
{
	coord[0] = ref_node[i][0];
	coord[1] = ref_node[i][1];
	coord[2] = ref_node[i][2];
}

int t10_get_ref_face2node_entry(const int i, const int j)
{
	return face2node[i][j];
}

gdn_real t10_get_wpg_ref_entry(const int i)
{
	return wpg[i];
}

gdn_real t10_get_mass_ref_entry(const int i, const int j)
{
	return mass[i][j];
}

gdn_real t10_get_inv_mass_ref_entry(const int i, const int j)
{
	return inv_mass[i][j];
}

gdn_real t10_get_internal_ref_entry(const int i, const int j, const int id_dir)
{
	return internal[i][j][id_dir];
}

/* src/mesh/t10.c:1428-1431: the exact face mass of face id_face */
gdn_real t10_get_flux_ref_entry(const int i, const int j, const int id_face)
{
	return flux_ref[i][j][id_face];
}

gdn_real t10_get_face_scale_entry(const int i, const int j)
{
	return face_scal[i][j];
}

int t4_get_ref_face2node_entry(const int i, const int j)
{
	return t4_face2node[i][j];
}

gdn_real t4_get_ref_normal_entry(const int i, const int j)
{
	return ref_normal[i][j];
}

gdn_real t4_get_wpg_ref_entry(const int i)
{
	(void)i;
	return 1.0 / 24.0;
}

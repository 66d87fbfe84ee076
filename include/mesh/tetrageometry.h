/*
 * Affine (P1) geometry of a tetrahedron. Replaces src/mesh/tetrageometry.h.
 *
 * gdn_tetra_ref2phy (src/mesh/tetrageometry.c:166-260): for the 4 vertices
 * physnode, optionally computes the image xphy of xref, the Jacobian
 * dtau[ii][jj] = d x_ii / d xref_jj, its cofactor matrix codtau (= det dtau^-T),
 * the physical gradient dphi = codtau dphiref and, for local face ifa >= 0, the
 * scaled outward normal vnds = codtau n_ref[ifa]. Any output may be NULL.
 */
#ifndef TETRAGEOMETRY_H
#define TETRAGEOMETRY_H

#include <gdon3d.h>

#ifdef __cplusplus
extern "C" {
#endif

gdn_real geometry_dot_product(const gdn_real a[3], const gdn_real b[3]);

void gdn_tetra_ref2phy(gdn_real physnode[4][3], gdn_real xref[3],
		       gdn_real dphiref[3], int ifa, gdn_real xphy[3],
		       gdn_real dtau[3][3], gdn_real codtau[3][3],
		       gdn_real dphi[3], gdn_real vnds[3]);

/* Inverse of the affine map: xref with gdn_tetra_ref2phy(xref) = xphy
 * (src/mesh/tetrageometry.c:262-292) */
void gdn_tetra_phy2ref(gdn_real physnode[4][3], gdn_real xphy[3], gdn_real xref[3]);

/* The same along the homotopy from the image of (1/2, 1/2, 1/2) to xphy in
 * 100 steps, stopping at the first step whose reference point leaves
 * [-eps, 1 + eps]^3 (src/mesh/tetrageometry.c:294-348) */
void gdn_tetra_robust_phy2ref(gdn_real physnode[4][3], gdn_real xphy[3], gdn_real xref[3]);

/* Heron-formula surface and edge-length volume (src/mesh/tetrageometry.c:122-164) */
void geometry_get_tetra_area_and_vol(gdn_real (*node)[3], gdn_real *surf,
				     gdn_real *vol);

#ifdef __cplusplus
}
#endif
#endif

/*
 * Affine tetrahedron geometry. Follows src/mesh/tetrageometry.c:122-260:
 * same quantities and the same order of floating-point operations, so that
 * codtau, det and the face normals are bit-identical to the reference's
 * (upwind signs are decided on them). Built with -ffp-contract=off.
 */
#include <math.h>
#include <stddef.h>
#include <mesh/tetrageometry.h>
#include <mesh/t10.h>

gdn_real geometry_dot_product(const gdn_real a[3], const gdn_real b[3])
{
	return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static gdn_real dist(const gdn_real a[3], const gdn_real b[3])
{
	gdn_real d0 = b[0] - a[0], d1 = b[1] - a[1], d2 = b[2] - a[2];
	return sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}

void geometry_get_tetra_area_and_vol(gdn_real (*node)[3], gdn_real *surf,
				     gdn_real *vol)
{
	/* edge lengths: 30 31 32 01 12 02 */
	const gdn_real a[6] = { dist(node[3], node[0]), dist(node[3], node[1]),
				dist(node[3], node[2]), dist(node[0], node[1]),
				dist(node[1], node[2]), dist(node[0], node[2]) };
	/* Heron on the four faces */
	const gdn_real s1 = (a[0] + a[1] + a[3]) / 2;
	const gdn_real s2 = (a[1] + a[2] + a[4]) / 2;
	const gdn_real s3 = (a[2] + a[5] + a[0]) / 2;
	const gdn_real s4 = (a[3] + a[4] + a[5]) / 2;
	*surf = sqrt(s1 * (s1 - a[0]) * (s1 - a[1]) * (s1 - a[3])) +
		sqrt(s2 * (s2 - a[1]) * (s2 - a[2]) * (s2 - a[4])) +
		sqrt(s3 * (s3 - a[2]) * (s3 - a[5]) * (s3 - a[0])) +
		sqrt(s4 * (s4 - a[3]) * (s4 - a[4]) * (s4 - a[5]));

	/* volume from the six edge lengths (Cayley-Menger form) */
	const gdn_real q1 = a[0] * a[0], q2 = a[1] * a[1], q3 = a[2] * a[2];
	const gdn_real q4 = a[3] * a[3], q5 = a[4] * a[4], q6 = a[5] * a[5];
	const gdn_real v2 = q1 * q5 * (q2 + q3 + q4 + q6 - q1 - q5) +
			    q2 * q6 * (q1 + q3 + q4 + q5 - q2 - q6) +
			    q3 * q4 * (q1 + q2 + q5 + q6 - q3 - q4) -
			    q1 * q2 * q4 - q2 * q3 * q5 - q1 * q3 * q6 -
			    q4 * q5 * q6;
	*vol = sqrt(v2 / 144.);
}

void gdn_tetra_ref2phy(gdn_real physnode[4][3], gdn_real xref[3],
		       gdn_real dphiref[3], int ifa, gdn_real xphy[3],
		       gdn_real dtau[3][3], gdn_real codtau[3][3],
		       gdn_real dphi[3], gdn_real vnds[3])
{
	/* P1 shape functions: gradient columns 0..2, value in column 3 */
	const gdn_real x = xref ? xref[0] : 0, y = xref ? xref[1] : 0,
		       z = xref ? xref[2] : 0;
	const gdn_real shape[4][4] = { { -1, -1, -1, 1 - x - y - z },
				       { 1, 0, 0, x },
				       { 0, 1, 0, y },
				       { 0, 0, 1, z } };

	if (xphy) {
		for (int d = 0; d < 3; d++) {
			xphy[d] = 0;
			for (int i = 0; i < 4; i++)
				xphy[d] += shape[i][3] * physnode[i][d];
		}
	}
	if (dtau) {
		for (int r = 0; r < 3; r++) {
			for (int c = 0; c < 3; c++)
				dtau[r][c] = 0;
			for (int i = 0; i < 4; i++)
				for (int c = 0; c < 3; c++)
					dtau[r][c] += physnode[i][r] * shape[i][c];
		}
	}
	if (codtau) {
		codtau[0][0] = dtau[1][1] * dtau[2][2] - dtau[1][2] * dtau[2][1];
		codtau[0][1] = -dtau[1][0] * dtau[2][2] + dtau[1][2] * dtau[2][0];
		codtau[0][2] = dtau[1][0] * dtau[2][1] - dtau[1][1] * dtau[2][0];
		codtau[1][0] = -dtau[0][1] * dtau[2][2] + dtau[0][2] * dtau[2][1];
		codtau[1][1] = dtau[0][0] * dtau[2][2] - dtau[0][2] * dtau[2][0];
		codtau[1][2] = -dtau[0][0] * dtau[2][1] + dtau[0][1] * dtau[2][0];
		codtau[2][0] = dtau[0][1] * dtau[1][2] - dtau[0][2] * dtau[1][1];
		codtau[2][1] = -dtau[0][0] * dtau[1][2] + dtau[0][2] * dtau[1][0];
		codtau[2][2] = dtau[0][0] * dtau[1][1] - dtau[0][1] * dtau[1][0];
	}
	if (dphi) {
		for (int r = 0; r < 3; r++) {
			dphi[r] = 0;
			for (int c = 0; c < 3; c++)
				dphi[r] += codtau[r][c] * dphiref[c];
		}
	}
	if (vnds && ifa >= 0) {
		for (int r = 0; r < 3; r++) {
			vnds[r] = 0.0;
			for (int c = 0; c < 3; c++)
				vnds[r] += codtau[r][c] *
					   t10_get_ref_normal_entry(ifa, c);
		}
	}
}

/*
 * Inverse of the affine map (src/mesh/tetrageometry.c:262-292). The
 * reference runs 20 Newton steps from xref = (1/2, 1/2, 1/2); the map being
 * affine, the first step already lands on the solution and the others only
 * re-round it. Here: the step from (1/2, 1/2, 1/2), xref = c - codtau^T
 * (x(c) - xphy) / det, then one correction step from that point.
 */
void gdn_tetra_phy2ref(gdn_real physnode[4][3], gdn_real xphy[3], gdn_real xref[3])
{
	gdn_real dtau[3][3], codtau[3][3], x[3];
	xref[0] = xref[1] = xref[2] = 0.5;
	for (int step = 0; step < 2; step++) {
		gdn_tetra_ref2phy(physnode, xref, NULL, -1, x, dtau, codtau, NULL, NULL);
		const gdn_real det = geometry_dot_product(dtau[0], codtau[0]);
		const gdn_real r[3] = { x[0] - xphy[0], x[1] - xphy[1], x[2] - xphy[2] };
		for (int i = 0; i < 3; i++)
			xref[i] -= (codtau[0][i] * r[0] + codtau[1][i] * r[1] +
				    codtau[2][i] * r[2]) / det;
	}
}

/*
 * src/mesh/tetrageometry.c:294-348: the reference follows the homotopy
 * xphy(theta) = theta xphy + (1 - theta) x(1/2, 1/2, 1/2), theta = k / 100,
 * and returns at the first theta whose reference point leaves
 * [-eps, 1 + eps]^3 (eps = DBL_EPSILON). Affine maps take that segment to the
 * segment xref(theta) = theta xref + (1 - theta) (1/2, 1/2, 1/2), so the
 * inverse is formed once and the path walked on the reference side.
 */
void gdn_tetra_robust_phy2ref(gdn_real physnode[4][3], gdn_real xphy[3], gdn_real xref[3])
{
	gdn_real xr[3];
	gdn_tetra_phy2ref(physnode, xphy, xr);
	const gdn_real dtheta = 1. / 100;
	for (int k = 0; k <= 100; k++) {
		const gdn_real theta = k * dtheta;
		int inside = 1;
		for (int i = 0; i < 3; i++) {
			xref[i] = theta * xr[i] + (1 - theta) * 0.5;
			inside &= xref[i] >= _GDN_XREF_LOWBOUND && xref[i] <= _GDN_XREF_UPBOUND;
		}
		if (!inside)
			return;
	}
}

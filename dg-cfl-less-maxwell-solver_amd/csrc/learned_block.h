/*
 * The diagonal block of the transport solve rebuilt from its seven numbers,
 * and the refinement of a predicted inverse, shared by the set-up check
 * (learned.hip) and the transport sweep (sweep.hip) so that both compute the
 * same values bit for bit.
 *
 * Item (v, e): D = I + theta dt/det Minv F_ee(v) (src/dg/relaxation.c:64-87,
 * src/simulation/simulation.c:298-308, 531-562) is
 *   D = I - sum_k x_k (Minv C_k) + sum_f x_{3+f} (Minv L_f S),
 * x_k = theta dt/det (v . gamma_k) with gamma_k the k-th column of codtau,
 * x_{3+f} = theta dt/det (v . vn_f) on the outflow faces (0 elsewhere),
 * Minv C_k the inverse mass times the volume operator of direction k and
 * Minv L_f S the inverse mass times the lift of the face mass.
 * The inverse is refined from a prediction X0 by Newton-Schulz,
 * R = I - D X, X <- X + X R: the residual squares at every step.
 */
#ifndef GDH_LEARNED_BLOCK_H
#define GDH_LEARNED_BLOCK_H

#include "physics.h"

namespace gdh {

__constant__ static const double c_lb_minv_int[3][10][10] = T10_MINV_INTERNAL_INIT;
__constant__ static const double c_lb_minv_lift[4][10][6] = T10_MINV_LIFT_INIT;
/* position b of local node j on face f (face-node order of c_face2node),
 * or -1 */
__constant__ static const signed char c_lb_slot[4][10] = {
	{ 0, -1, 2, 1, -1, -1, 5, 3, -1, 4 },
	{ 0, 1, -1, 2, 3, -1, -1, 5, 4, -1 },
	{ 0, 2, 1, -1, 5, 4, 3, -1, -1, -1 },
	{ -1, 0, 1, 2, -1, 3, -1, -1, 5, 4 },
};

/* Newton-Schulz steps applied to a prediction: S steps give an inverse
 * whose residual is the square of the last R; each item takes the fewest
 * steps (2 .. LB_MAX_STEPS) whose last R is at most the set-up tolerance
 * (1e-8: the result then lies within rounding of the exact inverse), found
 * at set-up, kept in the descriptor flags (bits 30-31: S - 2) */
#define LB_MAX_STEPS 5

/* entry (i, j) of D from the features x[0..6] */
__device__ inline double lb_entry(const double *x, int i, int j)
{
	double s = i == j ? 1.0 : 0.0;
	s -= x[0] * c_lb_minv_int[0][i][j] + x[1] * c_lb_minv_int[1][i][j] +
	     x[2] * c_lb_minv_int[2][i][j];
#pragma unroll
	for (int f = 0; f < 4; f++) {
		const int b = c_lb_slot[f][j];
		if (b >= 0)
			s += x[3 + f] * c_lb_minv_lift[f][i][b];
	}
	return s;
}

/* a stored prediction entry: (X0 - I)_ij in bfloat16 (the upper half of a
 * binary32) */
__device__ inline double lb_pred(unsigned short h)
{
	return (double)__uint_as_float((unsigned)h << 16);
}

} // namespace gdh

#endif

/*
 * Device physics: Maxwell fluxes, exact solutions, kinetic equilibria, and
 * the P2 reference element tables as device constants.
 *
 * Each function is the device twin of a host callback of the library
 * (src/maxwell.c, src/d3q4.c, src/d3q6.c) and follows the formulas of the
 * reference's src/models/maxwell.c:72-223, src/models/d3q4.c:82-301 and
 * src/models/d3q6.c:85-356.
 */
#ifndef GDH_PHYSICS_H
#define GDH_PHYSICS_H

#include <hip/hip_runtime.h>
#include <gdn_hip.h>
#include "../src/t10_tables.h"

namespace gdh {

/* ---------------------------------------------------------- T10 tables */
__constant__ static const double c_wpg[10] = T10_WPG_INIT;
__constant__ static const double c_inv_mass[10][10] = T10_INV_MASS_INIT;
__constant__ static const double c_mass[10][10] = T10_MASS_INIT;
__constant__ static const double c_internal[10][10][3] = T10_INTERNAL_INIT;
__constant__ static const double c_face_scal[6][6] = T10_FACE_SCAL_INIT;

/* face -> local nodes, (v0, v1, v2, e01, e12, e20) */
__constant__ static const int c_face2node[4][6] = { { 0, 3, 2, 7, 9, 6 },
						    { 0, 1, 3, 4, 8, 7 },
						    { 0, 2, 1, 6, 5, 4 },
						    { 1, 2, 3, 5, 9, 8 } };

/* the same map as a host/device constexpr for unrolled loops */
__host__ __device__ constexpr int face_node(int f, int a)
{
	return f == 0 ? (a == 0 ? 0 : a == 1 ? 3 : a == 2 ? 2 : a == 3 ? 7 : a == 4 ? 9 : 6) :
	       f == 1 ? (a == 0 ? 0 : a == 1 ? 1 : a == 2 ? 3 : a == 3 ? 4 : a == 4 ? 8 : 7) :
	       f == 2 ? (a == 0 ? 0 : a == 1 ? 2 : a == 2 ? 1 : a == 3 ? 6 : a == 4 ? 5 : 4) :
			(a == 0 ? 1 : a == 1 ? 2 : a == 2 ? 3 : a == 3 ? 5 : a == 4 ? 9 : 8);
}

/* face-local index of local node i on face f, -1 if i is not on f */
__host__ __device__ constexpr int node_on_face(int f, int i)
{
	return face_node(f, 0) == i ? 0 : face_node(f, 1) == i ? 1 :
	       face_node(f, 2) == i ? 2 : face_node(f, 3) == i ? 3 :
	       face_node(f, 4) == i ? 4 : face_node(f, 5) == i ? 5 : -1;
}

/* reference normals; face 3 uses the binary64 value of the reference's
 * literal 0.5773502691896258420811705, and the face-3 normal is scaled by
 * sqrt(3) (src/simulation/simulation.c:377-379) */
#define GDH_INV_SQRT3 0.5773502691896258420811705
#define GDH_SQRT3 1.7320508075688772

/* Scaled outward normal of local face f: cr * codtau * n_ref[f], in the
 * reference's order of operations. */
__device__ inline void face_normal(const double ct[9], int f, double vn[3])
{
	if (f < 3) {
		/* n_ref = -e_f: the products with 0 vanish exactly */
		for (int r = 0; r < 3; r++)
			vn[r] = -ct[3 * r + f];
	} else {
#pragma clang fp contract(off)
		for (int r = 0; r < 3; r++) {
			double s = 0.0;
			s = s + ct[3 * r + 0] * GDH_INV_SQRT3;
			s = s + ct[3 * r + 1] * GDH_INV_SQRT3;
			s = s + ct[3 * r + 2] * GDH_INV_SQRT3;
			vn[r] = s * GDH_SQRT3;
		}
	}
}

/* ------------------------------------------------------------- fluxes */

/* Upwind Maxwell flux (src/models/maxwell.c:152-192) with precomputed
 * overr = 1/(|n| + 1e-16); T = double (the plain form) or float (the FP32
 * form), same operation order. */
template <typename T>
__device__ inline void maxwell_upwind(const T wL[6], const T wR[6], const T n[3],
				      T overr, T fl[6])
{
	T avg[6], jmp[6];
#pragma unroll
	for (int q = 0; q < 6; q++) {
		avg[q] = T(0.5) * (wR[q] + wL[q]);
		jmp[q] = T(0.5) * (wR[q] - wL[q]);
	}
	const T nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
#pragma unroll
	for (int blk = 0; blk < 2; blk++) {
		const T *u = jmp + 3 * blk;
		const T *c = avg + 3 * (1 - blk);
		const T ndu = n[0] * u[0] + n[1] * u[1] + n[2] * u[2];
		const T cr0 = n[1] * c[2] - n[2] * c[1];
		const T cr1 = n[2] * c[0] - n[0] * c[2];
		const T cr2 = n[0] * c[1] - n[1] * c[0];
		const T sg = blk == 0 ? T(-1) : T(1);
		fl[3 * blk + 0] = sg * cr0 + overr * (n[0] * ndu - nn * u[0]);
		fl[3 * blk + 1] = sg * cr1 + overr * (n[1] * ndu - nn * u[1]);
		fl[3 * blk + 2] = sg * cr2 + overr * (n[2] * ndu - nn * u[2]);
	}
}

/* (-n x H_L, 0) (src/models/maxwell.c:205-223) */
template <typename T>
__device__ inline void maxwell_metal(const T wL[6], const T n[3], T fl[6])
{
	fl[0] = n[2] * wL[4] - n[1] * wL[5];
	fl[1] = n[0] * wL[5] - n[2] * wL[3];
	fl[2] = n[1] * wL[3] - n[0] * wL[4];
	fl[3] = fl[4] = fl[5] = T(0);
}

/* ----------------------------------------------------- exact solutions */

#define GDH_PI10 0.3141592654e1

__device__ inline double st_g(double x)
{
	return sin(2. * M_PI * (x - 0.5));
}
__device__ inline double st_f(double y, double z)
{
	return exp(-2 * M_PI * (y - 0.5)) * sin(2 * M_PI * (z - 0.5)) +
	       exp(-2 * M_PI * (z - 0.5)) * sin(2 * M_PI * (y - 0.5));
}
__device__ inline double st_df1(double y, double z)
{
	return -2. * GDH_PI10 * exp(-2. * GDH_PI10 * (y - 0.5)) *
		       sin(2. * GDH_PI10 * (z - 0.5)) +
	       2. * exp(-2. * GDH_PI10 * (z - 0.5)) * GDH_PI10 *
		       cos(2. * GDH_PI10 * (y - 0.5));
}
__device__ inline double st_df2(double y, double z)
{
	return 2. * exp(-2. * GDH_PI10 * (y - 0.5)) * GDH_PI10 *
		       cos(2. * GDH_PI10 * (z - 0.5)) -
	       2. * exp(-2. * GDH_PI10 * (z - 0.5)) * GDH_PI10 *
		       sin(2. * GDH_PI10 * (y - 0.5));
}

struct ImposedPar {
	int kind;
	double p[GDH_IMPOSED_NPAR];
};

/* Exact macro state at x, t for the recognised kinds (not GDH_IMPOSED_HOST) */
__device__ inline void imposed(const ImposedPar &ip, const double x[3], double t,
			       double w[6])
{
	switch (ip.kind) {
	case GDH_IMPOSED_MACRO_COS: {
		const double s = cos(5 * M_PI * (x[0] - t));
		w[0] = 0; w[1] = 0; w[2] = s; w[3] = 0; w[4] = -s; w[5] = 0;
		break;
	}
	case GDH_IMPOSED_CAVITY_TM: {
		const double r2 = sqrt(2.);
		const double sx = sin(GDH_PI10 * x[0]), cx = cos(GDH_PI10 * x[0]);
		const double sy = sin(GDH_PI10 * x[1]), cy = cos(GDH_PI10 * x[1]);
		const double st = sin(r2 * GDH_PI10 * t);
		w[0] = 0;
		w[1] = 0;
		w[2] = sx * sy * cos(r2 * GDH_PI10 * t);
		w[3] = -st * r2 * sx * cy / 2.;
		w[4] = st * r2 * cx * sy / 2.;
		w[5] = 0;
		break;
	}
	case GDH_IMPOSED_STATIONARY:
		w[0] = st_g(x[0]) + st_f(x[1], x[2]);
		w[1] = st_g(x[1]) + st_f(x[2], x[0]);
		w[2] = st_g(x[2]) + st_f(x[0], x[1]);
		w[3] = (st_df1(x[2], x[0]) - st_df2(x[0], x[1])) * t;
		w[4] = (st_df1(x[0], x[1]) - st_df2(x[1], x[2])) * t;
		w[5] = (st_df1(x[1], x[2]) - st_df2(x[2], x[0])) * t;
		break;
	case GDH_IMPOSED_PLANE_WAVE: {
		const double *k = ip.p, *e = ip.p + 3, *h = ip.p + 6;
		const double s = sin(k[0] * x[0] + k[1] * x[1] + k[2] * x[2] -
				     ip.p[9] * t + ip.p[10]);
		for (int d = 0; d < 3; d++) {
			w[d] = e[d] * s;
			w[3 + d] = h[d] * s;
		}
		break;
	}
	case GDH_IMPOSED_CAVITY_MODE: {
		const double *k = ip.p, *a = ip.p + 3, om = ip.p[6];
		double s[3], c[3];
		for (int d = 0; d < 3; d++) {
			s[d] = sin(k[d] * x[d]);
			c[d] = cos(k[d] * x[d]);
		}
		const double ct = cos(om * t), st = sin(om * t) / om;
		w[0] = a[0] * c[0] * s[1] * s[2] * ct;
		w[1] = a[1] * s[0] * c[1] * s[2] * ct;
		w[2] = a[2] * s[0] * s[1] * c[2] * ct;
		w[3] = -st * (k[1] * a[2] - k[2] * a[1]) * s[0] * c[1] * c[2];
		w[4] = -st * (k[2] * a[0] - k[0] * a[2]) * c[0] * s[1] * c[2];
		w[5] = -st * (k[0] * a[1] - k[1] * a[0]) * c[0] * c[1] * s[2];
		break;
	}
	case GDH_IMPOSED_GAUSSIAN_PULSE: {
		const double d0 = x[0] - ip.p[0], d1 = x[1] - ip.p[1], d2 = x[2] - ip.p[2];
		const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
		for (int q = 0; q < 6; q++)
			w[q] = 0.0;
		w[2] = ip.p[4] * exp(-r2 / (2 * ip.p[3] * ip.p[3]));
		break;
	}
	default:
		for (int q = 0; q < 6; q++)
			w[q] = 0.0;
	}
}

/* ---------------------------------------------------- kinetic equilibria */

struct KinPar {
	int kind; /* GDH_KIN_D3Q4 or GDH_KIN_D3Q6 */
	int nv;
	double vi[3 * GDH_MAX_V];
	double lambda, omega;
	/* boundary state: GDH_FR_FEQ f_eq of the imposed data, GDH_FR_PROJECTOR
	 * the projector form, with the model's projectors P = c0 I + c1 A^2 + c2 A
	 * (A = A(n / |n|)): pos (P^{+,0} or the like) and neg */
	int fr_kind;
	double proj_pos[3], proj_neg[3];
	/* permittivity: 1/eps per mesh element (null: eps = 1 everywhere); the
	 * kinetic state then carries the conserved fields (D, B) = (eps E, H),
	 * and the physical flux of the equilibrium is taken on (D / eps, B).
	 * bface: element and local face of each boundary face (ghost DoFs take
	 * the permittivity of their element). */
	const double *ieps;
	const int2 *bface;
	/* learned block inverses (null: the stored exact ones everywhere):
	 * per sweep item 50 words of two bfloat16 each, (X0 - I) row-major, and
	 * the block's 8 numbers in binary64 (7 used) */
	const unsigned *lb_p;
	const double *lb_x;
};

/* 1/eps at DoF ipg (element DoFs, or ghost DoFs through their element) */
__device__ inline double kin_ieps(const KinPar &kp, long ipg, long ne)
{
	if (!kp.ieps)
		return 1.0;
	const long e = ipg < 10 * ne ? ipg / 10 : kp.bface[(ipg - 10 * ne) / 6].x;
	return kp.ieps[e];
}

/* Physical flux q_k = A(e_k) w of Maxwell (the central part of the upwind
 * flux with n = e_k; its jump terms vanish for wL = wR). q[iw][k]. */
__device__ inline void maxwell_qM(const double w[6], double q[6][3])
{
	/* E-flux = -e_k x H, H-flux = e_k x E */
	q[0][0] = 0;	 q[0][1] = -w[5]; q[0][2] = w[4];
	q[1][0] = w[5];	 q[1][1] = 0;	  q[1][2] = -w[3];
	q[2][0] = -w[4]; q[2][1] = w[3];  q[2][2] = 0;
	q[3][0] = 0;	 q[3][1] = w[2];  q[3][2] = -w[1];
	q[4][0] = -w[2]; q[4][1] = 0;	  q[4][2] = w[0];
	q[5][0] = w[1];	 q[5][1] = -w[0]; q[5][2] = 0;
}

/* f_eq(w)[iw * nv + v] (src/models/d3q4.c:137-157,261-273,
 * src/models/d3q6.c:148-209,319-330) */
__device__ inline double feq_comp(const KinPar &kp, const double w[6],
				  const double q[6][3], int iw, int v)
{
	const double *vv = kp.vi + 3 * v;
	if (kp.kind == GDH_KIN_D3Q4) {
		const double l2 = kp.lambda * kp.lambda;
		const double qv = q[iw][0] * vv[0] + q[iw][1] * vv[1] + q[iw][2] * vv[2];
		return (w[iw] + qv / l2) / kp.nv;
	}
	/* D3Q6: v = +-lambda e_k, index k or 3 + k */
	const int k = v % 3;
	const double a = w[iw] / 6., b = q[iw][k] / (2. * kp.lambda);
	return v < 3 ? a + b : a - b;
}

/* P w = c0 w + c1 A^2 w + c2 A w with A = A(n/|n|): A w = (-m x H, m x E),
 * A^2 w = ((I - m m^T) E, (I - m m^T) H) (the callbacks
 * maxwell_get_proj_*_prod_w, src/models/maxwell.c:234-575) */
__device__ inline void maxwell_proj(const double w[6], const double n[3], const double c[3],
				    double r[6])
{
	const double nrm = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
	const double m[3] = { n[0] / nrm, n[1] / nrm, n[2] / nrm };
	const double *E = w, *H = w + 3;
	const double aw[6] = { -(m[1] * H[2] - m[2] * H[1]), -(m[2] * H[0] - m[0] * H[2]),
			       -(m[0] * H[1] - m[1] * H[0]), m[1] * E[2] - m[2] * E[1],
			       m[2] * E[0] - m[0] * E[2], m[0] * E[1] - m[1] * E[0] };
	double a2w[6];
	for (int blk = 0; blk < 2; blk++) {
		const double *u = w + 3 * blk;
		const double mu = m[0] * u[0] + m[1] * u[1] + m[2] * u[2];
		for (int d = 0; d < 3; d++)
			a2w[3 * blk + d] = u[d] - m[d] * mu;
	}
	for (int q = 0; q < 6; q++)
		r[q] = c[0] * w[q] + c[1] * a2w[q] + c[2] * aw[q];
}

/*
 * Projector boundary state (the reference's get_fR with use_feq = false,
 * src/models/d3q4.c:303-332, src/models/d3q6.c:358-379): (wL, yL) = Q(fL)
 * with y = q - qM(w) the distance of the velocity moments q to the physical
 * flux; wR = P^-(vn) wI (D3Q4: the reference keeps only that term) or
 * P^{+,0}(vn) wL + P^-(vn) wI (D3Q6); fR = Q^-1(wR, yL) = M^-1(wR, yL +
 * qM(wR)). fL[iw][v], fR[iw][v]; NV = kp.nv.
 */
template <int NV>
__device__ inline void fr_projector(const KinPar &kp, const double (&fL)[6][NV],
				    const double vn[3], const double wI[6], double (&fR)[6][NV])
{
	double wL[6], y[6][5];
	for (int iw = 0; iw < 6; iw++) {
		double s = 0, r[5] = { 0, 0, 0, 0, 0 };
		for (int v = 0; v < NV; v++) {
			const double *vv = kp.vi + 3 * v;
			const double f = fL[iw][v];
			s += f;
			r[0] += f * vv[0];
			r[1] += f * vv[1];
			r[2] += f * vv[2];
			r[3] += f * (vv[0] * vv[0] - vv[1] * vv[1]);
			r[4] += f * (vv[0] * vv[0] - vv[2] * vv[2]);
		}
		wL[iw] = s;
		for (int k = 0; k < 5; k++)
			y[iw][k] = r[k];
	}
	double qL[6][3];
	maxwell_qM(wL, qL);
	for (int iw = 0; iw < 6; iw++)
		for (int k = 0; k < 3; k++)
			y[iw][k] -= qL[iw][k];
	double wR[6];
	maxwell_proj(wI, vn, kp.proj_neg, wR);
	if (kp.kind == GDH_KIN_D3Q6) {
		double pp[6];
		maxwell_proj(wL, vn, kp.proj_pos, pp);
		for (int q = 0; q < 6; q++)
			wR[q] = pp[q] + wR[q];
	}
	double qR[6][3];
	maxwell_qM(wR, qR);
	const double l = kp.lambda, l2 = l * l;
	for (int iw = 0; iw < 6; iw++) {
		const double q0 = y[iw][0] + qR[iw][0], q1 = y[iw][1] + qR[iw][1],
			     q2 = y[iw][2] + qR[iw][2];
		if (kp.kind == GDH_KIN_D3Q4) {
			for (int v = 0; v < NV; v++) {
				const double *vv = kp.vi + 3 * v;
				const double qv = q0 * vv[0] + q1 * vv[1] + q2 * vv[2];
				fR[iw][v] = (wR[iw] + qv / l2) / NV;
			}
		} else {
			/* v = s lambda e_k at index k (s = +1) or 3 + k (s = -1) */
			const double qk[3] = { q0, q1, q2 };
			const double c3 = y[iw][3] / (6. * l2), c4 = y[iw][4] / (6. * l2);
			const double a[3] = { 1, -2, 1 }, b[3] = { 1, 1, -2 };
			for (int v = 0; v < NV; v++) {
				const int k = v % 3;
				const double b0 = wR[iw] / 6., b1 = qk[k] / (2. * l);
				fR[iw][v] = (v < 3 ? b0 + b1 : b0 - b1) + a[k] * c3 + b[k] * c4;
			}
		}
	}
}

} // namespace gdh

#endif

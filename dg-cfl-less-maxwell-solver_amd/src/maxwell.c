/*
 * Maxwell model (eps = mu = 1): numerical fluxes, eigen-projectors and exact
 * solutions. Restates src/models/maxwell.c:14-223; the parameterised plane
 * wave and cavity mode are additions.
 *
 * These host functions are the callbacks a gdn_model points to. The GPU
 * solver recognises them by address and runs the same formulas in its
 * kernels (csrc/physics.h); their parameters are read through
 * gdn_imposed_params().
 */
#include <math.h>
#include <string.h>

#include <models/maxwell.h>
#include "host_internal.h"

/* ------------------------------------------------------------- fluxes */

/* -n x {H} + n x n x [E] / r  |  n x {E} + n x n x [H] / r,
 * {.} = (R + L)/2, [.] = (R - L)/2, 1/r = 1/(|n| + 1e-16)
 * (src/models/maxwell.c:152-192) */
void maxwell_num_flux_upwind(const gdn_real *wL, const gdn_real *wR,
			     const gdn_real *vnorm, gdn_real *flux)
{
	const gdn_real n[3] = { vnorm[0], vnorm[1], vnorm[2] };
	const gdn_real overr =
		1.0 / (sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) + 1e-16);
	gdn_real avg[6], jmp[6];
	for (int q = 0; q < 6; q++) {
		avg[q] = 0.5 * (wR[q] + wL[q]);
		jmp[q] = 0.5 * (wR[q] - wL[q]);
	}
	/* n x (n x u) / r = (n (n.u) - |n|^2 u) / r */
	for (int blk = 0; blk < 2; blk++) {
		const gdn_real *u = jmp + 3 * blk;
		const gdn_real ndu = n[0] * u[0] + n[1] * u[1] + n[2] * u[2];
		const gdn_real nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
		/* curl part acts on the other block: E-flux uses H, H-flux uses E */
		const gdn_real *c = avg + 3 * (1 - blk);
		const gdn_real sgn = blk == 0 ? -1.0 : 1.0;
		const gdn_real cr[3] = { n[1] * c[2] - n[2] * c[1],
					 n[2] * c[0] - n[0] * c[2],
					 n[0] * c[1] - n[1] * c[0] };
		for (int d = 0; d < 3; d++)
			flux[3 * blk + d] =
				sgn * cr[d] + overr * (n[d] * ndu - nn * u[d]);
	}
}

/* (-n x H_L, 0) (src/models/maxwell.c:205-223) */
void maxwell_num_flux_boundary_metal(const gdn_real *wL, const gdn_real *wR,
				     const gdn_real *vnorm, gdn_real *flux)
{
	(void)wR;
	const gdn_real *n = vnorm, *H = wL + 3;
	flux[0] = n[2] * H[1] - n[1] * H[2];
	flux[1] = n[0] * H[2] - n[2] * H[0];
	flux[2] = n[1] * H[0] - n[0] * H[1];
	flux[3] = flux[4] = flux[5] = 0;
}

/* ---------------------------------------------------------- projectors */

/* A(m) w = (-m x H, m x E) for the unit normal m; A^2 w = ((I - m m^T) E,
 * (I - m m^T) H). */
static void maxwell_A(const gdn_real *w, const gdn_real m[3], gdn_real *aw)
{
	const gdn_real *E = w, *H = w + 3;
	aw[0] = -(m[1] * H[2] - m[2] * H[1]);
	aw[1] = -(m[2] * H[0] - m[0] * H[2]);
	aw[2] = -(m[0] * H[1] - m[1] * H[0]);
	aw[3] = m[1] * E[2] - m[2] * E[1];
	aw[4] = m[2] * E[0] - m[0] * E[2];
	aw[5] = m[0] * E[1] - m[1] * E[0];
}

static void maxwell_proj(const gdn_real *w, const gdn_real *n, gdn_real *res,
			 gdn_real c_id, gdn_real c_a2, gdn_real c_a)
{
	const gdn_real nrm = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
	const gdn_real m[3] = { n[0] / nrm, n[1] / nrm, n[2] / nrm };
	gdn_real aw[6], a2w[6];
	maxwell_A(w, m, aw);
	for (int blk = 0; blk < 2; blk++) {
		const gdn_real *u = w + 3 * blk;
		const gdn_real mu = m[0] * u[0] + m[1] * u[1] + m[2] * u[2];
		for (int d = 0; d < 3; d++)
			a2w[3 * blk + d] = u[d] - m[d] * mu;
	}
	for (int q = 0; q < 6; q++)
		res[q] = c_id * w[q] + c_a2 * a2w[q] + c_a * aw[q];
}

void maxwell_get_proj_neg_prod_w(const gdn_real *w, const gdn_real *n,
				 gdn_real *res)
{
	maxwell_proj(w, n, res, 0.0, 0.5, -0.5); /* (A^2 - A)/2 */
}

void maxwell_get_proj_neg_null_prod_w(const gdn_real *w, const gdn_real *n,
				      gdn_real *res)
{
	maxwell_proj(w, n, res, 1.0, -0.5, -0.5); /* (A^2 - A)/2 + I - A^2 */
}

void maxwell_get_proj_pos_prod_w(const gdn_real *w, const gdn_real *n,
				 gdn_real *res)
{
	maxwell_proj(w, n, res, 0.0, 0.5, 0.5); /* (A^2 + A)/2 */
}

void maxwell_get_proj_pos_null_prod_w(const gdn_real *w, const gdn_real *n,
				      gdn_real *res)
{
	maxwell_proj(w, n, res, 1.0, -0.5, 0.5); /* (A^2 + A)/2 + I - A^2 */
}

/* ---------------------------------------------- reference exact solutions */

/* src/models/maxwell.c:16-60, 72-84: a static field built from harmonic
 * functions; pi appears both as M_PI and as the 10-digit 0.3141592654e1 */
#define PI10 0.3141592654e1

static gdn_real st_g(gdn_real x)
{
	return sin(2. * M_PI * (x - 0.5));
}

static gdn_real st_f(gdn_real y, gdn_real z)
{
	return exp(-2 * M_PI * (y - 0.5)) * sin(2 * M_PI * (z - 0.5)) +
	       exp(-2 * M_PI * (z - 0.5)) * sin(2 * M_PI * (y - 0.5));
}

static gdn_real st_df1(gdn_real y, gdn_real z)
{
	return -2. * PI10 * exp(-2. * PI10 * (y - 0.5)) *
		       sin(2. * PI10 * (z - 0.5)) +
	       2. * exp(-2. * PI10 * (z - 0.5)) * PI10 *
		       cos(2. * PI10 * (y - 0.5));
}

static gdn_real st_df2(gdn_real y, gdn_real z)
{
	return 2. * exp(-2. * PI10 * (y - 0.5)) * PI10 *
		       cos(2. * PI10 * (z - 0.5)) -
	       2. * exp(-2. * PI10 * (z - 0.5)) * PI10 *
		       sin(2. * PI10 * (y - 0.5));
}

void maxwell_imposed_stationary(const gdn_real *x, const gdn_real t,
				gdn_real *w)
{
	w[0] = st_g(x[0]) + st_f(x[1], x[2]);
	w[1] = st_g(x[1]) + st_f(x[2], x[0]);
	w[2] = st_g(x[2]) + st_f(x[0], x[1]);
	w[3] = (st_df1(x[2], x[0]) - st_df2(x[0], x[1])) * t;
	w[4] = (st_df1(x[0], x[1]) - st_df2(x[1], x[2])) * t;
	w[5] = (st_df1(x[1], x[2]) - st_df2(x[2], x[0])) * t;
}

/* Ez = cos(5 pi (x - t)), Hy = -Ez (src/models/maxwell.c:95-107) */
void maxwell_imposed_macro_cos(const gdn_real *x, const gdn_real t, gdn_real *w)
{
	const gdn_real s = cos(5 * M_PI * (x[0] - t));
	w[0] = 0;
	w[1] = 0;
	w[2] = s;
	w[3] = 0;
	w[4] = -s;
	w[5] = 0;
}

/* TM(1,1) cavity field (src/models/maxwell.c:118-133) */
void maxwell_imposed_macro_cavity(const gdn_real *x, const gdn_real t,
				  gdn_real *w)
{
	const gdn_real r2 = sqrt(2.);
	const gdn_real sx = sin(PI10 * x[0]), cx = cos(PI10 * x[0]);
	const gdn_real sy = sin(PI10 * x[1]), cy = cos(PI10 * x[1]);
	const gdn_real st = sin(r2 * PI10 * t);
	w[0] = 0;
	w[1] = 0;
	w[2] = sx * sy * cos(r2 * PI10 * t);
	w[3] = -st * r2 * sx * cy / 2.;
	w[4] = st * r2 * cx * sy / 2.;
	w[5] = 0;
}

/* --------------------------------------------- parameterised solutions */

static struct gdn_imposed_params g_par = {
	/* plane wave: k = 2 pi x, e0 = y, h0 = z */
	.pw_k = { 2 * M_PI, 0, 0 },
	.pw_e = { 0, 1, 0 },
	.pw_h = { 0, 0, 1 },
	.pw_omega = 2 * M_PI,
	.pw_phase = 0,
	/* cavity mode (1,1,1), a = (1,-1,0) */
	.cm_k = { M_PI, M_PI, M_PI },
	.cm_a = { 1, -1, 0 },
	.cm_omega = M_PI * 1.7320508075688772,
	/* Gaussian pulse: centred in the unit cube, sigma 0.08 */
	.gp_c = { 0.5, 0.5, 0.5 },
	.gp_sigma = 0.08,
	.gp_amp = 1.0,
};

const struct gdn_imposed_params *gdn_imposed_params(void)
{
	return &g_par;
}

void maxwell_plane_wave_set(const gdn_real k[3], const gdn_real e0[3],
			    gdn_real phase)
{
	const gdn_real kk = sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
	for (int d = 0; d < 3; d++) {
		g_par.pw_k[d] = k[d];
		g_par.pw_e[d] = e0[d];
	}
	/* h0 = (k/|k|) x e0 */
	g_par.pw_h[0] = (k[1] * e0[2] - k[2] * e0[1]) / kk;
	g_par.pw_h[1] = (k[2] * e0[0] - k[0] * e0[2]) / kk;
	g_par.pw_h[2] = (k[0] * e0[1] - k[1] * e0[0]) / kk;
	g_par.pw_omega = kk;
	g_par.pw_phase = phase;
}

/* the same wave in a medium of permittivity eps (mu = 1): speed
 * 1 / sqrt(eps), H = sqrt(eps) (k/|k|) x e0 s */
void maxwell_plane_wave_set_eps(const gdn_real k[3], const gdn_real e0[3], gdn_real phase,
				gdn_real eps)
{
	maxwell_plane_wave_set(k, e0, phase);
	const gdn_real r = sqrt(eps);
	for (int d = 0; d < 3; d++)
		g_par.pw_h[d] *= r;
	g_par.pw_omega /= r;
}

void maxwell_gaussian_pulse_set(const gdn_real center[3], gdn_real sigma, gdn_real amp)
{
	for (int d = 0; d < 3; d++)
		g_par.gp_c[d] = center[d];
	g_par.gp_sigma = sigma;
	g_par.gp_amp = amp;
}

/* E_z = amp exp(-|x - c|^2 / (2 sigma^2)), every other field 0, at any t:
 * the initial state of a pulse and the data kept on the boundary */
void maxwell_imposed_gaussian_pulse(const gdn_real *x, const gdn_real t, gdn_real *w)
{
	(void)t;
	const struct gdn_imposed_params *p = &g_par;
	const gdn_real d0 = x[0] - p->gp_c[0], d1 = x[1] - p->gp_c[1], d2 = x[2] - p->gp_c[2];
	const gdn_real r2 = d0 * d0 + d1 * d1 + d2 * d2;
	for (int q = 0; q < 6; q++)
		w[q] = 0;
	w[2] = p->gp_amp * exp(-r2 / (2 * p->gp_sigma * p->gp_sigma));
}

void maxwell_imposed_plane_wave(const gdn_real *x, const gdn_real t,
				gdn_real *w)
{
	const struct gdn_imposed_params *p = &g_par;
	const gdn_real s = sin(p->pw_k[0] * x[0] + p->pw_k[1] * x[1] +
			       p->pw_k[2] * x[2] - p->pw_omega * t + p->pw_phase);
	for (int d = 0; d < 3; d++) {
		w[d] = p->pw_e[d] * s;
		w[3 + d] = p->pw_h[d] * s;
	}
}

void maxwell_cavity_mode_set(int l, int m, int n, const gdn_real a[3])
{
	g_par.cm_k[0] = l * M_PI;
	g_par.cm_k[1] = m * M_PI;
	g_par.cm_k[2] = n * M_PI;
	for (int d = 0; d < 3; d++)
		g_par.cm_a[d] = a[d];
	g_par.cm_omega = M_PI * sqrt((double)(l * l + m * m + n * n));
}

void maxwell_imposed_cavity_mode(const gdn_real *x, const gdn_real t,
				 gdn_real *w)
{
	const struct gdn_imposed_params *p = &g_par;
	gdn_real s[3], c[3];
	for (int d = 0; d < 3; d++) {
		s[d] = sin(p->cm_k[d] * x[d]);
		c[d] = cos(p->cm_k[d] * x[d]);
	}
	const gdn_real ct = cos(p->cm_omega * t);
	const gdn_real st = sin(p->cm_omega * t) / p->cm_omega;
	const gdn_real *a = p->cm_a, *k = p->cm_k;
	w[0] = a[0] * c[0] * s[1] * s[2] * ct;
	w[1] = a[1] * s[0] * c[1] * s[2] * ct;
	w[2] = a[2] * s[0] * s[1] * c[2] * ct;
	/* H = -(sin(om t)/om) curl E0 */
	w[3] = -st * (k[1] * a[2] - k[2] * a[1]) * s[0] * c[1] * c[2];
	w[4] = -st * (k[2] * a[0] - k[0] * a[2]) * c[0] * s[1] * c[2];
	w[5] = -st * (k[0] * a[1] - k[1] * a[0]) * c[0] * c[1] * s[2];
}

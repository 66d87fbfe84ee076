/*
 * Maxwell's equations in vacuum (eps = mu = 1), w = (Ex,Ey,Ez,Hx,Hy,Hz):
 *   d_t E - curl H = 0,  d_t H + curl E = 0.
 * Replaces src/models/maxwell.h:14-39 and adds parameterised exact solutions.
 */
#ifndef MAXWELL_H
#define MAXWELL_H

#include <gdon3d.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Exact solutions of the reference (src/models/maxwell.c:72-133). */
void maxwell_imposed_stationary(const gdn_real *x, const gdn_real t,
				gdn_real *w);
void maxwell_imposed_macro_cos(const gdn_real *x, const gdn_real t,
			       gdn_real *w);
void maxwell_imposed_macro_cavity(const gdn_real *x, const gdn_real t,
				  gdn_real *w);

/* Upwind flux with un-normalised normal, 1/r = 1/(|n| + 1e-16)
 * (src/models/maxwell.c:152-192). */
void maxwell_num_flux_upwind(const gdn_real *wL, const gdn_real *wR,
			     const gdn_real *vnorm, gdn_real *flux);

/* Perfect-conductor boundary flux (-n x H_L, 0) (src/models/maxwell.c:205-223). */
void maxwell_num_flux_boundary_metal(const gdn_real *wL, const gdn_real *wR,
				     const gdn_real *vnorm, gdn_real *flux);

/* Projections of w on the eigenspaces of A(n/|n|), whose eigenvalues are
 * -1, 0, +1 (src/models/maxwell.c:234-575): neg = (A^2-A)/2,
 * pos = (A^2+A)/2, *_null adds the kernel projector I - A^2. */
void maxwell_get_proj_neg_prod_w(const gdn_real *w, const gdn_real *n,
				 gdn_real *res);
void maxwell_get_proj_neg_null_prod_w(const gdn_real *w, const gdn_real *n,
				      gdn_real *res);
void maxwell_get_proj_pos_prod_w(const gdn_real *w, const gdn_real *n,
				 gdn_real *res);
void maxwell_get_proj_pos_null_prod_w(const gdn_real *w, const gdn_real *n,
				      gdn_real *res);

/* ----- Parameterised exact solutions (no counterpart in the reference) ----- */

/* Plane wave E = e0 s, H = (k/|k|) x e0 s with s = sin(k.x - |k| t + phase).
 * Parameters are process-global; set them before the solve. e0 must be
 * orthogonal to k. Default: k = 2 pi (1,0,0), e0 = (0,1,0), phase 0, i.e.
 * Ey = Hz = sin(2 pi (x - t)). */
void maxwell_plane_wave_set(const gdn_real k[3], const gdn_real e0[3],
			    gdn_real phase);
void maxwell_imposed_plane_wave(const gdn_real *x, const gdn_real t,
				gdn_real *w);

/* The plane wave in a medium of permittivity eps (mu = 1): speed
 * 1/sqrt(eps), H = sqrt(eps) (k/|k|) x e0 s, s = sin(k.x - |k| t / sqrt(eps)
 * + phase); the exact solution for a uniform permittivity eps. */
void maxwell_plane_wave_set_eps(const gdn_real k[3], const gdn_real e0[3], gdn_real phase,
				gdn_real eps);

/* Gaussian pulse E_z = amp exp(-|x - c|^2 / (2 sigma^2)), every other field
 * zero, the same at every t (initial state; boundary data that stay at the
 * pulse's tail). Default: c = (1/2, 1/2, 1/2), sigma = 0.08, amp = 1. */
void maxwell_gaussian_pulse_set(const gdn_real center[3], gdn_real sigma, gdn_real amp);
void maxwell_imposed_gaussian_pulse(const gdn_real *x, const gdn_real t, gdn_real *w);

/* Resonant mode (l,m,n) of the perfectly conducting unit cube:
 *   E = (a0 cx sy sz, a1 sx cy sz, a2 sx sy cz) cos(om t),
 *   H = -(sin(om t)/om) curl(E(t=0)),
 * cx = cos(l pi x), sx = sin(l pi x), ..., om = pi sqrt(l^2+m^2+n^2); the
 * amplitude must satisfy l a0 + m a1 + n a2 = 0. Default: mode (1,1,1),
 * a = (1,-1,0). */
void maxwell_cavity_mode_set(int l, int m, int n, const gdn_real a[3]);
void maxwell_imposed_cavity_mode(const gdn_real *x, const gdn_real t,
				 gdn_real *w);

#ifdef __cplusplus
}
#endif
#endif

/*
 * D3Q6 kinetic velocity set for Maxwell (src/models/d3q6.h:14-49, every
 * declaration kept): six axis
 * velocities +-lambda e_k, lambda = 3, omega = 1.999999999999.
 */
#ifndef D3Q6_H
#define D3Q6_H

#include <gdon3d.h>

#ifdef __cplusplus
extern "C" {
#endif

void d3q6_set_model(void *self);
int d3q6_get_nb_v(void);
gdn_real d3q6_get_lambda(void);
gdn_real *d3q6_get_vi(void);
void d3q6_get_w(const void *self, const gdn_real *f, gdn_real *w);
/* moments w and q (X, Y, Z, X^2 - Y^2, X^2 - Z^2; q[5 iw + k]) and back */
void d3q6_apply_M(const void *self, const gdn_real *f, gdn_real *w, gdn_real *q);
void d3q6_apply_M_inv(const void *self, const gdn_real *w, const gdn_real *q, gdn_real *f);
/* physical flux in moments, zero second-order moments */
void d3q6_get_qM(const void *self, const gdn_real *w, gdn_real *qM);
/* (w, y) = Q(f), y = q - qM(w); f = Q^-1(w, y) */
void d3q6_apply_Q(const void *self, const gdn_real *f, gdn_real *w, gdn_real *y);
void d3q6_apply_Q_inv(const void *self, const gdn_real *w, const gdn_real *y, gdn_real *f);
void d3q6_get_feq(const void *self, const gdn_real *w, gdn_real *feq);
/* boundary state f_eq(wI), the reference's get_fR (use_feq = true) */
void d3q6_get_fR(const void *self, const gdn_real *fL, const gdn_real *vn,
		 const gdn_real *wI, gdn_real *fR);
/* boundary state Q^-1(P^{+,0}(vn) wL + P^-(vn) wI, yL), the branch the
 * reference's get_fR takes with use_feq = false (src/models/d3q6.c:358-379);
 * no separate name in the reference: a model selects it through get_fR */
void d3q6_get_fR_projector(const void *self, const gdn_real *fL, const gdn_real *vn,
			   const gdn_real *wI, gdn_real *fR);

#ifdef __cplusplus
}
#endif
#endif

/*
 * D3Q4 kinetic velocity set for Maxwell (src/models/d3q4.h:14-44, every
 * declaration kept):
 * v = lambda (1,1,1), lambda (1,-1,-1), lambda (-1,1,-1), lambda (-1,-1,1),
 * lambda = sqrt(3), omega = 1.999999999999. Micro-field order inside a node is
 * iw * nb_v + iv.
 */
#ifndef D3Q4_H
#define D3Q4_H

#include <gdon3d.h>

#ifdef __cplusplus
extern "C" {
#endif

void d3q4_set_model(void *self);
int d3q4_get_nb_v(void);
gdn_real d3q4_get_lambda(void);
gdn_real *d3q4_get_vi(void);
void d3q4_get_w(const void *self, const gdn_real *f, gdn_real *w);
void d3q4_apply_M(const void *self, const gdn_real *f, gdn_real *w, gdn_real *q);
void d3q4_apply_M_inv(const void *self, const gdn_real *w, const gdn_real *q,
		      gdn_real *f);
void d3q4_get_qM(const void *self, const gdn_real *w, gdn_real *qM);
/* (w, y) = Q(f), y = q - qM(w) (src/models/d3q4.h:32-33) */
void d3q4_apply_Q(const void *self, const gdn_real *f, gdn_real *w, gdn_real *y);
/* f = Q^-1(w, y) (src/models/d3q4.h:35-36) */
void d3q4_apply_Q_inv(const void *self, const gdn_real *w, const gdn_real *y, gdn_real *f);
void d3q4_get_feq(const void *self, const gdn_real *w, gdn_real *feq);
/* boundary state f_eq(wI), the reference's get_fR (use_feq = true) */
void d3q4_get_fR(const void *self, const gdn_real *fL, const gdn_real *vn,
		 const gdn_real *wI, gdn_real *fR);
/* boundary state Q^-1(P^-(vn) wI, yL) from the projectors, the branch the
 * reference's get_fR takes with use_feq = false (src/models/d3q4.c:303-332);
 * no separate name in the reference: a model selects it through get_fR */
void d3q4_get_fR_projector(const void *self, const gdn_real *fL, const gdn_real *vn,
			   const gdn_real *wI, gdn_real *fR);
/* src/models/d3q4.h:43-44 */
void test_philippe_d3q4_get_f(const void *self, const gdn_real *w_inc, gdn_real *f);

#ifdef __cplusplus
}
#endif
#endif

/*
 * D3Q4 kinetic velocity set for Maxwell (src/models/d3q4.h:14-44):
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
void d3q4_get_feq(const void *self, const gdn_real *w, gdn_real *feq);
void d3q4_get_fR(const void *self, const gdn_real *fL, const gdn_real *vn,
		 const gdn_real *wI, gdn_real *fR);

#ifdef __cplusplus
}
#endif
#endif

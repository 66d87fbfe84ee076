/*
 * D3Q6 kinetic velocity set for Maxwell (src/models/d3q6.h): six axis
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
void d3q6_get_feq(const void *self, const gdn_real *w, gdn_real *feq);
void d3q6_get_fR(const void *self, const gdn_real *fL, const gdn_real *vn,
		 const gdn_real *wI, gdn_real *fR);

#ifdef __cplusplus
}
#endif
#endif

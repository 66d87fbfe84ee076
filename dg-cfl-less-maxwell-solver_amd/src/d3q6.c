/*
 * D3Q6 kinetic model (src/models/d3q6.c:19-379): six velocities
 * +-lambda e_k, lambda = 3, moments {1, X, Y, Z, X^2-Y^2, X^2-Z^2}, the last
 * two set to 0 at equilibrium:
 * f_eq(w)_{iw,v} = w_iw / 6 + (q_iw . v) / (2 lambda^2).
 */
#include <models/d3q6.h>
#include <models/model.h>

#define NV 6
#define LAMBDA 3.

static gdn_real vel[3 * NV] = { LAMBDA, 0, 0, 0, LAMBDA,  0, 0, 0, LAMBDA,
				-LAMBDA, 0, 0, 0, -LAMBDA, 0, 0, 0, -LAMBDA };

void d3q6_set_model(void *self)
{
	gdn_model *m = self;
	m->nb_v = NV;
	m->lambda = LAMBDA;
	m->vi = vel;
	m->get_w = d3q6_get_w;
	m->get_feq = d3q6_get_feq;
	m->get_fR = d3q6_get_fR;
	m->omega = 1.999999999999;
}

int d3q6_get_nb_v(void)
{
	return NV;
}

gdn_real d3q6_get_lambda(void)
{
	return LAMBDA;
}

gdn_real *d3q6_get_vi(void)
{
	return vel;
}

void d3q6_get_w(const void *self, const gdn_real *f, gdn_real *w)
{
	const gdn_model *m = self;
	for (int iw = 0; iw < m->nb_w; iw++) {
		gdn_real s = 0;
		for (int iv = 0; iv < m->nb_v; iv++)
			s += f[iw * m->nb_v + iv];
		w[iw] = s;
	}
}

void d3q6_get_feq(const void *self, const gdn_real *w, gdn_real *feq)
{
	const gdn_model *m = self;
	const gdn_real e[3][3] = { { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } };
	gdn_real fl[3][64];
	for (int k = 0; k < 3; k++)
		m->get_num_flux(w, w, e[k], fl[k]);
	const gdn_real l = m->lambda;
	for (int iw = 0; iw < m->nb_w; iw++)
		for (int k = 0; k < 3; k++) {
			const gdn_real a = w[iw] / 6., b = fl[k][iw] / (2. * l);
			feq[iw * NV + k] = a + b;
			feq[iw * NV + 3 + k] = a - b;
		}
}

/* f_R = f_eq of the imposed data (use_feq = true, src/models/d3q6.c:350-356) */
void d3q6_get_fR(const void *self, const gdn_real *fL, const gdn_real *vn,
		 const gdn_real *wI, gdn_real *fR)
{
	(void)fL;
	(void)vn;
	d3q6_get_feq(self, wI, fR);
}

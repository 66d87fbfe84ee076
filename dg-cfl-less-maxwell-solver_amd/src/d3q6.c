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

/* M(f) = (w, q): w = sum_v f and, per field, the moments X, Y, Z,
 * X^2 - Y^2, X^2 - Z^2 of f (src/models/d3q6.c:100-141), q[5 iw + k] */
void d3q6_apply_M(const void *self, const gdn_real *f, gdn_real *w, gdn_real *q)
{
	const gdn_model *m = self;
	for (int iw = 0; iw < m->nb_w; iw++) {
		gdn_real s = 0, r[5] = { 0, 0, 0, 0, 0 };
		for (int iv = 0; iv < NV; iv++) {
			const gdn_real fi = f[iw * NV + iv];
			const gdn_real *v = m->vi + 3 * iv;
			s += fi;
			r[0] += fi * v[0];
			r[1] += fi * v[1];
			r[2] += fi * v[2];
			r[3] += fi * (v[0] * v[0] - v[1] * v[1]);
			r[4] += fi * (v[0] * v[0] - v[2] * v[2]);
		}
		w[iw] = s;
		for (int k = 0; k < 5; k++)
			q[5 * iw + k] = r[k];
	}
}

/* f = M^-1(w, q) (src/models/d3q6.c:143-208): for v = s lambda e_k,
 * f = w/6 + s q_k/(2 lambda) + a_k q_3/(6 lambda^2) + b_k q_4/(6 lambda^2)
 * with (a_k, b_k) = (1, 1), (-2, 1), (1, -2) for k = x, y, z */
void d3q6_apply_M_inv(const void *self, const gdn_real *w, const gdn_real *q, gdn_real *f)
{
	const gdn_model *m = self;
	const gdn_real l = m->lambda, l2 = m->lambda * m->lambda;
	static const gdn_real a[3] = { 1, -2, 1 }, b[3] = { 1, 1, -2 };
	for (int iw = 0; iw < m->nb_w; iw++) {
		const gdn_real *r = q + 5 * iw;
		const gdn_real c3 = r[3] / (6. * l2), c4 = r[4] / (6. * l2);
		for (int k = 0; k < 3; k++) {
			const gdn_real b0 = w[iw] / 6., b1 = r[k] / (2. * l);
			f[iw * NV + k] = b0 + b1 + a[k] * c3 + b[k] * c4;
			f[iw * NV + 3 + k] = b0 - b1 + a[k] * c3 + b[k] * c4;
		}
	}
}

/* qM(w): the physical flux A(e_k) w in moments X, Y, Z and zero in the two
 * second-order moments (src/models/d3q6.c:210-244) */
void d3q6_get_qM(const void *self, const gdn_real *w, gdn_real *qM)
{
	const gdn_model *m = self;
	const gdn_real e[3][3] = { { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } };
	gdn_real fl[3][64];
	for (int k = 0; k < 3; k++)
		m->get_num_flux(w, w, e[k], fl[k]);
	for (int iw = 0; iw < m->nb_w; iw++) {
		for (int k = 0; k < 3; k++)
			qM[5 * iw + k] = fl[k][iw];
		qM[5 * iw + 3] = qM[5 * iw + 4] = 0;
	}
}

/* (w, y) = Q(f), y = q - qM(w) (src/models/d3q6.c:246-276) */
void d3q6_apply_Q(const void *self, const gdn_real *f, gdn_real *w, gdn_real *y)
{
	const gdn_model *m = self;
	gdn_real q[5 * 64], qM[5 * 64];
	d3q6_apply_M(m, f, w, q);
	d3q6_get_qM(m, w, qM);
	for (int k = 0; k < 5 * m->nb_w; k++)
		y[k] = q[k] - qM[k];
}

/* f = Q^-1(w, y) = M^-1(w, y + qM(w)) (src/models/d3q6.c:278-304) */
void d3q6_apply_Q_inv(const void *self, const gdn_real *w, const gdn_real *y, gdn_real *f)
{
	const gdn_model *m = self;
	gdn_real q[5 * 64];
	d3q6_get_qM(m, w, q);
	for (int k = 0; k < 5 * m->nb_w; k++)
		q[k] = y[k] + q[k];
	d3q6_apply_M_inv(m, w, q, f);
}

/* f_R = f_eq of the imposed data (use_feq = true, src/models/d3q6.c:350-356) */
void d3q6_get_fR(const void *self, const gdn_real *fL, const gdn_real *vn,
		 const gdn_real *wI, gdn_real *fR)
{
	(void)fL;
	(void)vn;
	d3q6_get_feq(self, wI, fR);
}

/* The reference's projector boundary state, the branch its use_feq = false
 * selects (src/models/d3q6.c:358-379): (wL, yL) = Q(fL),
 * wR = P^{+,0}(vn) wL + P^-(vn) wI through the model's get_proj_pos /
 * get_proj_neg, fR = Q^-1(wR, yL). A model selects it by pointing get_fR
 * here; the GPU solver runs the same on the device. */
void d3q6_get_fR_projector(const void *self, const gdn_real *fL, const gdn_real *vn,
			   const gdn_real *wI, gdn_real *fR)
{
	const gdn_model *m = self;
	gdn_real wL[64], wR[64], pp[64], pn[64], yL[5 * 64];
	d3q6_apply_Q(m, fL, wL, yL);
	m->get_proj_pos(wL, vn, pp);
	m->get_proj_neg(wI, vn, pn);
	for (int iw = 0; iw < m->nb_w; iw++)
		wR[iw] = pp[iw] + pn[iw];
	d3q6_apply_Q_inv(m, wR, yL, fR);
}

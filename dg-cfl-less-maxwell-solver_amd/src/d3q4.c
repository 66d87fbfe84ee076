/*
 * D3Q4 kinetic model (src/models/d3q4.c:17-332): four velocities
 * lambda (+-1,+-1,+-1) with an even number of minus signs, lambda = sqrt(3).
 * f_eq(w)_{iw,v} = (w_iw + (q_iw . v) / lambda^2) / 4 with
 * q_iw,k = (A(e_k) w)_iw the physical flux, so that sum_v f_eq = w and
 * sum_v v f_eq = q.
 */
#include <models/d3q4.h>
#include <models/model.h>

#define NV 4
#define LAMBDA 1.732050807568877293

static gdn_real vel[3 * NV] = { LAMBDA,	 LAMBDA,  LAMBDA,  LAMBDA,
				-LAMBDA, -LAMBDA, -LAMBDA, LAMBDA,
				-LAMBDA, -LAMBDA, -LAMBDA, LAMBDA };

void d3q4_set_model(void *self)
{
	gdn_model *m = self;
	m->nb_v = NV;
	m->lambda = LAMBDA;
	m->vi = vel;
	m->get_w = d3q4_get_w;
	m->get_feq = d3q4_get_feq;
	m->get_fR = d3q4_get_fR;
	m->omega = 1.999999999999;
}

int d3q4_get_nb_v(void)
{
	return NV;
}

gdn_real d3q4_get_lambda(void)
{
	return LAMBDA;
}

gdn_real *d3q4_get_vi(void)
{
	return vel;
}

void d3q4_get_w(const void *self, const gdn_real *f, gdn_real *w)
{
	const gdn_model *m = self;
	for (int iw = 0; iw < m->nb_w; iw++) {
		gdn_real s = 0;
		for (int iv = 0; iv < m->nb_v; iv++)
			s += f[iw * m->nb_v + iv];
		w[iw] = s;
	}
}

void d3q4_apply_M(const void *self, const gdn_real *f, gdn_real *w, gdn_real *q)
{
	const gdn_model *m = self;
	for (int iw = 0; iw < m->nb_w; iw++) {
		gdn_real s = 0, q0 = 0, q1 = 0, q2 = 0;
		for (int iv = 0; iv < m->nb_v; iv++) {
			const gdn_real fi = f[iw * m->nb_v + iv];
			s += fi;
			q0 += fi * m->vi[3 * iv + 0];
			q1 += fi * m->vi[3 * iv + 1];
			q2 += fi * m->vi[3 * iv + 2];
		}
		w[iw] = s;
		q[3 * iw + 0] = q0;
		q[3 * iw + 1] = q1;
		q[3 * iw + 2] = q2;
	}
}

void d3q4_apply_M_inv(const void *self, const gdn_real *w, const gdn_real *q,
		      gdn_real *f)
{
	const gdn_model *m = self;
	const gdn_real l2 = m->lambda * m->lambda;
	for (int iw = 0; iw < m->nb_w; iw++)
		for (int iv = 0; iv < m->nb_v; iv++) {
			const gdn_real *v = m->vi + 3 * iv;
			const gdn_real qv = q[3 * iw + 0] * v[0] +
					    q[3 * iw + 1] * v[1] +
					    q[3 * iw + 2] * v[2];
			f[iw * m->nb_v + iv] = (w[iw] + qv / l2) / m->nb_v;
		}
}

void d3q4_get_qM(const void *self, const gdn_real *w, gdn_real *qM)
{
	const gdn_model *m = self;
	const gdn_real e[3][3] = { { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } };
	gdn_real fl[3][64];
	for (int k = 0; k < 3; k++)
		m->get_num_flux(w, w, e[k], fl[k]);
	for (int iw = 0; iw < m->nb_w; iw++)
		for (int k = 0; k < 3; k++)
			qM[3 * iw + k] = fl[k][iw];
}

void d3q4_get_feq(const void *self, const gdn_real *w, gdn_real *feq)
{
	const gdn_model *m = self;
	gdn_real qM[3 * 64];
	d3q4_get_qM(m, w, qM);
	d3q4_apply_M_inv(m, w, qM, feq);
}

/* (w, y) = Q(f): the moments w = sum_v f, q = sum_v v f and their distance
 * y = q - qM(w) to the physical flux (src/models/d3q4.c:200-225) */
void d3q4_apply_Q(const void *self, const gdn_real *f, gdn_real *w, gdn_real *y)
{
	const gdn_model *m = self;
	gdn_real q[3 * 64], qM[3 * 64];
	d3q4_apply_M(m, f, w, q);
	d3q4_get_qM(m, w, qM);
	for (int k = 0; k < 3 * m->nb_w; k++)
		y[k] = q[k] - qM[k];
}

/* f = Q^-1(w, y) = M^-1(w, y + qM(w)) (src/models/d3q4.c:227-252) */
void d3q4_apply_Q_inv(const void *self, const gdn_real *w, const gdn_real *y, gdn_real *f)
{
	const gdn_model *m = self;
	gdn_real q[3 * 64];
	d3q4_get_qM(m, w, q);
	for (int k = 0; k < 3 * m->nb_w; k++)
		q[k] = y[k] + q[k];
	d3q4_apply_M_inv(m, w, q, f);
}

/* src/models/d3q4.c:334-352: f = Q^-1(w_inc, y) with every component of
 * y_iw equal to w_inc[iw] */
void test_philippe_d3q4_get_f(const void *self, const gdn_real *w_inc, gdn_real *f)
{
	const gdn_model *m = self;
	gdn_real y[3 * 64];
	for (int iw = 0; iw < m->nb_w; iw++)
		y[3 * iw + 0] = y[3 * iw + 1] = y[3 * iw + 2] = w_inc[iw];
	d3q4_apply_Q_inv(m, w_inc, y, f);
}

/* The reference evaluates the boundary state as f_eq of the imposed data
 * (use_feq = true, src/models/d3q4.c:295-301); fL and vn are unused. */
void d3q4_get_fR(const void *self, const gdn_real *fL, const gdn_real *vn,
		 const gdn_real *wI, gdn_real *fR)
{
	(void)fL;
	(void)vn;
	d3q4_get_feq(self, wI, fR);
}

/* The reference's other boundary state, the branch its use_feq = false
 * selects (src/models/d3q4.c:303-332): (wL, yL) = Q(fL), wR = P^-(vn) wI
 * through the model's get_proj_neg (the reference overwrites the sum
 * P^{+,0} wL + P^- wI with P^- wI), fR = Q^-1(wR, yL). A model selects it by
 * pointing get_fR here; the GPU solver runs the same on the device. */
void d3q4_get_fR_projector(const void *self, const gdn_real *fL, const gdn_real *vn,
			   const gdn_real *wI, gdn_real *fR)
{
	const gdn_model *m = self;
	gdn_real wL[64], wR[64], yL[3 * 64];
	d3q4_apply_Q(m, fL, wL, yL);
	m->get_proj_neg(wI, vn, wR);
	d3q4_apply_Q_inv(m, wR, yL, fR);
}

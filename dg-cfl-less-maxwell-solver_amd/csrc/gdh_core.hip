/*
 * GPU layer of gdon3d-hip: context, state transfers, initial and boundary
 * data, the fused RK3 stage, the kinetic transport sweep and relaxation, and
 * the L2 error reduction. Host-side C ABI declared in include/gdn_hip.h.
 *
 * Data in HBM (FP64):
 *  - geometry, element-major: geom[e][10] = codtau (row-major) + det,
 *    nbr[e] (int4: neighbour or -1-bf), nbmap[e] (int4: 4 bits per face node);
 *  - RK3 state in the reference AoS layout w[ipg * 6 + iw] (each element's
 *    60 values are one contiguous 480 B block), ping-pong buffers + k2;
 *  - kinetic state in the reference layout f[(iw * nv + iv) * neq + ipg].
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "physics.h"

#define HIPCHECK(call)                                                        \
	do {                                                                  \
		hipError_t err_ = (call);                                     \
		if (err_ != hipSuccess) {                                     \
			fprintf(stderr, "gdon3d-hip: %s failed: %s (%s:%d)\n", \
				#call, hipGetErrorString(err_), __FILE__,     \
				__LINE__);                                    \
			abort();                                              \
		}                                                             \
	} while (0)

using namespace gdh;

/* ===================================================================== */
/*                               context                                 */
/* ===================================================================== */

struct gdh_ctx {
	int ne, nb, neq, m, scheme;
	gdh_model_desc md;
	ImposedPar ip;
	KinPar kp;
	hipStream_t stream;
	hipEvent_t ev0, ev1;

	double *geom;
	int4 *nbr, *nbmap;
	double *xnode, *xghost;
	double *w[2];
	int cur;
	double *k2;
	double *hostvals; /* exact values uploaded for GDH_IMPOSED_HOST */
	size_t hostvals_len;
	double *partials;
	int npartials;

	/* kinetic solver */
	double dt;
	int nlevels;
	std::vector<int> level_ptr;
	int2 *items; /* (v, e) in level order */
	int nitems;
	double *dinv; /* per item: 10x10 inverse of the diagonal block */
	uint8_t *cls; /* per (e, f, v): 0 none, 1 outflow, 2 inflow */

	/* optional event timing of the main kernels: pairs of events recorded
	 * on the stream around each kernel group, resolved when read */
	int prof;
	std::vector<hipEvent_t> ev_pool;
	std::vector<int> ev_slot; /* slot of each recorded pair */
	std::vector<hipEvent_t> ev_pairs; /* begin, end, begin, end, ... */
	double prof_s[3];
	long prof_launches;
};

static hipEvent_t ev_get(gdh_ctx *c)
{
	if (c->ev_pool.empty()) {
		hipEvent_t e;
		HIPCHECK(hipEventCreate(&e));
		return e;
	}
	hipEvent_t e = c->ev_pool.back();
	c->ev_pool.pop_back();
	return e;
}

static void prof_begin(gdh_ctx *c, int slot)
{
	if (!c->prof)
		return;
	hipEvent_t e = ev_get(c);
	HIPCHECK(hipEventRecord(e, c->stream));
	c->ev_pairs.push_back(e);
	c->ev_slot.push_back(slot);
}

static void prof_end(gdh_ctx *c, int slot)
{
	(void)slot;
	if (!c->prof)
		return;
	hipEvent_t e = ev_get(c);
	HIPCHECK(hipEventRecord(e, c->stream));
	c->ev_pairs.push_back(e);
}

/* wait for the recorded pairs and accumulate their durations */
static void prof_resolve(gdh_ctx *c)
{
	HIPCHECK(hipStreamSynchronize(c->stream));
	for (size_t k = 0; k < c->ev_slot.size(); k++) {
		float ms = 0;
		HIPCHECK(hipEventElapsedTime(&ms, c->ev_pairs[2 * k],
					     c->ev_pairs[2 * k + 1]));
		c->prof_s[c->ev_slot[k]] += ms * 1e-3;
	}
	for (hipEvent_t e : c->ev_pairs)
		c->ev_pool.push_back(e);
	c->ev_pairs.clear();
	c->ev_slot.clear();
}

static void *dmalloc(size_t bytes)
{
	void *p = nullptr;
	HIPCHECK(hipMalloc(&p, bytes ? bytes : 16));
	return p;
}

extern "C" int gdh_device_count(void)
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess)
		return 0;
	return n;
}

extern "C" gdh_ctx *gdh_create(const gdh_mesh_desc *mesh,
			       const gdh_model_desc *model, int scheme)
{
	if (gdh_device_count() < 1) {
		fprintf(stderr, "gdon3d-hip: no HIP device available\n");
		abort();
	}
	gdh_ctx *c = new gdh_ctx();
	c->ne = mesh->nelem;
	c->nb = mesh->nbface;
	c->neq = 10 * c->ne + 6 * c->nb;
	c->scheme = scheme;
	c->md = *model;
	c->m = scheme ? model->nb_v * 6 : 6;
	c->ip.kind = model->imposed_kind;
	memcpy(c->ip.p, model->imposed_par, sizeof(c->ip.p));
	memset(&c->kp, 0, sizeof(c->kp));
	c->kp.kind = model->kinetic_kind;
	c->kp.nv = model->nb_v;
	memcpy(c->kp.vi, model->vi, sizeof(c->kp.vi));
	c->kp.lambda = model->lambda;
	c->kp.omega = model->omega;
	if (scheme && c->kp.kind == GDH_KIN_NONE) {
		fprintf(stderr, "gdon3d-hip: kinetic scheme without velocity set\n");
		abort();
	}
	HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
	HIPCHECK(hipEventCreate(&c->ev0));
	HIPCHECK(hipEventCreate(&c->ev1));
	c->prof = 0;
	c->prof_s[0] = c->prof_s[1] = c->prof_s[2] = 0;
	c->prof_launches = 0;

	const size_t ne = c->ne, nb = c->nb;
	c->geom = (double *)dmalloc(sizeof(double) * 10 * ne);
	c->nbr = (int4 *)dmalloc(sizeof(int4) * ne);
	c->nbmap = (int4 *)dmalloc(sizeof(int4) * ne);
	c->xnode = (double *)dmalloc(sizeof(double) * 30 * ne);
	c->xghost = (double *)dmalloc(sizeof(double) * 18 * (nb ? nb : 1));
	HIPCHECK(hipMemcpy(c->geom, mesh->geom, sizeof(double) * 10 * ne,
			   hipMemcpyHostToDevice));
	HIPCHECK(hipMemcpy(c->nbr, mesh->nbr, sizeof(int4) * ne,
			   hipMemcpyHostToDevice));
	HIPCHECK(hipMemcpy(c->nbmap, mesh->nbmap, sizeof(int4) * ne,
			   hipMemcpyHostToDevice));
	HIPCHECK(hipMemcpy(c->xnode, mesh->xnode, sizeof(double) * 30 * ne,
			   hipMemcpyHostToDevice));
	if (nb)
		HIPCHECK(hipMemcpy(c->xghost, mesh->xghost, sizeof(double) * 18 * nb,
				   hipMemcpyHostToDevice));

	const size_t wlen = (size_t)c->m * c->neq;
	for (int b = 0; b < 2; b++) {
		c->w[b] = (double *)dmalloc(sizeof(double) * wlen);
		HIPCHECK(hipMemset(c->w[b], 0, sizeof(double) * wlen));
	}
	c->cur = 0;
	c->k2 = nullptr;
	if (!scheme) {
		c->k2 = (double *)dmalloc(sizeof(double) * wlen);
		HIPCHECK(hipMemset(c->k2, 0, sizeof(double) * wlen));
	}
	c->hostvals = nullptr;
	c->hostvals_len = 0;
	c->npartials = 1024;
	c->partials = (double *)dmalloc(sizeof(double) * 8 * (c->npartials + 1));
	c->items = nullptr;
	c->dinv = nullptr;
	c->cls = nullptr;
	c->nlevels = 0;
	c->nitems = 0;
	return c;
}

extern "C" void gdh_destroy(gdh_ctx *c)
{
	if (!c)
		return;
	HIPCHECK(hipStreamSynchronize(c->stream));
	void *ptrs[] = { c->geom, c->nbr,   c->nbmap,	 c->xnode,
			 c->xghost, c->w[0], c->w[1],	 c->k2,
			 c->hostvals, c->partials, c->items, c->dinv,
			 c->cls };
	for (void *p : ptrs)
		if (p)
			HIPCHECK(hipFree(p));
	HIPCHECK(hipEventDestroy(c->ev0));
	HIPCHECK(hipEventDestroy(c->ev1));
	prof_resolve(c);
	for (hipEvent_t e : c->ev_pool)
		HIPCHECK(hipEventDestroy(e));
	HIPCHECK(hipStreamDestroy(c->stream));
	delete c;
}

extern "C" void gdh_upload_state(gdh_ctx *c, const double *wn)
{
	HIPCHECK(hipMemcpyAsync(c->w[c->cur], wn,
				sizeof(double) * (size_t)c->m * c->neq,
				hipMemcpyHostToDevice, c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
}

extern "C" void gdh_download_state(gdh_ctx *c, double *wn)
{
	HIPCHECK(hipMemcpyAsync(wn, c->w[c->cur],
				sizeof(double) * (size_t)c->m * c->neq,
				hipMemcpyDeviceToHost, c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
}

extern "C" void gdh_synchronize(gdh_ctx *c)
{
	HIPCHECK(hipStreamSynchronize(c->stream));
}

extern "C" void gdh_timer_start(gdh_ctx *c)
{
	HIPCHECK(hipEventRecord(c->ev0, c->stream));
}

extern "C" double gdh_timer_stop(gdh_ctx *c)
{
	HIPCHECK(hipEventRecord(c->ev1, c->stream));
	HIPCHECK(hipEventSynchronize(c->ev1));
	float ms = 0;
	HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
	return ms * 1e-3;
}

extern "C" void gdh_profile(gdh_ctx *c, int enable, double *out)
{
	prof_resolve(c);
	if (out) {
		out[0] = c->prof_s[0];
		out[1] = c->prof_s[1];
		out[2] = c->prof_s[2];
		out[3] = (double)c->prof_launches;
	}
	c->prof = enable;
	c->prof_s[0] = c->prof_s[1] = c->prof_s[2] = 0;
	c->prof_launches = 0;
}

/* copy host exact values (n doubles) to the device scratch of the context */
static const double *stage_hostvals(gdh_ctx *c, const double *h, size_t n)
{
	if (!h)
		return nullptr;
	if (c->hostvals_len < n) {
		if (c->hostvals)
			HIPCHECK(hipFree(c->hostvals));
		c->hostvals = (double *)dmalloc(sizeof(double) * n);
		c->hostvals_len = n;
	}
	HIPCHECK(hipMemcpyAsync(c->hostvals, h, sizeof(double) * n,
				hipMemcpyHostToDevice, c->stream));
	/* the host buffer may be freed as soon as this returns */
	HIPCHECK(hipStreamSynchronize(c->stream));
	return c->hostvals;
}

static inline int nblocks(long n, int bs)
{
	return (int)((n + bs - 1) / bs);
}

/* ===================================================================== */
/*                   initial condition and ghost data                    */
/* ===================================================================== */

/* Write the macro state wv (6 values) of DoF ipg into the state buffer, as
 * w itself (RK3) or as f_eq(w) (kinetic). */
__device__ inline void store_dof(double *st, int scheme, const KinPar &kp,
				 int neq, long ipg, const double wv[6])
{
	if (!scheme) {
#pragma unroll
		for (int q = 0; q < 6; q++)
			st[ipg * 6 + q] = wv[q];
		return;
	}
	double qm[6][3];
	maxwell_qM(wv, qm);
	for (int iw = 0; iw < 6; iw++)
		for (int v = 0; v < kp.nv; v++)
			st[(long)(iw * kp.nv + v) * neq + ipg] =
				feq_comp(kp, wv, qm, iw, v);
}

/* n DoFs starting at ipg0 with coordinates x[3 k]; exact values from the
 * closed form of ip, or from hv[6 k] when hv != NULL */
__global__ void k_fill_exact(double *st, int scheme, KinPar kp, ImposedPar ip,
			     int neq, long ipg0, long n, const double *__restrict__ x,
			     const double *__restrict__ hv, double t)
{
	const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n)
		return;
	double wv[6];
	if (hv) {
		for (int q = 0; q < 6; q++)
			wv[q] = hv[6 * k + q];
	} else {
		const double xx[3] = { x[3 * k], x[3 * k + 1], x[3 * k + 2] };
		imposed(ip, xx, t, wv);
	}
	store_dof(st, scheme, kp, neq, ipg0 + k, wv);
}

extern "C" void gdh_init_solution(gdh_ctx *c, const double *wexact)
{
	const long n = 10L * c->ne;
	if (c->ip.kind == GDH_IMPOSED_HOST && !wexact) {
		fprintf(stderr, "gdon3d-hip: host imposed data missing\n");
		abort();
	}
	const double *hv = stage_hostvals(c, wexact, 6 * n);
	k_fill_exact<<<nblocks(n, 256), 256, 0, c->stream>>>(
		c->w[c->cur], c->scheme, c->kp, c->ip, c->neq, 0, n, c->xnode, hv,
		0.0);
	HIPCHECK(hipGetLastError());
}

extern "C" void gdh_update_boundary(gdh_ctx *c, double t, const double *wghost)
{
	const long n = 6L * c->nb;
	if (!n)
		return;
	if (c->ip.kind == GDH_IMPOSED_HOST && !wghost) {
		fprintf(stderr, "gdon3d-hip: host ghost data missing\n");
		abort();
	}
	const double *hv = stage_hostvals(c, wghost, 6 * n);
	prof_begin(c, 2);
	k_fill_exact<<<nblocks(n, 256), 256, 0, c->stream>>>(
		c->w[c->cur], c->scheme, c->kp, c->ip, c->neq, 10L * c->ne, n,
		c->xghost, hv, t);
	HIPCHECK(hipGetLastError());
	prof_end(c, 2);
}

/* ===================================================================== */
/*                           fused RK3 stage                             */
/* ===================================================================== */

#define RK3_EPB 32 /* elements per block; 10 threads per element */

/* Bijection from the dispatch order of blocks to tiles such that the blocks
 * the hardware sends to one XCD (block ids congruent mod 8) get a contiguous
 * range of tiles: face neighbours then mostly share that XCD's L2. */
__device__ inline int xcd_tile(int hw, int nblk)
{
	const int q = nblk / 8, r = nblk % 8, x = hw % 8;
	return x * q + min(x, r) + hw / 8;
}

/* lists of (face, face-local index) for each local node */
__constant__ static const int c_node_faces[10][3][2] = {
	{ { 0, 0 }, { 1, 0 }, { 2, 0 } }, /* 0 */
	{ { 1, 1 }, { 2, 2 }, { 3, 0 } }, /* 1 */
	{ { 0, 2 }, { 2, 1 }, { 3, 1 } }, /* 2 */
	{ { 0, 1 }, { 1, 2 }, { 3, 2 } }, /* 3 */
	{ { 1, 3 }, { 2, 5 }, { -1, -1 } }, /* 4 */
	{ { 2, 4 }, { 3, 3 }, { -1, -1 } }, /* 5 */
	{ { 0, 5 }, { 2, 3 }, { -1, -1 } }, /* 6 */
	{ { 0, 3 }, { 1, 5 }, { -1, -1 } }, /* 7 */
	{ { 1, 4 }, { 3, 5 }, { -1, -1 } }, /* 8 */
	{ { 0, 4 }, { 3, 4 }, { -1, -1 } }, /* 9 */
};

/*
 * One Williamson LSRK stage, fused: dw = M^-1 F w (volume + upwind faces),
 * k2 = A k2 - dt dw, w_out = w_in + B k2. Thread (element, node). The
 * volume term is sum_j A(c_ij) w_j with c_ij = codtau int(d phi_i phi_j)
 * (src/simulation/simulation.c:263-324), the face term lifts the upwind flux
 * with the face mass (src/simulation/simulation.c:356-491) and the mass
 * inverse is inv_mass_ref / det (src/simulation/simulation.c:218-261).
 * Ghost values of boundary faces are copied through to w_out.
 */
__global__ __launch_bounds__(RK3_EPB * 10) void k_rk3_stage(
	const double *__restrict__ w_in, double *__restrict__ w_out,
	double *__restrict__ k2, const double *__restrict__ geom,
	const int4 *__restrict__ nbr, const int4 *__restrict__ nbmap, int ne,
	double Acoef, double Bcoef, double dt, int first, int bflux)
{
	__shared__ double sW[RK3_EPB][60]; /* own states, later R = S Phi - V */
	__shared__ double sPhi[RK3_EPB][4][6][6];
	__shared__ double sGeo[RK3_EPB][10];
	__shared__ double sN[RK3_EPB][4][4]; /* scaled normal + 1/(|n|+1e-16) */
	__shared__ double sI[10][10][3];
	__shared__ double sMi[10][10];

	const int tid = threadIdx.x;
	const int nthr = RK3_EPB * 10;
	/* XCD-aware tiling: consecutive tiles on the same XCD (8 XCDs,
	 * round-robin dispatch) so that face neighbours share an L2 */
	const int tile = xcd_tile(blockIdx.x, gridDim.x);
	const int e0 = tile * RK3_EPB;

	for (int k = tid; k < 300; k += nthr)
		(&sI[0][0][0])[k] = (&c_internal[0][0][0])[k];
	for (int k = tid; k < 100; k += nthr)
		(&sMi[0][0])[k] = (&c_inv_mass[0][0])[k];
	const int nel = min(RK3_EPB, ne - e0);
	for (int k = tid; k < nel * 60; k += nthr)
		(&sW[0][0])[k] = w_in[(long)e0 * 60 + k];
	for (int k = tid; k < nel * 10; k += nthr)
		(&sGeo[0][0])[k] = geom[(long)e0 * 10 + k];
	__syncthreads();

	const int le = tid / 10, m = tid % 10;
	const int e = e0 + le;
	const bool active = le < nel;

	if (active && m < 4) {
		double vn[3];
		face_normal(sGeo[le], m, vn);
		sN[le][m][0] = vn[0];
		sN[le][m][1] = vn[1];
		sN[le][m][2] = vn[2];
		sN[le][m][3] = 1.0 / (sqrt(vn[0] * vn[0] + vn[1] * vn[1] +
					   vn[2] * vn[2]) + 1e-16);
	}
	__syncthreads();

	double V[6] = { 0, 0, 0, 0, 0, 0 };
	double wm[6];
	if (active) {
		/* volume: G_q = sum_j c_mj w_jq with c_mj = codtau int_ref */
		double g[6][3];
#pragma unroll
		for (int q = 0; q < 6; q++)
			g[q][0] = g[q][1] = g[q][2] = 0.0;
#pragma unroll
		for (int j = 0; j < 10; j++) {
			const double i0 = sI[m][j][0], i1 = sI[m][j][1], i2 = sI[m][j][2];
#pragma unroll
			for (int q = 0; q < 6; q++) {
				const double wj = sW[le][6 * j + q];
				g[q][0] += i0 * wj;
				g[q][1] += i1 * wj;
				g[q][2] += i2 * wj;
			}
		}
		const double *ct = sGeo[le];
		double G[6][3];
#pragma unroll
		for (int q = 0; q < 6; q++)
#pragma unroll
			for (int l = 0; l < 3; l++)
				G[q][l] = ct[3 * l + 0] * g[q][0] +
					  ct[3 * l + 1] * g[q][1] +
					  ct[3 * l + 2] * g[q][2];
		/* A(c) w = (-c x H, c x E) summed over j */
		V[0] = -(G[5][1] - G[4][2]);
		V[1] = -(G[3][2] - G[5][0]);
		V[2] = -(G[4][0] - G[3][1]);
		V[3] = G[2][1] - G[1][2];
		V[4] = G[0][2] - G[2][0];
		V[5] = G[1][0] - G[0][1];
#pragma unroll
		for (int q = 0; q < 6; q++)
			wm[q] = sW[le][6 * m + q];

		/* face fluxes at this node */
		const int4 nb4 = nbr[e], map4 = nbmap[e];
		const int nbv[4] = { nb4.x, nb4.y, nb4.z, nb4.w };
		const int mpv[4] = { map4.x, map4.y, map4.z, map4.w };
#pragma unroll
		for (int s = 0; s < 3; s++) {
			const int f = c_node_faces[m][s][0];
			const int a = c_node_faces[m][s][1];
			if (f < 0)
				continue;
			const int nb = nbv[f];
			double wr[6];
			const double *src;
			if (nb >= 0) {
				const int j = (mpv[f] >> (4 * a)) & 15;
				src = w_in + ((long)nb * 10 + j) * 6;
			} else {
				const long g = 10L * ne + 6L * (-1 - nb) + a;
				src = w_in + g * 6;
			}
#pragma unroll
			for (int q = 0; q < 6; q++)
				wr[q] = src[q];
			const double n[3] = { sN[le][f][0], sN[le][f][1], sN[le][f][2] };
			double fl[6];
			if (nb < 0 && bflux == GDH_FLUX_MAXWELL_METAL)
				maxwell_metal(wm, n, fl);
			else
				maxwell_upwind(wm, wr, n, sN[le][f][3], fl);
#pragma unroll
			for (int q = 0; q < 6; q++)
				sPhi[le][f][a][q] = fl[q];
			if (nb < 0) {
				/* carry the ghost DoF over to the output buffer */
				const long g = 10L * ne + 6L * (-1 - nb) + a;
#pragma unroll
				for (int q = 0; q < 6; q++)
					w_out[g * 6 + q] = wr[q];
			}
		}
	}
	__syncthreads();

	if (active) {
		/* lift: R_m = sum_faces S[a][b] Phi_b - V_m */
		double R[6];
#pragma unroll
		for (int q = 0; q < 6; q++)
			R[q] = -V[q];
#pragma unroll
		for (int s = 0; s < 3; s++) {
			const int f = c_node_faces[m][s][0];
			const int a = c_node_faces[m][s][1];
			if (f < 0)
				continue;
#pragma unroll
			for (int b = 0; b < 6; b++) {
				const double sab = c_face_scal[a][b];
				if (sab == 0.0)
					continue;
#pragma unroll
				for (int q = 0; q < 6; q++)
					R[q] += sab * sPhi[le][f][b][q];
			}
		}
		/* every read of sW (volume term) happened before the last
		 * barrier: reuse it for R */
#pragma unroll
		for (int q = 0; q < 6; q++)
			sW[le][6 * m + q] = R[q];
	}
	__syncthreads();

	if (active) {
		const double invdet = 1.0 / sGeo[le][9];
		double dw[6] = { 0, 0, 0, 0, 0, 0 };
#pragma unroll
		for (int j = 0; j < 10; j++) {
			const double mij = sMi[m][j];
#pragma unroll
			for (int q = 0; q < 6; q++)
				dw[q] += mij * sW[le][6 * j + q];
		}
		const long base = ((long)e * 10 + m) * 6;
#pragma unroll
		for (int q = 0; q < 6; q++) {
			const double d = dw[q] * invdet;
			const double kk = first ? -dt * d : Acoef * k2[base + q] - dt * d;
			k2[base + q] = kk;
			w_out[base + q] = wm[q] + Bcoef * kk;
		}
	}
}

extern "C" void gdh_rk3_stage(gdh_ctx *c, int stage, double dt)
{
	static const double A[3] = { 0., -5. / 9., -153. / 128. };
	static const double B[3] = { 1. / 3., 15. / 16., 8. / 15. };
	const int nblk = nblocks(c->ne, RK3_EPB);
	prof_begin(c, 0);
	c->prof_launches++;
	k_rk3_stage<<<nblk, RK3_EPB * 10, 0, c->stream>>>(
		c->w[c->cur], c->w[1 - c->cur], c->k2, c->geom, c->nbr, c->nbmap,
		c->ne, A[stage], B[stage], dt, stage == 0, c->md.bflux_kind);
	HIPCHECK(hipGetLastError());
	prof_end(c, 0);
	c->cur = 1 - c->cur;
}

/* ===================================================================== */
/*                      kinetic relaxation scheme                        */
/* ===================================================================== */

#define THETA 0.5 /* Crank-Nicolson, src/dg/relaxation.c:133 */

/* own outward flux coefficient s_f = v . vn(e, f) */
__device__ inline double face_coef(const double ct[9], int f, const double *v)
{
	double vn[3];
	face_normal(ct, f, vn);
	return v[0] * vn[0] + v[1] * vn[1] + v[2] * vn[2];
}

/*
 * Diagonal block of M1 = I + theta dt M^-1 F_v for item (v, e), inverted by
 * Gauss-Jordan elimination with partial pivoting. F_v restricted to the
 * element: volume -v.c_ij (src/simulation/simulation.c:298-308) plus, on
 * outflow faces, s_f S (src/simulation/simulation.c:531-562).
 */
__global__ void k_block_inverse(const int2 *__restrict__ items, int nitems,
				const double *__restrict__ geom,
				const uint8_t *__restrict__ cls, KinPar kp,
				double dt, double *__restrict__ dinv)
{
	const int it = blockIdx.x * blockDim.x + threadIdx.x;
	if (it >= nitems)
		return;
	const int v = items[it].x, e = items[it].y;
	double ct[9];
	for (int k = 0; k < 9; k++)
		ct[k] = geom[10L * e + k];
	const double det = geom[10L * e + 9];
	const double *vv = kp.vi + 3 * v;
	double a[3];
	for (int k = 0; k < 3; k++)
		a[k] = ct[0 + k] * vv[0] + ct[3 + k] * vv[1] + ct[6 + k] * vv[2];

	double F[10][10];
	for (int i = 0; i < 10; i++)
		for (int j = 0; j < 10; j++)
			F[i][j] = -(a[0] * c_internal[i][j][0] +
				    a[1] * c_internal[i][j][1] +
				    a[2] * c_internal[i][j][2]);
	for (int f = 0; f < 4; f++) {
		if (cls[(4L * e + f) * kp.nv + v] != 1)
			continue;
		const double s = face_coef(ct, f, vv);
		for (int p = 0; p < 6; p++)
			for (int q = 0; q < 6; q++)
				F[c_face2node[f][p]][c_face2node[f][q]] +=
					s * c_face_scal[p][q];
	}
	const double tau = THETA * dt / det;
	double A[10][20];
	for (int i = 0; i < 10; i++) {
		for (int j = 0; j < 10; j++) {
			double s = 0.0;
			for (int k = 0; k < 10; k++)
				s += c_inv_mass[i][k] * F[k][j];
			A[i][j] = (i == j ? 1.0 : 0.0) + tau * s;
			A[i][10 + j] = i == j ? 1.0 : 0.0;
		}
	}
	for (int c = 0; c < 10; c++) {
		int p = c;
		double best = fabs(A[c][c]);
		for (int r = c + 1; r < 10; r++)
			if (fabs(A[r][c]) > best) {
				best = fabs(A[r][c]);
				p = r;
			}
		if (p != c)
			for (int k = 0; k < 20; k++) {
				const double t = A[c][k];
				A[c][k] = A[p][k];
				A[p][k] = t;
			}
		const double inv = 1.0 / A[c][c];
		for (int k = 0; k < 20; k++)
			A[c][k] *= inv;
		for (int r = 0; r < 10; r++) {
			if (r == c)
				continue;
			const double l = A[r][c];
			if (l == 0.0)
				continue;
			for (int k = 0; k < 20; k++)
				A[r][k] -= l * A[c][k];
		}
	}
	double *out = dinv + 100L * it;
	for (int i = 0; i < 10; i++)
		for (int j = 0; j < 10; j++)
			out[10 * i + j] = A[i][10 + j];
}

extern "C" void gdh_relax_setup(gdh_ctx *c, double dt, const uint8_t *cls,
				const int *items_ve, const int *level_ptr,
				int nlevels)
{
	if (!c->scheme) {
		fprintf(stderr, "gdon3d-hip: relaxation set-up on an RK3 context\n");
		abort();
	}
	const int ne = c->ne, nv = c->kp.nv;
	c->dt = dt;
	c->nlevels = nlevels;
	c->level_ptr.assign(level_ptr, level_ptr + nlevels + 1);
	c->nitems = level_ptr[nlevels];
	if (c->nitems != nv * ne) {
		fprintf(stderr, "gdon3d-hip: sweep schedule has %d items, expected %d\n",
			c->nitems, nv * ne);
		abort();
	}
	std::vector<int2> items(c->nitems);
	for (int k = 0; k < c->nitems; k++) {
		const int ve = items_ve[k];
		if (ve < 0 || ve >= nv * ne) {
			fprintf(stderr, "gdon3d-hip: invalid sweep item %d\n", ve);
			abort();
		}
		items[k] = make_int2(ve / ne, ve % ne);
	}
	c->items = (int2 *)dmalloc(sizeof(int2) * items.size());
	HIPCHECK(hipMemcpy(c->items, items.data(), sizeof(int2) * items.size(),
			   hipMemcpyHostToDevice));
	c->cls = (uint8_t *)dmalloc((size_t)ne * 4 * nv);
	HIPCHECK(hipMemcpy(c->cls, cls, (size_t)ne * 4 * nv,
			   hipMemcpyHostToDevice));
	c->dinv = (double *)dmalloc(sizeof(double) * 100 * items.size());
	k_block_inverse<<<nblocks(c->nitems, 64), 64, 0, c->stream>>>(
		c->items, c->nitems, c->geom, c->cls, c->kp, dt, c->dinv);
	HIPCHECK(hipGetLastError());
}

/*
 * One upwind level of the transport solve sol = M1^-1 (M2 f): one item
 * (velocity v, element e) per 64-lane block, lane (iw, i) = (field, node).
 *   y = (1-theta) F_ee f_e + sum_inflow s_f S ((1-theta) f_nb + theta sol_nb)
 *   rhs = f_e - dt/det Minv_ref y,   sol_e = D^-1 rhs
 * (src/dg/relaxation.c:93-110: M2 = I - (1-theta) dt M^-1 F,
 *  M1 = I + theta dt M^-1 F, block-lower-triangular in level order).
 * Ghost DoFs are identity rows of both matrices: their sol is f.
 */
__global__ __launch_bounds__(64) void k_sweep_level(
	const int2 *__restrict__ items, int begin, const double *__restrict__ fo,
	double *__restrict__ fn, const double *__restrict__ geom,
	const int4 *__restrict__ nbr, const int4 *__restrict__ nbmap,
	const uint8_t *__restrict__ cls, const double *__restrict__ dinv,
	KinPar kp, int ne, int neq, double dt)
{
	__shared__ double sF[6][10], sY[6][10], sR[6][10];
	const int it = begin + blockIdx.x;
	const int v = items[it].x, e = items[it].y;
	const int t = threadIdx.x;
	const int iw = t / 10, i = t % 10;
	const bool lane = t < 60;
	const long base = (long)(iw * kp.nv + v) * neq;

	if (lane)
		sF[iw][i] = fo[base + 10L * e + i];
	double ct[9];
	for (int k = 0; k < 9; k++)
		ct[k] = geom[10L * e + k];
	const double det = geom[10L * e + 9];
	const double *vv = kp.vi + 3 * v;
	__syncthreads();

	if (lane) {
		double a[3];
		for (int k = 0; k < 3; k++)
			a[k] = ct[0 + k] * vv[0] + ct[3 + k] * vv[1] + ct[6 + k] * vv[2];
		double own = 0.0;
		for (int j = 0; j < 10; j++)
			own -= (a[0] * c_internal[i][j][0] + a[1] * c_internal[i][j][1] +
				a[2] * c_internal[i][j][2]) *
			       sF[iw][j];
		double inflow = 0.0;
		const int4 n4 = nbr[e], m4 = nbmap[e];
		const int nbv[4] = { n4.x, n4.y, n4.z, n4.w };
		const int mpv[4] = { m4.x, m4.y, m4.z, m4.w };
		for (int f = 0; f < 4; f++) {
			int ai = -1;
			for (int p = 0; p < 6; p++)
				if (c_face2node[f][p] == i)
					ai = p;
			if (ai < 0)
				continue;
			const uint8_t cl = cls[(4L * e + f) * kp.nv + v];
			if (!cl)
				continue;
			const double s = face_coef(ct, f, vv);
			if (cl == 1) {
				for (int b = 0; b < 6; b++)
					own += s * c_face_scal[ai][b] *
					       sF[iw][c_face2node[f][b]];
			} else {
				const int nb = nbv[f];
				for (int b = 0; b < 6; b++) {
					double u;
					if (nb >= 0) {
						const int j = (mpv[f] >> (4 * b)) & 15;
						const long g = base + 10L * nb + j;
						u = (1.0 - THETA) * fo[g] + THETA * fn[g];
					} else {
						u = fo[base + 10L * ne + 6L * (-1 - nb) + b];
					}
					inflow += s * c_face_scal[ai][b] * u;
				}
			}
		}
		sY[iw][i] = (1.0 - THETA) * own + inflow;
	}
	__syncthreads();
	if (lane) {
		double s = 0.0;
		for (int mm = 0; mm < 10; mm++)
			s += c_inv_mass[i][mm] * sY[iw][mm];
		sR[iw][i] = sF[iw][i] - dt / det * s;
	}
	__syncthreads();
	if (lane) {
		const double *D = dinv + 100L * it + 10 * i;
		double s = 0.0;
		for (int j = 0; j < 10; j++)
			s += D[j] * sR[iw][j];
		fn[base + 10L * e + i] = s;
	}
}

/* f <- omega f_eq(w(f)) + (1 - omega) f at the element DoFs
 * (src/dg/dg.c:109-132, src/simulation/simulation.c:42-54) */
__global__ void k_relax(double *__restrict__ f, KinPar kp, long n, int neq)
{
	const long ipg = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (ipg >= n)
		return;
	const int nv = kp.nv;
	double fo[6 * GDH_MAX_V];
	double w[6];
	for (int iw = 0; iw < 6; iw++) {
		double s = 0.0;
		for (int v = 0; v < nv; v++) {
			const double x = f[(long)(iw * nv + v) * neq + ipg];
			fo[iw * nv + v] = x;
			s += x;
		}
		w[iw] = s;
	}
	double q[6][3];
	maxwell_qM(w, q);
	const double om = kp.omega;
	for (int iw = 0; iw < 6; iw++)
		for (int v = 0; v < nv; v++)
			f[(long)(iw * nv + v) * neq + ipg] =
				om * feq_comp(kp, w, q, iw, v) +
				(1. - om) * fo[iw * nv + v];
}

extern "C" void gdh_relax_step(gdh_ctx *c)
{
	double *fo = c->w[c->cur], *fn = c->w[1 - c->cur];
	/* Ghost DoFs are identity rows of M1 and M2; the sweep reads them from
	 * fo and nothing reads the ghost region of fn before the next ghost
	 * update overwrites it. */
	prof_begin(c, 0);
	for (int L = 0; L < c->nlevels; L++) {
		const int b = c->level_ptr[L], n = c->level_ptr[L + 1] - b;
		if (!n)
			continue;
		c->prof_launches++;
		k_sweep_level<<<n, 64, 0, c->stream>>>(
			c->items, b, fo, fn, c->geom, c->nbr, c->nbmap, c->cls,
			c->dinv, c->kp, c->ne, c->neq, c->dt);
	}
	HIPCHECK(hipGetLastError());
	prof_end(c, 0);
	const long n = 10L * c->ne;
	prof_begin(c, 1);
	k_relax<<<nblocks(n, 256), 256, 0, c->stream>>>(fn, c->kp, n, c->neq);
	HIPCHECK(hipGetLastError());
	prof_end(c, 1);
	c->cur = 1 - c->cur;
}

/* relaxation of every DoF, ghosts included (src/dg/dg.c:109-147) */
extern "C" void gdh_apply_relaxation(gdh_ctx *c)
{
	const long n = c->neq;
	k_relax<<<nblocks(n, 256), 256, 0, c->stream>>>(c->w[c->cur], c->kp, n,
							    c->neq);
	HIPCHECK(hipGetLastError());
}

/* ===================================================================== */
/*                              L2 error                                 */
/* ===================================================================== */

#define ERR_BS 256

/* per-block partial sums of (w - w_exact)^2 wpg_i det_e per component
 * (src/simulation/simulation.c:624-650) */
__global__ __launch_bounds__(ERR_BS) void k_err_partial(
	const double *__restrict__ st, int scheme, KinPar kp, ImposedPar ip,
	const double *__restrict__ geom, const double *__restrict__ xnode,
	const double *__restrict__ hv, int ne, int neq, double t,
	double *__restrict__ partials)
{
	__shared__ double red[6][ERR_BS];
	double acc[6] = { 0, 0, 0, 0, 0, 0 };
	for (int e = blockIdx.x * ERR_BS + threadIdx.x; e < ne;
	     e += gridDim.x * ERR_BS) {
		const double det = geom[10L * e + 9];
		for (int i = 0; i < 10; i++) {
			const long ipg = 10L * e + i;
			double wx[6], wl[6];
			if (hv) {
				for (int q = 0; q < 6; q++)
					wx[q] = hv[6 * ipg + q];
			} else {
				const double *x = xnode + 3 * ipg;
				const double xx[3] = { x[0], x[1], x[2] };
				imposed(ip, xx, t, wx);
			}
			if (!scheme) {
				for (int q = 0; q < 6; q++)
					wl[q] = st[6 * ipg + q];
			} else {
				for (int q = 0; q < 6; q++) {
					double s = 0.0;
					for (int v = 0; v < kp.nv; v++)
						s += st[(long)(q * kp.nv + v) * neq + ipg];
					wl[q] = s;
				}
			}
			for (int q = 0; q < 6; q++) {
				const double d = wl[q] - wx[q];
				acc[q] += d * d * c_wpg[i] * det;
			}
		}
	}
	for (int q = 0; q < 6; q++)
		red[q][threadIdx.x] = acc[q];
	__syncthreads();
	for (int s = ERR_BS / 2; s > 0; s >>= 1) {
		if (threadIdx.x < s)
			for (int q = 0; q < 6; q++)
				red[q][threadIdx.x] += red[q][threadIdx.x + s];
		__syncthreads();
	}
	if (threadIdx.x < 6)
		partials[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

/* fixed-order final sum: mean over components of sqrt|sum| */
__global__ void k_err_final(const double *__restrict__ partials, int np,
			    double *__restrict__ out)
{
	if (threadIdx.x != 0 || blockIdx.x != 0)
		return;
	double err = 0.0;
	for (int q = 0; q < 6; q++) {
		double s = 0.0;
		for (int b = 0; b < np; b++)
			s += partials[6 * b + q];
		err += sqrt(fabs(s)) / 6;
	}
	out[0] = err;
}

extern "C" double gdh_error_l2(gdh_ctx *c, double t, const double *wexact)
{
	if (c->ip.kind == GDH_IMPOSED_HOST && !wexact) {
		fprintf(stderr, "gdon3d-hip: host exact values missing\n");
		abort();
	}
	const double *hv = stage_hostvals(c, wexact, 60L * c->ne);
	const int np = std::min(c->npartials, std::max(1, nblocks(c->ne, ERR_BS)));
	k_err_partial<<<np, ERR_BS, 0, c->stream>>>(
		c->w[c->cur], c->scheme, c->kp, c->ip, c->geom, c->xnode, hv,
		c->ne, c->neq, t, c->partials);
	HIPCHECK(hipGetLastError());
	k_err_final<<<1, 64, 0, c->stream>>>(c->partials, np,
					     c->partials + 6L * c->npartials);
	HIPCHECK(hipGetLastError());
	double err = 0;
	HIPCHECK(hipMemcpyAsync(&err, c->partials + 6L * c->npartials,
				sizeof(double), hipMemcpyDeviceToHost, c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
	return err;
}

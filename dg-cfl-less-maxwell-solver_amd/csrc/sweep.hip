/*
 * Kinetic relaxation scheme on the GPU: per-element block factors of
 * M1 = I + dt/2 M^-1 F_v, the upwind transport sweep sol = M1^-1 (M2 f), and
 * the node-local relaxation.
 */
#include "ctx.h"

using namespace gdh;

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
	sweep_setup(c);
}

/* ------------------------------------------------------------------ */
/* Cross-workgroup handoff inside one launch (persistent sweep): values   */
/* produced by another workgroup are read and written with relaxed       */
/* agent-scope atomics, which go through the per-XCD L2 to the coherent   */
/* memory side, and ordered with workgroup-scope fences plus a wait for   */
/* outstanding memory operations (no L2 flush or invalidate).             */
/* ------------------------------------------------------------------ */
typedef unsigned long long u64;

__device__ inline double coherent_load(const double *p)
{
	const u64 v = __hip_atomic_load((const u64 *)p, __ATOMIC_RELAXED,
					__HIP_MEMORY_SCOPE_AGENT);
	return __longlong_as_double((long long)v);
}

__device__ inline void coherent_store(double *p, double x)
{
	__hip_atomic_store((u64 *)p, (u64)__double_as_longlong(x), __ATOMIC_RELAXED,
			   __HIP_MEMORY_SCOPE_AGENT);
}

/* wait for all outstanding vector memory operations of the wave */
__device__ inline void vmem_wait()
{
	__builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8));
}

#define SWEEP_SPIN_LIMIT (1 << 20) /* polls before a wait is declared stuck */

/* Operator tables, copied to LDS once per block. */
struct SweepTables {
	double minv[10][10];
	double S[6][6];
};

/* One prepared item: everything that does not depend on the inflow
 * neighbours' new values. */
struct ItemBuf {
	double base[6][10];
	double G[4][10][6]; /* theta G_f for the interior inflow faces */
	int v, e;
	int nb[4], map[4], act[4]; /* act: interior inflow face to wait on */
};

/* Work area of prepare(). */
struct PrepScratch {
	double F[6][10];
	double D[100];
	double H[10][10];
	double U[4][6][6];
	double Gf[4][10][6];
};

/*
 * Transport of item (v, e), split around the dependency wait. With
 * D = I + theta dt/det Minv F_ee(v) (its inverse stored per item), theta = 1/2
 * and u_f the inflow trace ((1-theta) f_nb + theta sol_nb, or the ghost
 * value), one block row of M1 sol = M2 f (src/dg/relaxation.c:64-110) reads
 *   sol = D^-1 [ f - dt/det Minv ((1-theta) F_ee f + sum_f s_f L_f S u_f) ]
 *       = 2 D^-1 f - f - sum_f G_f u_f,  G_f = s_f (dt/det) D^-1 Minv L_f S
 * since (dt/det) Minv (1-theta) F_ee = D - I for theta = 1/2. prepare()
 * forms 2 D^-1 f - f - sum_f G_f ((1-theta) f_nb or ghost) and theta G_f;
 * after the wait, finish() subtracts theta G_f sol_nb. Lanes: (iw, i) =
 * (field, node) for the vectors, (i, b) for the 10x6 matrices G_f.
 */
__device__ inline void sweep_prepare(int it, const int2 *__restrict__ items,
				     const double *__restrict__ fo,
				     const double *__restrict__ geom,
				     const int4 *__restrict__ nbr,
				     const int4 *__restrict__ nbmap,
				     const uint8_t *__restrict__ cls,
				     const double *__restrict__ dinv,
				     const KinPar &kp, int ne, int neq, double dt,
				     const SweepTables &T, ItemBuf &B, PrepScratch &P)
{
	const int2 ve = items[it];
	const int v = ve.x, e = ve.y;
	const int t = threadIdx.x;
	const int iw = t / 10, i = t % 10;
	const int nv = kp.nv;

	if (t < 60)
		P.F[iw][i] = fo[(long)(iw * nv + v) * neq + 10L * e + i];
	for (int k = t; k < 100; k += 64)
		P.D[k] = dinv[100L * it + k];
	double ct[9];
#pragma unroll
	for (int k = 0; k < 9; k++)
		ct[k] = geom[10L * e + k];
	const double det = geom[10L * e + 9];
	const int4 n4 = nbr[e], m4 = nbmap[e];
	const int nbv[4] = { n4.x, n4.y, n4.z, n4.w };
	const int mpv[4] = { m4.x, m4.y, m4.z, m4.w };
	bool in[4];
#pragma unroll
	for (int f = 0; f < 4; f++)
		in[f] = cls[(4L * e + f) * nv + v] == 2;
	/* inflow traces from the previous state (or the ghosts) */
	if (t < 36) {
		const int q = t / 6, b = t % 6;
		const long bq = (long)(q * nv + v) * neq;
#pragma unroll
		for (int f = 0; f < 4; f++) {
			if (!in[f])
				continue;
			const int nb = nbv[f];
			P.U[f][q][b] = nb >= 0 ?
				(1.0 - THETA) * fo[bq + 10L * nb + ((mpv[f] >> (4 * b)) & 15)] :
				fo[bq + 10L * ne + 6L * (-1 - nb) + b];
		}
	}
	if (t == 0) {
		B.v = v;
		B.e = e;
#pragma unroll
		for (int f = 0; f < 4; f++) {
			B.nb[f] = nbv[f];
			B.map[f] = mpv[f];
			B.act[f] = in[f] && nbv[f] >= 0;
		}
	}
	__syncthreads();

	/* 2 D^-1 f - f, and H = dt/det D^-1 Minv */
	double base = 0.0;
	if (t < 60) {
		double s = 0.0;
#pragma unroll
		for (int j = 0; j < 10; j++)
			s += P.D[10 * i + j] * P.F[iw][j];
		base = 2.0 * s - P.F[iw][i];
	}
	const double sc = dt / det;
	for (int k = t; k < 100; k += 64) {
		const int r = k / 10, c = k % 10;
		double s = 0.0;
#pragma unroll
		for (int l = 0; l < 10; l++)
			s += P.D[10 * r + l] * T.minv[l][c];
		P.H[r][c] = sc * s;
	}
	__syncthreads();

	/* G_f = s_f H[:, face f] S for each inflow face */
	if (t < 60) {
		const int r = t / 6, b = t % 6;
		const double *vv = kp.vi + 3 * v;
#pragma unroll
		for (int f = 0; f < 4; f++) {
			if (!in[f])
				continue;
			const double sf = face_coef(ct, f, vv);
			double s = 0.0;
#pragma unroll
			for (int a = 0; a < 6; a++)
				s += P.H[r][face_node(f, a)] * T.S[a][b];
			const double g = sf * s;
			P.Gf[f][r][b] = g;
			B.G[f][r][b] = THETA * g;
		}
	}
	__syncthreads();

	if (t < 60) {
#pragma unroll
		for (int f = 0; f < 4; f++) {
			if (!in[f])
				continue;
#pragma unroll
			for (int b = 0; b < 6; b++)
				base -= P.Gf[f][i][b] * P.U[f][iw][b];
		}
		B.base[iw][i] = base;
	}
	__syncthreads();
}

/* true when every interior inflow neighbour of the prepared item is done */
__device__ inline bool sweep_ready(const ItemBuf &B, const int *done, int ne,
				   int epoch)
{
	const int t = threadIdx.x;
	bool wait = false;
	if (t < 4 && B.act[t])
		wait = __hip_atomic_load(done + (long)B.v * ne + B.nb[t],
					 __ATOMIC_RELAXED,
					 __HIP_MEMORY_SCOPE_AGENT) != epoch;
	return __ballot(wait) == 0;
}

template <bool PERSIST>
__device__ inline void sweep_finish(const ItemBuf &B, double *fn,
				    const KinPar &kp, int ne, int neq, int *done,
				    int epoch, int *status)
{
	const int t = threadIdx.x;
	const int iw = t / 10, i = t % 10;
	const int v = B.v, e = B.e;
	if (PERSIST) {
		if (t < 4 && B.act[t]) {
			const int *flag = done + (long)v * ne + B.nb[t];
			int spins = 0;
			while (__hip_atomic_load(flag, __ATOMIC_RELAXED,
						 __HIP_MEMORY_SCOPE_AGENT) != epoch) {
				/* bounded: after a timeout anywhere every wait ends */
				if ((++spins & 255) == 0 &&
				    (spins > SWEEP_SPIN_LIMIT ||
				     __hip_atomic_load(status, __ATOMIC_RELAXED,
						       __HIP_MEMORY_SCOPE_AGENT))) {
					__hip_atomic_store(status, 1, __ATOMIC_RELAXED,
							   __HIP_MEMORY_SCOPE_AGENT);
					break;
				}
				__builtin_amdgcn_s_sleep(1);
			}
		}
		vmem_wait();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
	}
	if (t < 60) {
		const long bq = (long)(iw * kp.nv + v) * neq;
		double s = B.base[iw][i];
#pragma unroll
		for (int f = 0; f < 4; f++) {
			if (!B.act[f])
				continue;
			const long g0 = bq + 10L * B.nb[f];
			const int map = B.map[f];
#pragma unroll
			for (int b = 0; b < 6; b++) {
				const long g = g0 + ((map >> (4 * b)) & 15);
				const double x = PERSIST ? coherent_load(fn + g) : fn[g];
				s -= B.G[f][i][b] * x;
			}
		}
		const long g = bq + 10L * e + i;
		if (PERSIST)
			coherent_store(fn + g, s);
		else
			fn[g] = s;
	}
	if (PERSIST) {
		/* every lane's store is complete before the flag is raised */
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
		vmem_wait();
		if (t == 0)
			__hip_atomic_store(done + (long)v * ne + e, epoch, __ATOMIC_RELAXED,
					   __HIP_MEMORY_SCOPE_AGENT);
	}
	__syncthreads();
}

__device__ inline void load_tables(SweepTables &T)
{
	for (int k = threadIdx.x; k < 100; k += blockDim.x)
		(&T.minv[0][0])[k] = (&c_inv_mass[0][0])[k];
	for (int k = threadIdx.x; k < 36; k += blockDim.x)
		(&T.S[0][0])[k] = (&c_face_scal[0][0])[k];
	__syncthreads();
}

/* One upwind level per launch, one item per 64-lane block. */
__global__ __launch_bounds__(64) void k_sweep_level(
	const int2 *__restrict__ items, int begin, const double *__restrict__ fo,
	double *__restrict__ fn, const double *__restrict__ geom,
	const int4 *__restrict__ nbr, const int4 *__restrict__ nbmap,
	const uint8_t *__restrict__ cls, const double *__restrict__ dinv,
	KinPar kp, int ne, int neq, double dt)
{
	__shared__ SweepTables T;
	__shared__ ItemBuf B;
	__shared__ PrepScratch P;
	load_tables(T);
	sweep_prepare(begin + blockIdx.x, items, fo, geom, nbr, nbmap, cls, dinv, kp,
		      ne, neq, dt, T, B, P);
	sweep_finish<false>(B, fn, kp, ne, neq, nullptr, 0, nullptr);
}

/*
 * The whole sweep in one launch. Items are in level order; block b takes
 * items b, b + G, b + 2G, ... and waits on the done flags of its inflow
 * neighbours (always earlier items). All G blocks are resident (cooperative
 * launch), so the earliest unfinished item always has its inputs ready and
 * the sweep cannot deadlock; a bounded spin reports a timeout in *status
 * instead of hanging. While an item's inputs are not ready, the block
 * prepares its next item.
 */
__global__ __launch_bounds__(64) void k_sweep_persistent(
	const int2 *__restrict__ items, int nitems, const double *__restrict__ fo,
	double *fn, const double *__restrict__ geom, const int4 *__restrict__ nbr,
	const int4 *__restrict__ nbmap, const uint8_t *__restrict__ cls,
	const double *__restrict__ dinv, KinPar kp, int ne, int neq, double dt,
	int *done, int epoch, int *status)
{
	__shared__ SweepTables T;
	__shared__ ItemBuf B[2];
	__shared__ PrepScratch P;
	load_tables(T);
	int it = blockIdx.x, k = 0;
	if (it < nitems)
		sweep_prepare(it, items, fo, geom, nbr, nbmap, cls, dinv, kp, ne, neq,
			      dt, T, B[0], P);
	for (; it < nitems; it += gridDim.x, k ^= 1) {
		const int next = it + gridDim.x;
		bool prepared = false;
		if (next < nitems && !sweep_ready(B[k], done, ne, epoch)) {
			sweep_prepare(next, items, fo, geom, nbr, nbmap, cls, dinv, kp,
				      ne, neq, dt, T, B[k ^ 1], P);
			prepared = true;
		}
		sweep_finish<true>(B[k], fn, kp, ne, neq, done, epoch, status);
		if (next < nitems && !prepared)
			sweep_prepare(next, items, fo, geom, nbr, nbmap, cls, dinv, kp,
				      ne, neq, dt, T, B[k ^ 1], P);
	}
}

void sweep_setup(gdh_ctx *c)
{
	c->done = (int *)dmalloc(sizeof(int) * (size_t)c->nitems);
	HIPCHECK(hipMemset(c->done, 0, sizeof(int) * (size_t)c->nitems));
	c->status = (int *)dmalloc(sizeof(int));
	HIPCHECK(hipMemset(c->status, 0, sizeof(int)));
	c->epoch = 0;
	int dev = 0, ncu = 0, per_cu = 0;
	HIPCHECK(hipGetDevice(&dev));
	HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
	HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sweep_persistent,
							      64, 0));
	c->sweep_grid = std::max(1, per_cu * ncu);
	const char *mode = getenv("GDN_SWEEP");
	c->sweep_mode = (mode && !strcmp(mode, "level")) ? 0 : 1;
}

void sweep_step(gdh_ctx *c, double *fo, double *fn)
{
	if (c->sweep_mode == 0) {
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
		return;
	}
	c->epoch++;
	c->prof_launches++;
	const int2 *items = c->items;
	int nitems = c->nitems, ne = c->ne, neq = c->neq, epoch = c->epoch;
	double dt = c->dt;
	KinPar kp = c->kp;
	const double *geom = c->geom, *dinv = c->dinv;
	const int4 *nbr = c->nbr, *nbmap = c->nbmap;
	const uint8_t *cls = c->cls;
	int *done = c->done, *status = c->status;
	void *args[] = { &items, &nitems, &fo, &fn, &geom, &nbr, &nbmap, &cls,
			 &dinv, &kp, &ne, &neq, &dt, &done, &epoch, &status };
	HIPCHECK(hipLaunchCooperativeKernel((const void *)k_sweep_persistent,
					    dim3(std::min(c->sweep_grid, nitems)),
					    dim3(64), args, 0, c->stream));
}

/* f <- omega f_eq(w(f)) + (1 - omega) f at DoFs [0, n)
 * (src/dg/dg.c:109-132, src/simulation/simulation.c:42-54); NV = kp.nv,
 * fixed at compile time so that the node's 6 NV values stay in registers. */
template <int NV>
__global__ __launch_bounds__(256) void k_relax(double *__restrict__ f, KinPar kp,
					       long n, int neq)
{
	const long ipg = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (ipg >= n)
		return;
	double fo[6][NV];
	double w[6];
#pragma unroll
	for (int iw = 0; iw < 6; iw++) {
		double s = 0.0;
#pragma unroll
		for (int v = 0; v < NV; v++) {
			const double x = f[(long)(iw * NV + v) * neq + ipg];
			fo[iw][v] = x;
			s += x;
		}
		w[iw] = s;
	}
	double q[6][3];
	maxwell_qM(w, q);
	const double om = kp.omega;
#pragma unroll
	for (int iw = 0; iw < 6; iw++)
#pragma unroll
		for (int v = 0; v < NV; v++)
			f[(long)(iw * NV + v) * neq + ipg] =
				om * feq_comp(kp, w, q, iw, v) + (1. - om) * fo[iw][v];
}

void relax_launch(gdh_ctx *c, double *f, long n)
{
	const int nb = nblocks(n, 256);
	if (c->kp.nv == 4)
		k_relax<4><<<nb, 256, 0, c->stream>>>(f, c->kp, n, c->neq);
	else if (c->kp.nv == 6)
		k_relax<6><<<nb, 256, 0, c->stream>>>(f, c->kp, n, c->neq);
	else {
		fprintf(stderr, "gdon3d-hip: no relaxation kernel for %d velocities\n",
			c->kp.nv);
		abort();
	}
	HIPCHECK(hipGetLastError());
}

extern "C" void gdh_relax_step(gdh_ctx *c)
{
	double *fo = c->w[c->cur], *fn = c->w[1 - c->cur];
	/* Ghost DoFs are identity rows of M1 and M2; the sweep reads them from
	 * fo and nothing reads the ghost region of fn before the next ghost
	 * update overwrites it. */
	prof_begin(c, 0);
	sweep_step(c, fo, fn);
	prof_end(c, 0);
	const long n = 10L * c->ne;
	prof_begin(c, 1);
	relax_launch(c, fn, n);
	prof_end(c, 1);
	c->cur = 1 - c->cur;
}

/* relaxation of every DoF, ghosts included (src/dg/dg.c:109-147) */
extern "C" void gdh_apply_relaxation(gdh_ctx *c)
{
	const long n = c->neq;
	relax_launch(c, c->w[c->cur], n);
}


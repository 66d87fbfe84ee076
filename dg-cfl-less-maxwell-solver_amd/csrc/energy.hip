/*
 * Discrete electromagnetic energy of a state: E(w) = 1/2 w^T M w with M the
 * assembled P2 mass matrix (src/simulation/simulation.c:218-261: per element
 * det_e times the reference mass of src/mesh/t10.c), summed over the element
 * DoFs, for the macro fields w (RK3) or w = sum_v f_v (kinetic scheme,
 * src/models/d3q4.c:82-95). The reference reports no energy; it is the
 * quantity Maxwell conserves and the upwind scheme dissipates, reported per
 * step by the benchmark's energy history.
 *
 * Two passes with a fixed reduction order (per-block partial sums, then one
 * 64-lane block), so a result does not depend on scheduling.
 */
#include <pthread.h>

#include "ctx.h"

#define EN_BS 256
#define EN_MAXB 1024 /* partial sums: at most this many blocks */

/* the reference layout of simu->wn on the device: RK3 wn[6 ipg + q]
 * (nv = 0), kinetic wn[(q nv + v) neq + ipg] */
struct EnergyRef {
	const double *wn;
	int nv;
	long neq;
	__device__ double weight_e(long) const
	{
		return 1.0;
	}
	__device__ double w(long ipg, int q) const
	{
		if (!nv)
			return wn[6 * ipg + q];
		double s = 0.0;
		for (int v = 0; v < nv; v++)
			s += wn[((long)q * nv + v) * neq + ipg];
		return s;
	}
};

/* the device state of a context: RK3 in storage columns (macro_ptr),
 * kinetic item-major (kin_at) */
template <typename T> struct EnergyDev {
	const T *kin;
	StateRef<T> sr;
	int nv; /* 0: RK3 */
	long neq;
	const double *ieps; /* 1/eps (kinetic: by element; RK3: by column) or null */
	__device__ double w(long ipg, int q) const
	{
		if (!nv)
			return (double)*macro_ptr(sr, ipg, q);
		double s = 0.0;
		for (int v = 0; v < nv; v++)
			s += (double)kin[kin_at(sr.ne, neq, nv, ipg, q, v)];
		return s;
	}
	/* weight of the E (D) components of element e: eps |E|^2 (RK3),
	 * |D|^2 / eps (kinetic) */
	__device__ double weight_e(long e) const
	{
		if (!ieps)
			return 1.0;
		return nv ? ieps[e] : 1.0 / ieps[sr.pos[e]];
	}
};

/* per-block partial sums of det_e sum_q w_q^T Mref w_q over elements
 * e = block * EN_BS + thread, + grid stride; det_e = det[e * dstride] */
template <class A>
__global__ __launch_bounds__(EN_BS) void k_energy_partial(A a, const double *__restrict__ det,
							  int dstride, long ne,
							  double *__restrict__ partials)
{
	__shared__ double red[EN_BS];
	double acc = 0.0;
	for (long e = (long)blockIdx.x * EN_BS + threadIdx.x; e < ne;
	     e += (long)gridDim.x * EN_BS) {
		double s = 0.0, se = 0.0;
		for (int q = 0; q < 6; q++) {
			if (q == 3) { /* E components done: their weight */
				se = s;
				s = 0.0;
			}
			double w[10];
#pragma unroll
			for (int i = 0; i < 10; i++)
				w[i] = a.w(10 * e + i, q);
#pragma unroll
			for (int i = 0; i < 10; i++) {
				double mw = 0.0;
#pragma unroll
				for (int j = 0; j < 10; j++)
					mw += c_mass[i][j] * w[j];
				s += w[i] * mw;
			}
		}
		acc += det[e * dstride] * (a.weight_e(e) * se + s);
	}
	red[threadIdx.x] = acc;
	__syncthreads();
	for (int h = EN_BS / 2; h > 0; h >>= 1) {
		if (threadIdx.x < h)
			red[threadIdx.x] += red[threadIdx.x + h];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		partials[blockIdx.x] = red[0];
}

/* out[0] = 1/2 sum of the np partials: 64 lanes in a fixed stride, then a
 * fixed-order tree */
__global__ __launch_bounds__(64) void k_energy_final(const double *__restrict__ partials, int np,
						     double *__restrict__ out)
{
	__shared__ double red[64];
	const int t = threadIdx.x;
	double s = 0.0;
	for (int b = t; b < np; b += 64)
		s += partials[b];
	red[t] = s;
	__syncthreads();
	for (int h = 32; h > 0; h >>= 1) {
		if (t < h)
			red[t] += red[t + h];
		__syncthreads();
	}
	if (t == 0)
		out[0] = 0.5 * red[0];
}

static int energy_blocks(long ne)
{
	return (int)std::min<long>(EN_MAXB, std::max<long>(1, (ne + EN_BS - 1) / EN_BS));
}

/* both passes on stream s into out (device), the partial sums (at most
 * EN_MAXB doubles) in part */
template <class A>
static void energy_launch(const A &a, const double *det, int dstride, long ne, double *part,
			  double *out, hipStream_t s)
{
	const int np = energy_blocks(ne);
	k_energy_partial<A><<<np, EN_BS, 0, s>>>(a, det, dstride, ne, part);
	HIPCHECK(hipGetLastError());
	k_energy_final<<<1, 64, 0, s>>>(part, np, out);
	HIPCHECK(hipGetLastError());
}

extern "C" double gdh_energy(gdh_ctx *c)
{
	relax_materialize(c);
	/* partial sums at the start of the context's reduction scratch (1024 + 1
	 * blocks of 8 doubles), the result after those of the L2 error */
	double *out = c->partials + 6L * c->npartials + 1;
	if (c->fp32) {
		EnergyDev<float> a;
		a.kin = (const float *)c->w[c->cur];
		a.sr.w = (float *)c->w[c->cur];
		a.sr.gh = (float *)c->gh;
		a.sr.pos = c->elem_pos;
		a.sr.ne = c->ne;
		a.sr.ngh = 6L * c->nb;
		a.nv = c->scheme ? c->kp.nv : 0;
		a.neq = c->neq;
		a.ieps = c->ieps;
		energy_launch(a, c->geom + 9, 10, c->ne, c->partials, out, c->stream);
	} else {
		EnergyDev<double> a;
		a.kin = (const double *)c->w[c->cur];
		a.sr.w = (double *)c->w[c->cur];
		a.sr.gh = (double *)c->gh;
		a.sr.pos = c->elem_pos;
		a.sr.ne = c->ne;
		a.sr.ngh = 6L * c->nb;
		a.nv = c->scheme ? c->kp.nv : 0;
		a.neq = c->neq;
		a.ieps = c->ieps;
		energy_launch(a, c->geom + 9, 10, c->ne, c->partials, out, c->stream);
	}
	double e = 0.0;
	HIPCHECK(hipMemcpyAsync(&e, out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
	check_sweep_status(c);
	return e;
}

/* argument checks shared by the device-pointer entry points */
static int energy_args_ok(const double *wn, int nv, int ne, int nb, const double *det,
			  const double *out)
{
	if (!wn || !det || !out || ne < 0 || nb < 0 || nv < 0 || nv > GDH_MAX_V) {
		fprintf(stderr, "gdon3d-hip: gdh_energy_dev: invalid arguments\n");
		return 0;
	}
	/* element DoF indices 10 ne + 9 and row offsets (q nv + v) neq stay
	 * within 64-bit arithmetic; the element count within int */
	return 1;
}

extern "C" int gdh_energy_dev(const double *wn, int nv, int ne, int nb, const double *det,
			      double *out, void *stream)
{
	if (!energy_args_ok(wn, nv, ne, nb, det, out))
		return -1;
	const hipStream_t s = (hipStream_t)stream;
	const EnergyRef a{ wn, nv, 10L * ne + 6L * nb };
	double *part = (double *)dmalloc(sizeof(double) * EN_MAXB);
	energy_launch(a, det, 1, ne, part, out, s);
	HIPCHECK(hipStreamSynchronize(s));
	HIPCHECK(hipFree(part));
	return 0;
}

/*
 * The energy without any allocation: the partial sums go to a fixed device
 * array of EN_MAXB doubles (8 KiB, held on every GPU the library runs on
 * until the process ends) instead of a buffer allocated per call. Its device
 * address is looked up once per device, with that device current, at the
 * first call on it; host code uses that address, never the array's name.
 * Every partial sum is written before it is read, so nothing needs clearing;
 * calls that overlap on one GPU share the array and race.
 */
static __device__ __align__(256) double scratch[EN_MAXB];

#define EN_MAX_DEVICES 64
static double *scratch_dev[EN_MAX_DEVICES];
static pthread_mutex_t scratch_lock = PTHREAD_MUTEX_INITIALIZER;

/* the address of this device's copy of scratch, or null on failure */
static double *scratch_of_current_device()
{
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= EN_MAX_DEVICES) {
		fprintf(stderr, "gdon3d-hip: gdh_energy_dev_noalloc: no usable current device\n");
		return nullptr;
	}
	pthread_mutex_lock(&scratch_lock);
	if (!scratch_dev[dev]) {
		void *p = nullptr;
		const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (rc != hipSuccess)
			fprintf(stderr, "gdon3d-hip: hipGetSymbolAddress(scratch) failed: %s\n",
				hipGetErrorString(rc));
		else
			scratch_dev[dev] = (double *)p;
	}
	double *p = scratch_dev[dev];
	pthread_mutex_unlock(&scratch_lock);
	return p;
}

extern "C" int gdh_energy_dev_noalloc(const double *wn, int nv, int ne, int nb,
				      const double *det, double *out, void *stream)
{
	if (!energy_args_ok(wn, nv, ne, nb, det, out))
		return -1;
	if (energy_blocks(ne) > EN_MAXB) { /* the grid is capped: never taken */
		fprintf(stderr, "gdon3d-hip: gdh_energy_dev_noalloc: %d partial sums exceed the "
				"fixed scratch of %d\n", energy_blocks(ne), EN_MAXB);
		return -1;
	}
	double *part = scratch_of_current_device();
	if (!part)
		return -1;
	const EnergyRef a{ wn, nv, 10L * ne + 6L * nb };
	energy_launch(a, det, 1, ne, part, out, (hipStream_t)stream);
	return 0;
}

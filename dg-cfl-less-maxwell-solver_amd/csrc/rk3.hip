/*
 * Fused low-storage RK3 stage of the P2 DG Maxwell operator (FP64), one
 * thread per element over structure-of-arrays state.
 *
 * HBM layout (RK3 scheme): W[q][i][e] = field q at local node i of element e
 * (so a wave's 64 elements read 512 contiguous bytes per (q, i)), ghost
 * values GH[q][6 bf + a], geometry G[k][e] (codtau row-major, then det).
 *
 * Per element, with gamma_k = column k of codtau (= det J^-T):
 *   dw = (1/det) [ - sum_k A(gamma_k) (Minv D_k) w + sum_f (Minv L_f S) Phi_f ]
 * where (D_k)_mj = int d_k(phi_m) phi_j (src/simulation/simulation.c:291-319,
 * c_ij = codtau int_ref), A(n) w = (-n x H, n x E) is the Maxwell flux
 * (src/models/maxwell.c:152-192 with wL = wR), Phi_f the upwind flux at the
 * 6 face nodes with the scaled outward normal (src/simulation/simulation.c:
 * 377-487), L_f S the face-mass lift and Minv / det the inverse mass
 * (src/simulation/simulation.c:218-261). The products Minv D_k and Minv L_f S
 * are exact tables; they are uniform across the wave and live in scalar
 * registers. Then k2 = A k2 - dt dw, w_out = w + B k2 (src/dg/explicit.c:
 * 82-95).
 */
#include "ctx.h"

using namespace gdh;

/* The operator tables reach the kernel through a pointer, so that they are
 * read with scalar loads into SGPR operands instead of being materialised as
 * immediates held in vector registers. T = double (plain form) or float
 * (FP32 form: the exact rationals rounded once more, to binary32). */
template <typename T> struct Rk3Tables {
	T minv_int[3][10][10];
	T minv_lift[4][10][6];
};
static const Rk3Tables<double> h_rk3_tables = { T10_MINV_INTERNAL_INIT, T10_MINV_LIFT_INIT };
static Rk3Tables<double> *d_rk3_tables64 = nullptr;
static Rk3Tables<float> *d_rk3_tables32 = nullptr;


/* Bijection from the dispatch order of blocks to tiles such that the blocks
 * the hardware sends to one XCD (block ids congruent mod 8) get a contiguous
 * range of tiles: face neighbours then mostly share that XCD's L2. */
__device__ inline int xcd_tile(int hw, int nblk)
{
	const int q = nblk / 8, r = nblk % 8, x = hw % 8;
	return x * q + min(x, r) + hw / 8;
}

/* The same pointer value, hidden from the optimiser: loads through it are
 * not merged with earlier loads of the same addresses. Re-reading the
 * element's nodes from the L1/L2 is cheaper than keeping all of them live
 * in registers across the kernel. */
template <typename T> __device__ __forceinline__ const T *reload_ptr(const T *p)
{
	asm volatile("" : "+s"(p));
	return p;
}

/* Materialise the accumulators at a phase boundary: without it the
 * optimiser re-associates the sums across phases and sinks every product to
 * the final update, keeping all intermediate values live. */
template <typename T> __device__ __forceinline__ void pin(T (&acc)[6][10])
{
#pragma unroll
	for (int p = 0; p < 6; p++)
#pragma unroll
		for (int m = 0; m < 10; m++)
			asm volatile("" : "+v"(acc[p][m]));
	__builtin_amdgcn_sched_barrier(0);
}

/* acc_p -= A(n)[p][q] h for the two outputs p that field q feeds */
template <int Q, typename T>
__device__ __forceinline__ void curl_accumulate(T (&acc)[6][10], const T (&h)[10],
						T nx, T ny, T nz)
{
	/* A(n): E-flux = -n x H, H-flux = n x E */
	constexpr int p1 = Q == 0 ? 4 : Q == 1 ? 3 : Q == 2 ? 3 : Q == 3 ? 1 : Q == 4 ? 0 : 0;
	constexpr int p2 = Q == 0 ? 5 : Q == 1 ? 5 : Q == 2 ? 4 : Q == 3 ? 2 : Q == 4 ? 2 : 1;
	const T c1 = Q == 0 ? nz : Q == 1 ? -nz : Q == 2 ? ny : Q == 3 ? -nz : Q == 4 ? nz : -ny;
	const T c2 = Q == 0 ? -ny : Q == 1 ? nx : Q == 2 ? -nx : Q == 3 ? ny : Q == 4 ? -nx : nx;
#pragma unroll
	for (int m = 0; m < 10; m++) {
		acc[p1][m] -= c1 * h[m];
		acc[p2][m] -= c2 * h[m];
	}
}

template <int Q, typename T>
__device__ __forceinline__ void volume_field(T (&acc)[6][10],
					     const T *__restrict__ W, long ne,
					     int e, const double (&ct)[9],
					     const Rk3Tables<T> *__restrict__ Tb)
{
	T wq[10];
#pragma unroll
	for (int j = 0; j < 10; j++)
		wq[j] = W[(Q * 10 + j) * ne + e];
#pragma unroll
	for (int k = 0; k < 3; k++) {
		T h[10];
#pragma unroll
		for (int m = 0; m < 10; m++) {
			T s = T(0);
#pragma unroll
			for (int j = 0; j < 10; j++)
				s += Tb->minv_int[k][m][j] * wq[j];
			h[m] = s;
		}
		curl_accumulate<Q>(acc, h, (T)ct[0 + k], (T)ct[3 + k], (T)ct[6 + k]);
		pin(acc);
	}
}

/* Face f's contribution; the scaled normal and 1/(|n| + 1e-16) are formed
 * in FP64 from the FP64 geometry, then used in T. */
template <int F, typename T>
__device__ __forceinline__ void face_term(T (&acc)[6][10], const T *__restrict__ W,
					  const T *__restrict__ GH, long ne,
					  long ngh, int e, int nb, int map,
					  const double (&ctd)[9], int bflux,
					  const Rk3Tables<T> *__restrict__ Tb)
{
	double vnd[3];
	face_normal(ctd, F, vnd);
	const T overr = (T)(1.0 / (sqrt(vnd[0] * vnd[0] + vnd[1] * vnd[1] +
					vnd[2] * vnd[2]) + 1e-16));
	const T vn[3] = { (T)vnd[0], (T)vnd[1], (T)vnd[2] };
#pragma unroll
	for (int b = 0; b < 6; b++) {
		const int i = face_node(F, b);
		T wl[6], wr[6], fl[6];
#pragma unroll
		for (int q = 0; q < 6; q++)
			wl[q] = W[(q * 10 + i) * ne + e];
		if (nb >= 0) {
			const int j = (map >> (4 * b)) & 15;
#pragma unroll
			for (int q = 0; q < 6; q++)
				wr[q] = W[(q * 10 + j) * ne + nb];
		} else {
			const long k = 6L * (-1 - nb) + b;
#pragma unroll
			for (int q = 0; q < 6; q++)
				wr[q] = GH[q * ngh + k];
		}
		if (nb < 0 && bflux == GDH_FLUX_MAXWELL_METAL)
			maxwell_metal(wl, vn, fl);
		else
			maxwell_upwind(wl, wr, vn, overr, fl);
#pragma unroll
		for (int m = 0; m < 10; m++) {
			const T lift = Tb->minv_lift[F][m][b];
#pragma unroll
			for (int p = 0; p < 6; p++)
				acc[p][m] += lift * fl[p];
		}
		pin(acc);
	}
}

/* face_term with the tile's values in LDS (Ws, column t) and the neighbour
 * trace from LDS when the neighbour column is in the tile [c0, c0 + 64) */
template <int F, typename T>
__device__ __forceinline__ void face_term_tile(T (&acc)[6][10], const T *Ws,
					       const T *__restrict__ W,
					       const T *__restrict__ GH, long ne, long ngh,
					       int t, int c0, int nb, int map,
					       const double (&ctd)[9], int bflux,
					       const Rk3Tables<T> *__restrict__ Tb)
{
	double vnd[3];
	face_normal(ctd, F, vnd);
	const T overr = (T)(1.0 / (sqrt(vnd[0] * vnd[0] + vnd[1] * vnd[1] +
					vnd[2] * vnd[2]) + 1e-16));
	const T vn[3] = { (T)vnd[0], (T)vnd[1], (T)vnd[2] };
	/* the whole neighbour trace first, one load path for every lane (a
	 * neighbour inside the tile was just read by this block and comes from
	 * the cache): one round trip per face */
	(void)c0;
	T wra[6][6];
#pragma unroll
	for (int b = 0; b < 6; b++) {
		const int j = (map >> (4 * b)) & 15;
		if (nb >= 0) {
#pragma unroll
			for (int q = 0; q < 6; q++)
				wra[b][q] = W[(q * 10 + j) * ne + nb];
		} else {
			const long k = 6L * (-1 - nb) + b;
#pragma unroll
			for (int q = 0; q < 6; q++)
				wra[b][q] = GH[q * ngh + k];
		}
	}
#pragma unroll
	for (int b = 0; b < 6; b++) {
		const int i = face_node(F, b);
		T wl[6], fl[6];
		T(&wr)[6] = wra[b];
#pragma unroll
		for (int q = 0; q < 6; q++)
			wl[q] = Ws[(q * 10 + i) * 64 + t];
		if (nb < 0 && bflux == GDH_FLUX_MAXWELL_METAL)
			maxwell_metal(wl, vn, fl);
		else
			maxwell_upwind(wl, wr, vn, overr, fl);
#pragma unroll
		for (int m = 0; m < 10; m++) {
			const T lift = Tb->minv_lift[F][m][b];
#pragma unroll
			for (int p = 0; p < 6; p++)
				acc[p][m] += lift * fl[p];
		}
		pin(acc);
	}
}

/* One stage over storage columns e; T = double or float for the state and
 * the flux/volume arithmetic, geometry G always FP64; BS elements per block.
 * Batches: the time step of the column's instance (dt_tab[inst[e]]) and,
 * with ieps, dE/dt scaled by 1/eps of the column (eps dE/dt = curl H). */
template <typename T, int BS>
__global__ __launch_bounds__(BS) void k_rk3_soa(
	const T *__restrict__ W, T *__restrict__ Wo, T *__restrict__ K2,
	const T *__restrict__ GH, const double *__restrict__ G,
	const int4 *__restrict__ nbr, const int4 *__restrict__ nbmap,
	const Rk3Tables<T> *__restrict__ Tb, int ne_, int ngh_, double Ac,
	double Bc, double dt, int first, int bflux, const int *__restrict__ inst,
	const double *__restrict__ dt_tab, const double *__restrict__ ieps)
{
	const long ne = ne_, ngh = ngh_;
	const int tile = xcd_tile(blockIdx.x, gridDim.x);
	const int e = tile * BS + threadIdx.x;
	if (e >= ne_)
		return;
	double ctd[9];
#pragma unroll
	for (int k = 0; k < 9; k++)
		ctd[k] = G[k * ne + e];
	const double det = G[9 * ne + e];

	T acc[6][10];
#pragma unroll
	for (int p = 0; p < 6; p++)
#pragma unroll
		for (int m = 0; m < 10; m++)
			acc[p][m] = T(0);

	volume_field<0>(acc, reload_ptr(W), ne, e, ctd, Tb);
	pin(acc);
	volume_field<1>(acc, reload_ptr(W), ne, e, ctd, Tb);
	pin(acc);
	volume_field<2>(acc, reload_ptr(W), ne, e, ctd, Tb);
	pin(acc);
	volume_field<3>(acc, reload_ptr(W), ne, e, ctd, Tb);
	pin(acc);
	volume_field<4>(acc, reload_ptr(W), ne, e, ctd, Tb);
	pin(acc);
	volume_field<5>(acc, reload_ptr(W), ne, e, ctd, Tb);
	pin(acc);

	const int4 n4 = nbr[e], m4 = nbmap[e];
	face_term<0>(acc, reload_ptr(W), GH, ne, ngh, e, n4.x, m4.x, ctd, bflux, Tb);
	pin(acc);
	face_term<1>(acc, reload_ptr(W), GH, ne, ngh, e, n4.y, m4.y, ctd, bflux, Tb);
	pin(acc);
	face_term<2>(acc, reload_ptr(W), GH, ne, ngh, e, n4.z, m4.z, ctd, bflux, Tb);
	pin(acc);
	face_term<3>(acc, reload_ptr(W), GH, ne, ngh, e, n4.w, m4.w, ctd, bflux, Tb);
	pin(acc);

	const T invdet = (T)(1.0 / det);
	const T idE = ieps ? (T)(ieps[e] / det) : invdet;
	const T dtT = (T)(inst ? dt_tab[inst[e]] : dt), AcT = (T)Ac, BcT = (T)Bc;
	const T *Wr = reload_ptr(W);
#pragma unroll
	for (int q = 0; q < 6; q++)
#pragma unroll
		for (int i = 0; i < 10; i++) {
			const long g = (q * 10 + i) * ne + e;
			const T d = acc[q][i] * (q < 3 ? idE : invdet);
			const T k = first ? -dtT * d : AcT * K2[g] - dtT * d;
			K2[g] = k;
			Wo[g] = Wr[g] + BcT * k;
		}
}

/*
 * The same stage with the state of a tile of 64 consecutive storage columns
 * (a compact Morton block) staged in LDS by one coalesced pass: the
 * element's own values are read from LDS in every phase, and so is the
 * trace of a face neighbour inside the tile (about 70 % of the interior
 * faces at 64 elements per tile); neighbours outside the tile and ghosts
 * come from global memory. One 64-lane block per tile, one element per
 * lane; the arithmetic is that of k_rk3_soa.
 */
template <typename T>
__global__ __launch_bounds__(64) void k_rk3_tile(
	const T *__restrict__ W, T *__restrict__ Wo, T *__restrict__ K2,
	const T *__restrict__ GH, const double *__restrict__ G,
	const int4 *__restrict__ nbr, const int4 *__restrict__ nbmap,
	const Rk3Tables<T> *__restrict__ Tb, int ne_, int ngh_, double Ac,
	double Bc, double dt, int first, int bflux, const int *__restrict__ inst,
	const double *__restrict__ dt_tab, const double *__restrict__ ieps)
{
	__shared__ T Ws[60 * 64]; /* value (q, i) of tile column t at (10 q + i) 64 + t */
	const long ne = ne_, ngh = ngh_;
	const int tile = xcd_tile(blockIdx.x, gridDim.x);
	const int t = threadIdx.x, c0 = tile * 64, e = c0 + t;
	const bool live = e < ne_;
	{
		/* all 60 loads in flight, then the LDS stores */
		T v[60];
#pragma unroll
		for (int r = 0; r < 60; r++)
			v[r] = live ? W[r * ne + e] : T(0);
#pragma unroll
		for (int r = 0; r < 60; r++)
			Ws[r * 64 + t] = v[r];
	}
	__syncthreads();
	if (!live)
		return;
	double ctd[9];
#pragma unroll
	for (int k = 0; k < 9; k++)
		ctd[k] = G[k * ne + e];
	const double det = G[9 * ne + e];

	T acc[6][10];
#pragma unroll
	for (int p = 0; p < 6; p++)
#pragma unroll
		for (int m = 0; m < 10; m++)
			acc[p][m] = T(0);

	volume_field<0>(acc, Ws, 64, t, ctd, Tb);
	pin(acc);
	volume_field<1>(acc, Ws, 64, t, ctd, Tb);
	pin(acc);
	volume_field<2>(acc, Ws, 64, t, ctd, Tb);
	pin(acc);
	volume_field<3>(acc, Ws, 64, t, ctd, Tb);
	pin(acc);
	volume_field<4>(acc, Ws, 64, t, ctd, Tb);
	pin(acc);
	volume_field<5>(acc, Ws, 64, t, ctd, Tb);
	pin(acc);

	const int4 n4 = nbr[e], m4 = nbmap[e];
	face_term_tile<0>(acc, Ws, W, GH, ne, ngh, t, c0, n4.x, m4.x, ctd, bflux, Tb);
	pin(acc);
	face_term_tile<1>(acc, Ws, W, GH, ne, ngh, t, c0, n4.y, m4.y, ctd, bflux, Tb);
	pin(acc);
	face_term_tile<2>(acc, Ws, W, GH, ne, ngh, t, c0, n4.z, m4.z, ctd, bflux, Tb);
	pin(acc);
	face_term_tile<3>(acc, Ws, W, GH, ne, ngh, t, c0, n4.w, m4.w, ctd, bflux, Tb);
	pin(acc);

	const T invdet = (T)(1.0 / det);
	const T idE = ieps ? (T)(ieps[e] / det) : invdet;
	const T dtT = (T)(inst ? dt_tab[inst[e]] : dt), AcT = (T)Ac, BcT = (T)Bc;
#pragma unroll
	for (int q = 0; q < 6; q++)
#pragma unroll
		for (int i = 0; i < 10; i++) {
			const long g = (q * 10 + i) * ne + e;
			const T d = acc[q][i] * (q < 3 ? idE : invdet);
			const T k = first ? -dtT * d : AcT * K2[g] - dtT * d;
			K2[g] = k;
			Wo[g] = Ws[(q * 10 + i) * 64 + t] + BcT * k;
		}
}

template <typename T, int BS>
static void rk3_launch_bs(gdh_ctx *c, const Rk3Tables<T> *tab, double Ac, double Bc,
			  double dt, int first, const void *W, void *Wo)
{
	k_rk3_soa<T, BS><<<nblocks(c->ne, BS), BS, 0, c->stream>>>(
		(const T *)W, (T *)Wo, (T *)c->k2, (const T *)c->gh, c->geom_soa, c->nbr,
		c->nbmap, tab, c->ne, 6 * c->nb, Ac, Bc, dt, first, c->md.bflux_kind,
		c->bt.einst_col, c->bt.dt, c->bt.ieps_col);
	HIPCHECK(hipGetLastError());
}

template <typename T>
static void rk3_launch(gdh_ctx *c, const Rk3Tables<T> *tab, double Ac, double Bc,
		       double dt, int first, const void *W, void *Wo)
{
	if (c->rk3_tile) {
		k_rk3_tile<T><<<nblocks(c->ne, 64), 64, 0, c->stream>>>(
			(const T *)W, (T *)Wo, (T *)c->k2, (const T *)c->gh, c->geom_soa, c->nbr,
			c->nbmap, tab, c->ne, 6 * c->nb, Ac, Bc, dt, first, c->md.bflux_kind,
			c->bt.einst_col, c->bt.dt, c->bt.ieps_col);
		HIPCHECK(hipGetLastError());
		return;
	}
	switch (c->rk3_bs) {
	case 64: rk3_launch_bs<T, 64>(c, tab, Ac, Bc, dt, first, W, Wo); break;
	case 256: rk3_launch_bs<T, 256>(c, tab, Ac, Bc, dt, first, W, Wo); break;
	default: rk3_launch_bs<T, 128>(c, tab, Ac, Bc, dt, first, W, Wo); break;
	}
}

static void rk3_tables_init()
{
	if (d_rk3_tables64)
		return;
	d_rk3_tables64 = (Rk3Tables<double> *)dmalloc(sizeof(Rk3Tables<double>));
	HIPCHECK(hipMemcpy(d_rk3_tables64, &h_rk3_tables, sizeof(Rk3Tables<double>),
			   hipMemcpyHostToDevice));
	Rk3Tables<float> t32;
	for (int k = 0; k < 300; k++)
		(&t32.minv_int[0][0][0])[k] = (float)(&h_rk3_tables.minv_int[0][0][0])[k];
	for (int k = 0; k < 240; k++)
		(&t32.minv_lift[0][0][0])[k] = (float)(&h_rk3_tables.minv_lift[0][0][0])[k];
	d_rk3_tables32 = (Rk3Tables<float> *)dmalloc(sizeof(Rk3Tables<float>));
	HIPCHECK(hipMemcpy(d_rk3_tables32, &t32, sizeof(Rk3Tables<float>),
			   hipMemcpyHostToDevice));
}

/* One first-stage launch from the current state into the other buffer
 * (timing only: used by the tile autotuner before a run starts). */
void rk3_trial(gdh_ctx *c)
{
	rk3_tables_init();
	if (c->fp32)
		rk3_launch<float>(c, d_rk3_tables32, 0.0, 1.0 / 3.0, 1e-3, 1,
				  c->w[c->cur], c->w[1 - c->cur]);
	else
		rk3_launch<double>(c, d_rk3_tables64, 0.0, 1.0 / 3.0, 1e-3, 1,
				   c->w[c->cur], c->w[1 - c->cur]);
}

extern "C" void gdh_rk3_stage(gdh_ctx *c, int stage, double dt)
{
	static const double A[3] = { 0., -5. / 9., -153. / 128. };
	static const double B[3] = { 1. / 3., 15. / 16., 8. / 15. };
	rk3_tables_init();
	prof_begin(c, 0);
	c->prof_launches++;
	if (c->fp32)
		rk3_launch<float>(c, d_rk3_tables32, A[stage], B[stage], dt, stage == 0,
				  c->w[c->cur], c->w[1 - c->cur]);
	else
		rk3_launch<double>(c, d_rk3_tables64, A[stage], B[stage], dt, stage == 0,
				   c->w[c->cur], c->w[1 - c->cur]);
	prof_end(c, 0);
	c->cur = 1 - c->cur;
}

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

/* The stage's k2 stream (read once and written once per stage): with
 * non-temporal hints (NT) or ordinary loads and stores. Which is faster
 * depends on the workload (measured: non-temporal at 64^3, ordinary on
 * the batch of small instances in binary32), so both are built. */
template <bool NT, typename T> __device__ __forceinline__ T k2_load(const T *p)
{
	if constexpr (NT)
		return __builtin_nontemporal_load(p);
	else
		return *p;
}

template <bool NT, typename T> __device__ __forceinline__ void k2_store(T v, T *p)
{
	if constexpr (NT)
		__builtin_nontemporal_store(v, p);
	else
		*p = v;
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
template <int F, typename T, int TL>
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
	/* the whole neighbour trace first, one round trip per face: in binary32
	 * from LDS for a neighbour inside the tile (64^3: 0.68 vs 0.70 ms per
	 * stage), else, and always in binary64 (1.34 vs 1.32 ms), from global
	 * memory */
	const unsigned nt = (unsigned)(nb - c0);
	T wra[6][6];
#pragma unroll
	for (int b = 0; b < 6; b++) {
		const int j = (map >> (4 * b)) & 15;
		if (sizeof(T) == 4 && nt < (unsigned)TL) {
#pragma unroll
			for (int q = 0; q < 6; q++)
				wra[b][q] = Ws[(q * 10 + j) * TL + nt];
		} else if (nb >= 0) {
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
			wl[q] = Ws[(q * 10 + i) * TL + t];
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
template <typename T, int BS, bool NTK>
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
			const T k = first ? -dtT * d : AcT * k2_load<NTK>(K2 + g) - dtT * d;
			k2_store<NTK>(k, K2 + g);
			__builtin_nontemporal_store(Wr[g] + BcT * k, Wo + g);
		}
}

/*
 * The same stage with the state of a tile of 64 NW consecutive storage
 * columns (a compact Morton block) staged in LDS by one coalesced pass: the
 * element's own values are read from LDS in every phase, and so is the
 * trace of a face neighbour inside the tile (about 70 % of the interior
 * faces at 64 elements per tile, more in larger tiles); neighbours outside
 * the tile and ghosts come from global memory. One block of NW 64-lane
 * waves per tile, one element per lane; the arithmetic is that of
 * k_rk3_soa.
 */
template <typename T, bool NTK, int NW>
__global__ __launch_bounds__(64 * NW) void k_rk3_tile(
	const T *__restrict__ W, T *__restrict__ Wo, T *__restrict__ K2,
	const T *__restrict__ GH, const double *__restrict__ G,
	const int4 *__restrict__ nbr, const int4 *__restrict__ nbmap,
	const Rk3Tables<T> *__restrict__ Tb, int ne_, int ngh_, double Ac,
	double Bc, double dt, int first, int bflux, const int *__restrict__ inst,
	const double *__restrict__ dt_tab, const double *__restrict__ ieps)
{
	constexpr int TL = 64 * NW;
	__shared__ T Ws[60 * TL]; /* value (q, i) of tile column t at (10 q + i) TL + t */
	const long ne = ne_, ngh = ngh_;
	const int tile = xcd_tile(blockIdx.x, gridDim.x);
	const int t = threadIdx.x, c0 = tile * TL, e = c0 + t;
	const bool live = e < ne_;
	{
		/* all 60 loads in flight, then the LDS stores */
		T v[60];
#pragma unroll
		for (int r = 0; r < 60; r++)
			v[r] = live ? W[r * ne + e] : T(0);
#pragma unroll
		for (int r = 0; r < 60; r++)
			Ws[r * TL + t] = v[r];
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

	volume_field<0>(acc, Ws, TL, t, ctd, Tb);
	pin(acc);
	volume_field<1>(acc, Ws, TL, t, ctd, Tb);
	pin(acc);
	volume_field<2>(acc, Ws, TL, t, ctd, Tb);
	pin(acc);
	volume_field<3>(acc, Ws, TL, t, ctd, Tb);
	pin(acc);
	volume_field<4>(acc, Ws, TL, t, ctd, Tb);
	pin(acc);
	volume_field<5>(acc, Ws, TL, t, ctd, Tb);
	pin(acc);

	const int4 n4 = nbr[e], m4 = nbmap[e];
	face_term_tile<0, T, TL>(acc, Ws, W, GH, ne, ngh, t, c0, n4.x, m4.x, ctd, bflux, Tb);
	pin(acc);
	face_term_tile<1, T, TL>(acc, Ws, W, GH, ne, ngh, t, c0, n4.y, m4.y, ctd, bflux, Tb);
	pin(acc);
	face_term_tile<2, T, TL>(acc, Ws, W, GH, ne, ngh, t, c0, n4.z, m4.z, ctd, bflux, Tb);
	pin(acc);
	face_term_tile<3, T, TL>(acc, Ws, W, GH, ne, ngh, t, c0, n4.w, m4.w, ctd, bflux, Tb);
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
			const T k = first ? -dtT * d : AcT * k2_load<NTK>(K2 + g) - dtT * d;
			k2_store<NTK>(k, K2 + g);
			__builtin_nontemporal_store(Ws[(q * 10 + i) * TL + t] + BcT * k, Wo + g);
		}
}

template <typename T> __device__ __forceinline__ void pin3(T (&acc)[3][10])
{
#pragma unroll
	for (int p = 0; p < 3; p++)
#pragma unroll
		for (int m = 0; m < 10; m++)
			asm volatile("" : "+v"(acc[p][m]));
	__builtin_amdgcn_sched_barrier(0);
}

/*
 * Two lanes per element: each 64-lane wave takes 32 consecutive storage
 * columns (15 KB of LDS in FP64), a block of NW waves a tile of 32 NW
 * columns staged in LDS together (more face neighbours inside the tile:
 * 64^3 binary64 stage 1.133 ms with 32 columns, 1.091 with 64, 1.076 with
 * 128); lane half h = 0 forms the E rows of the update, h = 1 the H rows. The
 * code is the same for both halves: the curl terms of the H rows are those
 * of the E rows with E and H exchanged and the sign reversed (A(n): E-flux
 * = -n x H, H-flux = n x E), so a half differs only by its source fields
 * (H or E) and a sign sigma = -1 or +1. Both halves form the face fluxes;
 * each lifts its three components.
 */
template <int J, typename T, int TL>
__device__ __forceinline__ void volume_pair(T (&acc)[3][10], const T *Ws, int c, int qs0,
					    T sigma, const double (&ct)[9],
					    const Rk3Tables<T> *__restrict__ Tb)
{
	T wq[10];
#pragma unroll
	for (int j = 0; j < 10; j++)
		wq[j] = Ws[((qs0 + J) * 10 + j) * TL + c];
	/* source component J feeds outputs o1, o2 with (sign * n_comp) */
	constexpr int o1 = J == 0 ? 1 : 0, o2 = J == 2 ? 1 : 2;
#pragma unroll
	for (int k = 0; k < 3; k++) {
		const T nx = (T)ct[0 + k], ny = (T)ct[3 + k], nz = (T)ct[6 + k];
		/* J = 0: o1 -= s nz, o2 += s ny; J = 1: o0 += s nz, o2 -= s nx;
		 * J = 2: o0 -= s ny, o1 += s nx */
		const T c1 = J == 0 ? sigma * nz : J == 1 ? -sigma * nz : sigma * ny;
		const T c2 = J == 0 ? -sigma * ny : J == 1 ? sigma * nx : -sigma * nx;
#pragma unroll
		for (int m = 0; m < 10; m++) {
			T hm = T(0);
#pragma unroll
			for (int j = 0; j < 10; j++)
				hm += Tb->minv_int[k][m][j] * wq[j];
			acc[o1][m] -= c1 * hm;
			acc[o2][m] -= c2 * hm;
		}
		pin3(acc);
	}
}

/* exchange between lane halves: lanes 32-63 of a take lanes 0-31 of b and
 * the other way round (v_permlane32_swap), one dword at a time */
__device__ __forceinline__ void swap_halves(float &a, float &b)
{
	auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(unsigned, a),
						  __builtin_bit_cast(unsigned, b), false, false);
	a = __builtin_bit_cast(float, (unsigned)r[0]);
	b = __builtin_bit_cast(float, (unsigned)r[1]);
}

__device__ __forceinline__ void swap_halves(double &a, double &b)
{
	const unsigned long long ua = __builtin_bit_cast(unsigned long long, a);
	const unsigned long long ub = __builtin_bit_cast(unsigned long long, b);
	auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ua, (unsigned)ub, false, false);
	auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32),
						   false, false);
	a = __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi[0] << 32) | (unsigned)lo[0]);
	b = __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi[1] << 32) | (unsigned)lo[1]);
}

/* Neighbour trace of face F for the pair layout: lane half h gathers the
 * six fields at face nodes b = 3 h + b' (b' < 3), from the neighbour's
 * columns or the ghost DoFs of boundary face -1 - nb */
template <int F, typename T, int TL>
__device__ __forceinline__ void gather_pair(T (&wra)[3][6], const T *Ws, const T *__restrict__ W,
					    const T *__restrict__ GH, long ne, long ngh, int h,
					    int c0, int nb, int map)
{
	const unsigned nt = (unsigned)(nb - c0);
#pragma unroll
	for (int bp = 0; bp < 3; bp++) {
		const int b = 3 * h + bp;
		if (nt < (unsigned)TL) {
			const int j = (map >> (4 * b)) & 15;
#pragma unroll
			for (int q = 0; q < 6; q++)
				wra[bp][q] = Ws[(q * 10 + j) * TL + nt];
		} else if (nb >= 0) {
			const int j = (map >> (4 * b)) & 15;
#pragma unroll
			for (int q = 0; q < 6; q++)
				wra[bp][q] = W[(q * 10 + j) * ne + nb];
		} else {
			const long k = 6L * (-1 - nb) + b;
#pragma unroll
			for (int q = 0; q < 6; q++)
				wra[bp][q] = GH[q * ngh + k];
		}
	}
}

/* Face F's flux and lift for the pair layout, from the gathered trace
 * wra: each half forms the full six-component flux at its three nodes;
 * one swap per component then gives every lane, for its three components,
 * the flux at node b' (fl[p]) and node 3 + b' (fl[3 + p]): half 0 keeps
 * its E values of nodes 0-2 and takes its partner's of nodes 3-5, half 1
 * takes its partner's H values of nodes 0-2 and keeps its own of nodes
 * 3-5. */
template <int F, typename T, int TL>
__device__ __forceinline__ void face_pair(T (&acc)[3][10], const T *Ws, const T (&wra)[3][6],
					  int c, int h, int nb, const double (&ctd)[9], int bflux,
					  const Rk3Tables<T> *__restrict__ Tb)
{
	double vnd[3];
	face_normal(ctd, F, vnd);
	const T overr = (T)(1.0 / (sqrt(vnd[0] * vnd[0] + vnd[1] * vnd[1] +
					vnd[2] * vnd[2]) + 1e-16));
	const T vn[3] = { (T)vnd[0], (T)vnd[1], (T)vnd[2] };
#pragma unroll
	for (int bp = 0; bp < 3; bp++) {
		const int i = h ? face_node(F, 3 + bp) : face_node(F, bp);
		T wl[6], fl[6];
#pragma unroll
		for (int q = 0; q < 6; q++)
			wl[q] = Ws[(q * 10 + i) * TL + c];
		if (nb < 0 && bflux == GDH_FLUX_MAXWELL_METAL)
			maxwell_metal(wl, vn, fl);
		else
			maxwell_upwind(wl, wra[bp], vn, overr, fl);
#pragma unroll
		for (int p = 0; p < 3; p++)
			swap_halves(fl[p], fl[3 + p]);
#pragma unroll
		for (int m = 0; m < 10; m++) {
			const T la = Tb->minv_lift[F][m][bp], lb = Tb->minv_lift[F][m][3 + bp];
#pragma unroll
			for (int p = 0; p < 3; p++)
				acc[p][m] += la * fl[p] + lb * fl[3 + p];
		}
		pin3(acc);
	}
}

template <typename T, bool NTK, int NW>
__global__ __launch_bounds__(64 * NW) void k_rk3_pair(
	const T *__restrict__ W, T *__restrict__ Wo, T *__restrict__ K2,
	const T *__restrict__ GH, const double *__restrict__ G,
	const int4 *__restrict__ nbr, const int4 *__restrict__ nbmap,
	const Rk3Tables<T> *__restrict__ Tb, int ne_, int ngh_, double Ac,
	double Bc, double dt, int first, int bflux, const int *__restrict__ inst,
	const double *__restrict__ dt_tab, const double *__restrict__ ieps)
{
	constexpr int TL = 32 * NW;
	__shared__ T Ws[60 * TL]; /* value (q, i) of tile column c at (10 q + i) TL + c */
	const long ne = ne_, ngh = ngh_;
	const int tile = xcd_tile(blockIdx.x, gridDim.x);
	/* wave w holds columns 32 w .. 32 w + 31, both halves of each */
	const int l = threadIdx.x % 64;
	const int c = (threadIdx.x / 64) * 32 + l % 32, h = l / 32, c0 = tile * TL, e = c0 + c;
	const bool live = e < ne_;
	{
		/* each half stages 30 rows, all loads in flight */
		T v[30];
#pragma unroll
		for (int r = 0; r < 30; r++)
			v[r] = live ? W[(30 * h + r) * ne + e] : T(0);
#pragma unroll
		for (int r = 0; r < 30; r++)
			Ws[(30 * h + r) * TL + c] = v[r];
	}
	__syncthreads();
	if (!live)
		return;
	double ctd[9];
#pragma unroll
	for (int k = 0; k < 9; k++)
		ctd[k] = G[k * ne + e];
	const double det = G[9 * ne + e];
	T acc[3][10];
#pragma unroll
	for (int p = 0; p < 3; p++)
#pragma unroll
		for (int m = 0; m < 10; m++)
			acc[p][m] = T(0);
	/* E rows (h = 0) from the H fields, H rows (h = 1) from the E fields */
	const int qs0 = 3 - 3 * h;
	const T sigma = h ? T(1) : T(-1);
	/* each face's neighbour trace is gathered one phase ahead: face 0's
	 * during the volume term, face f + 1's during face f's flux */
	const int4 n4 = nbr[e], m4 = nbmap[e];
	T ga[3][6], gb[3][6];
	gather_pair<0, T, TL>(ga, Ws, W, GH, ne, ngh, h, c0, n4.x, m4.x);
	volume_pair<0, T, TL>(acc, Ws, c, qs0, sigma, ctd, Tb);
	volume_pair<1, T, TL>(acc, Ws, c, qs0, sigma, ctd, Tb);
	volume_pair<2, T, TL>(acc, Ws, c, qs0, sigma, ctd, Tb);
	gather_pair<1, T, TL>(gb, Ws, W, GH, ne, ngh, h, c0, n4.y, m4.y);
	face_pair<0, T, TL>(acc, Ws, ga, c, h, n4.x, ctd, bflux, Tb);
	gather_pair<2, T, TL>(ga, Ws, W, GH, ne, ngh, h, c0, n4.z, m4.z);
	face_pair<1, T, TL>(acc, Ws, gb, c, h, n4.y, ctd, bflux, Tb);
	gather_pair<3, T, TL>(gb, Ws, W, GH, ne, ngh, h, c0, n4.w, m4.w);
	face_pair<2, T, TL>(acc, Ws, ga, c, h, n4.z, ctd, bflux, Tb);
	face_pair<3, T, TL>(acc, Ws, gb, c, h, n4.w, ctd, bflux, Tb);
	const T invdet = (T)(1.0 / det);
	const T sc = h == 0 && ieps ? (T)(ieps[e] / det) : invdet;
	const T dtT = (T)(inst ? dt_tab[inst[e]] : dt), AcT = (T)Ac, BcT = (T)Bc;
#pragma unroll
	for (int p = 0; p < 3; p++)
#pragma unroll
		for (int i = 0; i < 10; i++) {
			const int r = (3 * h + p) * 10 + i;
			const long g = r * ne + e;
			const T d = acc[p][i] * sc;
			const T k = first ? -dtT * d : AcT * k2_load<NTK>(K2 + g) - dtT * d;
			k2_store<NTK>(k, K2 + g);
			__builtin_nontemporal_store(Ws[r * TL + c] + BcT * k, Wo + g);
		}
}

template <typename T, int BS, bool NTK>
static void rk3_launch_bs(gdh_ctx *c, const Rk3Tables<T> *tab, double Ac, double Bc,
			  double dt, int first, const void *W, void *Wo)
{
	k_rk3_soa<T, BS, NTK><<<nblocks(c->ne, BS), BS, 0, c->stream>>>(
		(const T *)W, (T *)Wo, (T *)c->k2, (const T *)c->gh, c->geom_soa, c->nbr,
		c->nbmap, tab, c->ne, 6 * c->nb, Ac, Bc, dt, first, c->md.bflux_kind,
		c->bt.einst_col, c->bt.dt, c->bt.ieps_col);
	HIPCHECK(hipGetLastError());
}

template <typename T, bool NTK, int NW>
static void rk3_launch_pair(gdh_ctx *c, const Rk3Tables<T> *tab, double Ac, double Bc,
			    double dt, int first, const void *W, void *Wo)
{
	k_rk3_pair<T, NTK, NW><<<nblocks(c->ne, 32 * NW), 64 * NW, 0, c->stream>>>(
		(const T *)W, (T *)Wo, (T *)c->k2, (const T *)c->gh, c->geom_soa, c->nbr,
		c->nbmap, tab, c->ne, 6 * c->nb, Ac, Bc, dt, first, c->md.bflux_kind,
		c->bt.einst_col, c->bt.dt, c->bt.ieps_col);
	HIPCHECK(hipGetLastError());
}

template <typename T, bool NTK, int NW>
static void rk3_launch_tile(gdh_ctx *c, const Rk3Tables<T> *tab, double Ac, double Bc,
			    double dt, int first, const void *W, void *Wo)
{
	k_rk3_tile<T, NTK, NW><<<nblocks(c->ne, 64 * NW), 64 * NW, 0, c->stream>>>(
		(const T *)W, (T *)Wo, (T *)c->k2, (const T *)c->gh, c->geom_soa, c->nbr,
		c->nbmap, tab, c->ne, 6 * c->nb, Ac, Bc, dt, first, c->md.bflux_kind,
		c->bt.einst_col, c->bt.dt, c->bt.ieps_col);
	HIPCHECK(hipGetLastError());
}

/* rk3_bs = elements per block: 32, 64 or 128 (pair), 64, 128 or, in
 * binary32, 256 (tile; 60 KB of LDS), 64, 128 or 256 (soa) */
template <typename T, bool NTK>
static void rk3_launch_nt(gdh_ctx *c, const Rk3Tables<T> *tab, double Ac, double Bc,
			  double dt, int first, const void *W, void *Wo)
{
	if (c->rk3_tile == 2) {
		switch (c->rk3_bs) {
		case 128: rk3_launch_pair<T, NTK, 4>(c, tab, Ac, Bc, dt, first, W, Wo); break;
		case 64: rk3_launch_pair<T, NTK, 2>(c, tab, Ac, Bc, dt, first, W, Wo); break;
		default: rk3_launch_pair<T, NTK, 1>(c, tab, Ac, Bc, dt, first, W, Wo); break;
		}
		return;
	}
	if (c->rk3_tile) {
		if constexpr (sizeof(T) == 4)
			if (c->rk3_bs == 256) {
				rk3_launch_tile<T, NTK, 4>(c, tab, Ac, Bc, dt, first, W, Wo);
				return;
			}
		if (c->rk3_bs >= 128)
			rk3_launch_tile<T, NTK, 2>(c, tab, Ac, Bc, dt, first, W, Wo);
		else
			rk3_launch_tile<T, NTK, 1>(c, tab, Ac, Bc, dt, first, W, Wo);
		return;
	}
	switch (c->rk3_bs) {
	case 64: rk3_launch_bs<T, 64, NTK>(c, tab, Ac, Bc, dt, first, W, Wo); break;
	case 256: rk3_launch_bs<T, 256, NTK>(c, tab, Ac, Bc, dt, first, W, Wo); break;
	default: rk3_launch_bs<T, 128, NTK>(c, tab, Ac, Bc, dt, first, W, Wo); break;
	}
}

template <typename T>
static void rk3_launch(gdh_ctx *c, const Rk3Tables<T> *tab, double Ac, double Bc,
		       double dt, int first, const void *W, void *Wo)
{
	if (c->rk3_ntk)
		rk3_launch_nt<T, true>(c, tab, Ac, Bc, dt, first, W, Wo);
	else
		rk3_launch_nt<T, false>(c, tab, Ac, Bc, dt, first, W, Wo);
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

/* One stage launch from the current state into the other buffer (timing
 * only: used by the tile autotuner before a run starts). A first stage does
 * not read k2; a later one (first = 0) does. Both overwrite only the other
 * state buffer and k2, which the run's first stage writes before it reads
 * them. */
void rk3_trial(gdh_ctx *c, int first)
{
	rk3_tables_init();
	const double A = first ? 0.0 : -5.0 / 9.0;
	if (c->fp32)
		rk3_launch<float>(c, d_rk3_tables32, A, 1.0 / 3.0, 1e-3, first,
				  c->w[c->cur], c->w[1 - c->cur]);
	else
		rk3_launch<double>(c, d_rk3_tables64, A, 1.0 / 3.0, 1e-3, first,
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

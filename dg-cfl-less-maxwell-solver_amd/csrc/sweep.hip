/*
 * Kinetic relaxation scheme on the GPU: per-element block factors of
 * M1 = I + dt/2 M^-1 F_v, the upwind transport sweep sol = M1^-1 (M2 f), and
 * the node-local relaxation.
 */
#include "ctx.h"
#include "learned_block.h"

#include <chrono>

using namespace gdh;

/* ===================================================================== */
/*                      kinetic relaxation scheme                        */
/* ===================================================================== */

#define THETA 0.5 /* Crank-Nicolson, src/dg/relaxation.c:133 */
/* lagged nodes of one velocity taken by the dense inverse of the
 * capacitance form (GDN_CAP_MAX_NODES overrides, up to CAP_MAX_DENSE: q^2
 * binary64 values per velocity); above, the matrix-free form */
#define CAP_MAX_NODES 4096
#define CAP_MAX_DENSE 16384

static void cap_setup(gdh_ctx *c);
static void relax_choose(gdh_ctx *c);
static void cluster_setup(gdh_ctx *c, const uint8_t *cls, const int *items_ve,
			  const int *level_ptr, int nlevels);
static int grid_cap(int g);

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
				double dt_, double *__restrict__ dinv, BatchTab bt)
{
	const int it = blockIdx.x * blockDim.x + threadIdx.x;
	if (it >= nitems)
		return;
	const int v = items[it].x, e = items[it].y;
	const double dt = bt.ninst ? bt.dt[bt.einst[e]] : dt_;
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

__global__ void k_flag_last(struct ItemDesc *__restrict__ desc, const int *__restrict__ last,
			    int ne);

extern "C" void gdh_relax_setup(gdh_ctx *c, double dt, const uint8_t *cls,
				const int *items_ve, const int *level_ptr,
				int nlevels)
{
	if (!c->scheme) {
		fprintf(stderr, "gdon3d-hip: relaxation set-up on an RK3 context\n");
		abort();
	}
	const int ne = c->ne, nv = c->kp.nv;
	/* GDN_VERBOSE: the wall time of each phase (the stream synchronized at
	 * each mark) */
	const bool verbose = getenv("GDN_VERBOSE") != nullptr;
	auto t_mark = std::chrono::steady_clock::now();
	auto mark = [&](const char *what) {
		if (!verbose)
			return;
		HIPCHECK(hipStreamSynchronize(c->stream));
		const auto now = std::chrono::steady_clock::now();
		fprintf(stderr, "gdon3d-hip: relaxation set-up, %s: %.3f s\n", what,
			std::chrono::duration<double>(now - t_mark).count());
		t_mark = now;
	};
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
		c->items, c->nitems, c->geom, c->cls, c->kp, dt, c->dinv, c->bt);
	HIPCHECK(hipGetLastError());
	mark("items and block inverses");
	/* lagged (cycle-breaking) inflow faces: capacitance form */
	c->nlagged = 0;
	for (size_t k = 0; k < (size_t)ne * 4 * nv; k++)
		c->nlagged += cls[k] == 3;
	c->lag_last = c->lag_max = 0;
	sweep_setup(c);
	{
		/* the last item of each element in the level order (descriptor
		 * pad bit 0): with the relaxation by the element's last item it
		 * relaxes every velocity of the element */
		std::vector<int> last(ne, -1);
		for (int k = 0; k < c->nitems; k++)
			last[items[k].y] = k;
		int *dlast = (int *)dmalloc(sizeof(int) * (size_t)ne);
		HIPCHECK(hipMemcpy(dlast, last.data(), sizeof(int) * (size_t)ne,
				   hipMemcpyHostToDevice));
		k_flag_last<<<nblocks(ne, 256), 256, 0, c->stream>>>(c->desc, dlast, ne);
		HIPCHECK(hipGetLastError());
		HIPCHECK(hipStreamSynchronize(c->stream));
		HIPCHECK(hipFree(dlast));
	}
	mark("item descriptors and residency");
	const int mode = c->sweep_mode;
	c->sweep_mode = mode == 0 ? 0 : 1; /* the capacitance set-up sweeps by items */
	if (c->nlagged) {
		const size_t blen = (size_t)c->m * c->neq;
		c->wx = dmalloc(c->rsz * blen);
		cap_setup(c);
		mark("upwind cycles");
	}
	if (mode != 0 && mode != 1) {
		cluster_setup(c, cls, items_ve, level_ptr, nlevels);
		mark("cluster plan");
	}
	if (mode == 2 && !c->ncl) {
		fprintf(stderr, "gdon3d-hip: GDN_SWEEP=cluster, but the mesh gives no cluster plan\n");
		abort();
	}
	c->sweep_mode = mode;
	relax_choose(c);
	mark("form chosen by timing");
}

/* ------------------------------------------------------------------ */
/* Persistent sweep: hand-off of values between workgroups               */
/* ------------------------------------------------------------------ */
/*
 * Each published value is one word of its own precision (binary64: 64
 * bits, binary32: 32 bits) whose lowest mantissa bit carries the parity of
 * the sweep's epoch (the value is rounded once more by that: at most one
 * unit in the last place). Words are written and read with relaxed
 * agent-scope atomics, single-copy atomic and written through the per-XCD
 * L2 to the coherent memory side, so a reader that sees the current parity
 * holds the whole value: no separate flag, no fence, no wait for the store
 * on the producer side, and the consumer polls the data it needs directly.
 * Every item that publishes does so in every sweep, so a word holds either
 * the previous sweep's value (the other parity) or this one's. The buffer
 * starts zeroed (parity 0) and the first epoch is 1; after a sweep in which
 * some items did not publish (the cluster sweep) and before one that reads
 * every item, the buffer is zeroed and the epochs restart.
 */
typedef unsigned long long u64;

struct ItemDesc {
	double cf[4]; /* s_f dt / det on the inflow faces, 0 elsewhere */
	int e, v;
	int flags; /* bit f: inflow face f; bit 4 + f: inflow from the ghosts;
		    * bit 29: the block inverse is refined in the sweep from the
		    * network's prediction (KinPar lb_p, lb_x), else it is the
		    * stored exact one; bits 30-31: its refinement steps - 2;
		    * bit 8 + f: lagged inflow (cycle broken; value given in xl);
		    * cluster order only, bits 12 + 4 f .. 15 + 4 f: where the
		    * trace of face f comes from, inflow face x (0..7) of the
		    * cluster from outside, or 8 + the in-cluster producer;
		    * bit 28: a consumer lies in another cluster (the item's
		    * values go to the hand-off words) */
	int pad; /* cluster order: the item's level-order index; level order:
		  * 1 on the last item of its element (the one that relaxes the
		  * element when the relaxation is done by the last item) */
	int nb[4]; /* neighbour element, or the boundary face index */
	int map[4]; /* 4 bits per face node: the neighbour's local node */
};

#define SWEEP_SPIN_LIMIT (1 << 20) /* polls before a wait is declared stuck */


__global__ void k_item_desc(const int2 *__restrict__ items, int nitems,
			    const double *__restrict__ geom,
			    const int4 *__restrict__ nbr,
			    const int4 *__restrict__ nbmap,
			    const uint8_t *__restrict__ cls, KinPar kp, double dt_,
			    ItemDesc *__restrict__ desc, BatchTab bt)
{
	const int it = blockIdx.x * blockDim.x + threadIdx.x;
	if (it >= nitems)
		return;
	const int v = items[it].x, e = items[it].y;
	const double dt = bt.ninst ? bt.dt[bt.einst[e]] : dt_;
	double ct[9];
	for (int k = 0; k < 9; k++)
		ct[k] = geom[10L * e + k];
	const double sc = dt / geom[10L * e + 9];
	const int4 n4 = nbr[e], m4 = nbmap[e];
	const int nbv[4] = { n4.x, n4.y, n4.z, n4.w };
	const int mpv[4] = { m4.x, m4.y, m4.z, m4.w };
	ItemDesc d;
	d.e = e;
	d.v = v;
	d.flags = 0;
	d.pad = 0;
	for (int f = 0; f < 4; f++) {
		const uint8_t cl = cls[(4L * e + f) * kp.nv + v];
		const bool in = cl == 2 || cl == 3;
		d.cf[f] = in ? face_coef(ct, f, kp.vi + 3 * v) * sc : 0.0;
		d.flags |= in ? (nbv[f] >= 0 ? 1 << f : (1 << f) | (1 << (4 + f))) : 0;
		d.flags |= cl == 3 ? 1 << (8 + f) : 0;
		d.nb[f] = nbv[f] >= 0 ? nbv[f] : -1 - nbv[f];
		d.map[f] = mpv[f];
	}
	desc[it] = d;
}

/* Minv L_f S: the inverse mass applied to the lift of the face mass of
 * face f, [f][m][b] (exact rationals; the face mass rounded as the
 * reference rounds it) */
__constant__ static const double c_minv_lift[4][10][6] = T10_MINV_LIFT_INIT;

/* row iw of maxwell_qM(w), by selects (iw differs between lanes) */
__device__ inline void maxwell_q_row(const double w[6], int iw, double r[3])
{
	double q[6][3];
	maxwell_qM(w, q);
	r[0] = r[1] = r[2] = 0.0;
#pragma unroll
	for (int k = 0; k < 6; k++)
		if (iw == k) {
			r[0] = q[k][0];
			r[1] = q[k][1];
			r[2] = q[k][2];
		}
}

/*
 * The node-local relaxation f <- omega f_eq(w(f)) + (1 - omega) f of value
 * (iw, i) = (t / 10, t % 10) of velocity v, from that value of every
 * velocity (fr[0..NV-1]); the node's six w components come from lanes i,
 * 10 + i, ... by shuffles, so every lane of the wave takes part. The same
 * arithmetic as k_relax (src/dg/dg.c:109-132, src/simulation/simulation.c:
 * 42-54).
 */
template <int NV, typename R>
__device__ inline double relax_value(const KinPar &kp, const R (&fr)[NV], int v, double ie)
{
	const int t = threadIdx.x % 64;
	const int tt = t < 60 ? t : 59, iw = tt / 10, i = tt % 10;
	double s = 0.0, fo = 0.0;
#pragma unroll
	for (int u = 0; u < NV; u++) {
		s += (double)fr[u];
		fo = u == v ? (double)fr[u] : fo;
	}
	double w[6];
#pragma unroll
	for (int q = 0; q < 6; q++)
		w[q] = __shfl(s, 10 * q + i, 64);
	/* the flux is taken on (D / eps, B); ie = 1 exactly without permittivity */
#pragma unroll
	for (int q = 0; q < 3; q++)
		w[q] *= ie;
	double r[3];
	maxwell_q_row(w, iw, r);
	/* the model's scalars taken where they are used (not values derived
	 * once and held in vector registers across the item loop) */
	double om = kp.omega, lam = kp.lambda;
	int nvk = kp.nv;
	asm volatile("" : "+s"(om), "+s"(lam), "+s"(nvk));
	if (kp.kind == GDH_KIN_D3Q4) {
		const double l2 = lam * lam;
		const double *vv = kp.vi + 3 * v;
		const double qv = r[0] * vv[0] + r[1] * vv[1] + r[2] * vv[2];
		return om * ((s + qv / l2) / nvk) + (1. - om) * fo;
	}
	const int k = v % 3;
	const double a = s / 6., b = r[k] / (2. * lam);
	return om * (v < 3 ? a + b : a - b) + (1. - om) * fo;
}

/* Per-block operator table in LDS */
struct SweepTables {
	double ML[4][10][6];
};

/*
 * One prepared item of the item and level sweeps: everything that does not
 * depend on the new values of its inflow neighbours. Slot s < nslot is the
 * s-th interior, not lagged, inflow face: its face index, coefficient c_f,
 * neighbour and node map.
 */
struct ItemBuf {
	double D[10][11]; /* D^-1 (rows padded: the ten rows read at once fall
			   * in distinct banks) */
	double Z[60]; /* after the wait: 2 f - sum_f c_f ML_f u_f */
	double cf[3];
	int v, e, nslot, last; /* last: the element's last item (pad bit 0) */
	int nb[3], map[3], face[3];
};

/* Per-block scratch of the item and level sweeps: one given trace at a time
 * (and the learned refinement's eight numbers); with learned inverses also
 * the refinement's X and D (200 values). */
template <bool LB> struct ItemScratch {
	double S[60];
	double L[LB ? 200 : 1];
};

/* Per-block scratch of the cluster sweep's items: D^-1, and 60 values (one
 * given trace at a time, then Z). */
struct PrepScratch {
	double D[10][11]; /* D^-1, rows padded */
	double S[60];
};

/* The global loads of one item's preparation, issued ahead of their use
 * and kept as loaded (no arithmetic on them until prep_compute, so that
 * none of them waits for another): lane t < 60 its value (iw, i) = (t / 10,
 * t % 10) of the item's velocity (fr[0]) or, when the state still has to
 * be relaxed, of every velocity (fr[0..NV-1]); D^-1 entries t and 64 + t.
 * The traces of ghost and lagged inflow faces (boundary items and cycles
 * only) are read by prep_compute itself. */
template <typename R, int NV> struct PrepLoads {
	R fr[NV];
	double d0, d1; /* exact D^-1 entries t, 64 + t; learned: d0 = number t < 8 */
	unsigned pw; /* learned: prediction entries 2 t, 2 t + 1 (t < 50) */
};

#define LB_FLAG (1 << 29)

/*
 * Transport of item (v, e), split around the dependency wait. With
 * D = I + theta dt/det Minv F_ee(v) (its inverse stored per item), theta = 1/2
 * and u_f the inflow trace ((1-theta) f_nb + theta sol_nb, or the ghost
 * value), one block row of M1 sol = M2 f (src/dg/relaxation.c:64-110) reads
 *   sol = D^-1 [ f - dt/det Minv ((1-theta) F_ee f + sum_f s_f L_f S u_f) ]
 *       = D^-1 (2 f - sum_f c_f ML_f u_f) - f,   c_f = s_f dt/det,
 * since (dt/det) Minv (1-theta) F_ee = D - I for theta = 1/2, with
 * ML_f = Minv L_f S. Every item publishes w = (1-theta) f + theta sol, which
 * is u_f for its downstream neighbours. Before the wait: z0 = 2 f -
 * sum_g c_g ML_g u_g over the ghost and lagged inflow faces g, and D^-1 into
 * the item's buffer (prep_issue: the loads; prep_compute: LDS only). After
 * it (sweep_finish):
 *   sol = D^-1 (z0 - sum_s c_s ML_s w_nb,s) - f
 * over the interior inflow slots s. This direct form does less work per item
 * than forming D^-1 ML_s before the wait (the 32^3 sweep is bound by how
 * fast the workgroups turn items over, not by the length of the dependency
 * chain: 0.754 -> 0.722 ms with the relaxation separate, 0.873 -> 0.845 ms
 * fused). Lanes: t = 10 iw + i (field, node) for vectors, t = 6 q + b
 * (field, face node) for traces.
 */
/* state index of the inflow trace value of lane t < 36 on face f */
__device__ inline long trace_index(const ItemDesc &d, int f, int t, int ne, int neq, int nv)
{
	if (d.flags & (16 << f)) /* ghost values of boundary face nb[f] */
		return 60L * nv * ne + 6L * d.v * (neq - 10L * ne) + 36L * d.nb[f] + t;
	const int q = t / 6, b = t % 6;
	return ((long)d.v * ne + d.nb[f]) * 60 + 10 * q + ((d.map[f] >> (4 * b)) & 15);
}

/* inflow faces whose trace is known before the wait: ghost or lagged */
__device__ inline int given_faces(const ItemDesc &d)
{
	return d.flags & ((d.flags >> 4) | (d.flags >> 8)) & 15;
}

/* interior inflow faces, not lagged: the slots */
__device__ inline int slot_faces(const ItemDesc &d)
{
	return d.flags & ~(d.flags >> 4) & ~(d.flags >> 8) & 15;
}

/* the lane index, opaque to the compiler: per-lane addresses are then formed
 * where they are used (uniform base + 32-bit lane offset) instead of being
 * hoisted out of the item loop as 64-bit values that occupy registers */
__device__ inline int lane_id()
{
	int t = threadIdx.x;
	asm volatile("" : "+v"(t));
	return t;
}

/* a value the whole wave holds (read from LDS), made scalar */
__device__ inline int uni(int x)
{
	return __builtin_amdgcn_readfirstlane(x);
}

__device__ inline double uni(double x)
{
	const long long b = __double_as_longlong(x);
	const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
	const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
	return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <typename R, int NV, bool LB = false>
__device__ inline void prep_issue(int it, const ItemDesc &d, const R *__restrict__ fo,
				  const double *__restrict__ dinv, int ne, int relax_in,
				  const KinPar &kp, PrepLoads<R, NV> &P)
{
	const int t = lane_id();
	if (relax_in) {
#pragma unroll
		for (int u = 0; u < NV; u++)
			P.fr[u] = t < 60 ? fo[((long)u * ne + d.e) * 60 + t] : (R)0;
	} else {
		/* the item's own values, read by this item only: non-temporal */
		P.fr[0] = t < 60 ? __builtin_nontemporal_load(fo + ((long)d.v * ne + d.e) * 60 + t)
				 : (R)0;
#pragma unroll
		for (int u = 1; u < NV; u++)
			P.fr[u] = (R)0;
	}
	if (LB && (d.flags & LB_FLAG)) {
		/* the prediction (200 B) and the block's numbers instead of the
		 * stored inverse (800 B) */
		P.pw = t < 50 ? kp.lb_p[50L * it + t] : 0u;
		P.d0 = t < 8 ? kp.lb_x[8L * it + t] : 0.0;
		P.d1 = 0.0;
		return;
	}
	/* read once per sweep: non-temporal, which leaves L2 to the hand-off
	 * words (32^3 binary64 sweep 5 % faster in both relaxation forms) */
	const double *D = dinv + 100L * it;
	P.d0 = __builtin_nontemporal_load(D + t);
	P.d1 = t < 36 ? __builtin_nontemporal_load(D + 64 + t) : 0.0;
	P.pw = 0u;
}

/*
 * D^-1 of a learned item into Dinv (LDS, rows padded to 11): D from its
 * seven numbers, X0 = I + the network's prediction, then nsteps
 * Newton-Schulz steps R = I - D X, X <- X + X R, every entry a sum over
 * l = 0..9 in order (the set-up check does the same arithmetic). Lane t
 * owns entries t and (t < 36) 64 + t. X and D live in scr (200 doubles of
 * the block's scratch), the numbers in xs (8 doubles).
 */
template <typename R, int NV>
__device__ inline void learned_inverse(const PrepLoads<R, NV> &P, double (*Dinv)[11],
				       double *scr, double *xs, int nsteps)
{
	const int t = threadIdx.x;
	double *X = scr, *Dm = scr + 100;
	const int ne_ = t < 36 ? 2 : 1; /* entries t and 64 + t */
	if (t < 8)
		xs[t] = P.d0;
	if (t < 50) {
		const int a = 2 * t, b = 2 * t + 1;
		X[a] = lb_pred((unsigned short)(P.pw & 0xffffu)) + (a / 10 == a % 10 ? 1.0 : 0.0);
		X[b] = lb_pred((unsigned short)(P.pw >> 16)) + (b / 10 == b % 10 ? 1.0 : 0.0);
	}
	__syncthreads();
#pragma unroll 1
	for (int m = 0; m < ne_; m++) {
		const int e = t + 64 * m;
		Dm[e] = lb_entry(xs, e / 10, e % 10);
	}
	__syncthreads();
#pragma unroll 1
	for (int step = 0; step < nsteps; step++) {
		/* R = I - D X, kept in Dinv */
#pragma unroll 1
		for (int m = 0; m < ne_; m++) {
			const int e = t + 64 * m, i = e / 10, j = e % 10;
			double r = i == j ? 1.0 : 0.0;
#pragma unroll
			for (int l = 0; l < 10; l++)
				r -= Dm[10 * i + l] * X[10 * l + j];
			Dinv[i][j] = r;
		}
		__syncthreads();
		/* X <- X + X R */
		double xn[2];
#pragma unroll 1
		for (int m = 0; m < ne_; m++) {
			const int e = t + 64 * m, i = e / 10, j = e % 10;
			double x = X[e];
#pragma unroll
			for (int l = 0; l < 10; l++)
				x += X[10 * i + l] * Dinv[l][j];
			xn[m] = x;
		}
		__syncthreads();
#pragma unroll 1
		for (int m = 0; m < ne_; m++) {
			const int e = t + 64 * m;
			if (step + 1 < nsteps)
				X[e] = xn[m];
			else
				Dinv[e / 10][e % 10] = xn[m];
		}
		__syncthreads();
	}
}

/* u = the ghost value, or on a lagged face (1 - theta) f_nb + theta x_nb
 * with x_nb from xl (none when xl is null: both halves of a lagged trace
 * are given values). With relax_in the item's values are relaxed first (the
 * state holds the previous transport's output) and rounded to R as a stored
 * state would be. Fills B (D^-1 and the slots) and fv (the lane's value of
 * the item, t < 60) and returns the lane's z0 = 2 f - sum over the given
 * faces. */
template <typename R, int NV, bool LB = false>
__device__ inline double prep_compute(const ItemDesc &d, const PrepLoads<R, NV> &P,
				      const R *__restrict__ fo, const R *__restrict__ xl,
				      int ne, int neq, int nv, int relax_in, const KinPar &kp,
				      const SweepTables &T, ItemBuf &B, ItemScratch<LB> &W, R &fv)
{
	const int t = threadIdx.x;
	fv = relax_in ? (R)relax_value<NV>(kp, P.fr, d.v, kp.ieps ? kp.ieps[d.e] : 1.0) : P.fr[0];
	const int gf = given_faces(d), sf = slot_faces(d);
	const int iw = t / 10, i = t % 10;
	if (LB && (d.flags & LB_FLAG)) {
		learned_inverse(P, B.D, W.L, W.S, ((unsigned)d.flags >> 30) + 2);
	} else {
		B.D[t / 10][t % 10] = P.d0;
		if (t < 36)
			B.D[(64 + t) / 10][(64 + t) % 10] = P.d1;
	}
	if (t == 0) {
		int s = 0;
		B.v = d.v;
		B.e = d.e;
		B.last = d.pad & 1;
		for (int f = 0; f < 4; f++) {
			if (!(sf & (1 << f)))
				continue;
			B.nb[s] = d.nb[f];
			B.map[s] = d.map[f];
			B.face[s] = f;
			B.cf[s] = d.cf[f];
			s++;
		}
		B.nslot = s;
	}
	double y = 0.0;
#pragma unroll
	for (int f = 0; f < 4; f++) {
		if (!(gf & (1 << f)))
			continue;
		if (t < 36) {
			const long g = trace_index(d, f, lane_id(), ne, neq, nv);
			double u = (double)fo[g];
			if (!(d.flags & (16 << f))) {
				u = (1.0 - THETA) * u;
				if (xl)
					u += THETA * (double)xl[g];
			}
			W.S[t] = u;
		}
		__syncthreads();
		if (t < 60) {
			double s = 0.0;
#pragma unroll
			for (int b = 0; b < 6; b++)
				s += T.ML[f][i][b] * W.S[6 * iw + b];
			y += d.cf[f] * s;
		}
		__syncthreads();
	}
	__syncthreads();
	return 2.0 * (double)fv - y;
}

__device__ inline u64 ll_load(const u64 *p)
{
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline unsigned ll_load(const unsigned *p)
{
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename W> __device__ inline void ll_store(W *p, W x)
{
	__hip_atomic_store(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* Hand-off words of the 60 values of one item (v, e): 60 words W at
 * ll + 60 (v ne + e) in units of W, value k at word k. */
template <typename R> struct Handoff;
template <> struct Handoff<double> {
	using W = u64;
	__device__ static W tag(double x, unsigned epoch)
	{
		return ((u64)__double_as_longlong(x) & ~1ull) | (u64)(epoch & 1u);
	}
	__device__ static void put(W *item, int k, double x, unsigned epoch)
	{
		ll_store(item + k, tag(x, epoch));
	}
	__device__ static bool fresh(W w, unsigned epoch)
	{
		return (unsigned)(w & 1ull) == (epoch & 1u);
	}
	__device__ static double value(W w)
	{
		return __longlong_as_double((long long)w);
	}
};
template <> struct Handoff<float> {
	using W = unsigned;
	__device__ static W tag(float x, unsigned epoch)
	{
		return (__float_as_uint(x) & ~1u) | (epoch & 1u);
	}
	__device__ static void put(W *item, int k, float x, unsigned epoch)
	{
		ll_store(item + k, tag(x, epoch));
	}
	__device__ static bool fresh(W w, unsigned epoch)
	{
		return (w & 1u) == (epoch & 1u);
	}
	__device__ static double value(W w)
	{
		return (double)__uint_as_float(w);
	}
};

/* hand-off words of item (v, e) */
template <typename R>
__device__ inline typename Handoff<R>::W *ll_item(u64 *ll, int v, int e, int ne)
{
	return reinterpret_cast<typename Handoff<R>::W *>(ll) + 60 * ((long)v * ne + e);
}

/* the value as a consumer of the hand-off reads it (tag bit included) */
template <typename R> __device__ inline R published(R x, unsigned epoch)
{
	const typename Handoff<R>::W w = Handoff<R>::tag(x, epoch);
	return (R)Handoff<R>::value(w);
}

/*
 * Lane s (< nslot) checks the first word of one value of slot s's
 * neighbour trace (field 0, face node 0): the full trace is read again only
 * once every slot shows the current epoch there, so that waiting blocks do
 * not flood the memory side with coherent loads.
 */
template <typename R>
__device__ inline bool ll_sentinels(const ItemBuf &B, u64 *ll, int ne, unsigned epoch)
{
	const int t = threadIdx.x;
	bool wait = false;
	if (t < B.nslot)
		wait = !Handoff<R>::fresh(ll_load(ll_item<R>(ll, B.v, B.nb[t], ne) + (B.map[t] & 15)),
					  epoch);
	return __ballot(wait) == 0;
}

/*
 * One pass over the inflow values that lane t = 6 q + b (t < 36) still
 * misses (bit s of pending): each complete one goes to X[s][t]. Returns the
 * slots still pending. All loads are issued before any is decoded.
 */
template <typename R>
__device__ inline unsigned ll_poll(const ItemBuf &B, u64 *ll, int ne, unsigned epoch,
				   unsigned pending, double (&X)[3][36])
{
	const int t = lane_id();
	const int q = t / 6, b = t % 6;
	typename Handoff<R>::W w[3];
	const int v = uni(B.v);
#pragma unroll
	for (int s = 0; s < 3; s++) {
		w[s] = 0;
		if (pending & (1u << s))
			w[s] = ll_load(ll_item<R>(ll, v, uni(B.nb[s]), ne) + q * 10 +
				       ((uni(B.map[s]) >> (4 * b)) & 15));
	}
#pragma unroll
	for (int s = 0; s < 3; s++) {
		if ((pending & (1u << s)) && Handoff<R>::fresh(w[s], epoch)) {
			X[s][t] = Handoff<R>::value(w[s]);
			pending &= ~(1u << s);
		}
	}
	return pending;
}

/* bounded wait: true once the limit is passed or another wave timed out */
__device__ inline bool spin_expired(int spins, int *status)
{
	if ((spins & 255) != 0)
		return false;
	if (spins > SWEEP_SPIN_LIMIT ||
	    __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
		if (threadIdx.x == 0)
			__hip_atomic_store(status, 1, __ATOMIC_RELAXED,
					   __HIP_MEMORY_SCOPE_AGENT);
		return true;
	}
	return false;
}

/* sol = D^-1 (z0 - sum_s c_s ML_s X_s) - fv with X[s] the published trace w
 * of slot s's neighbour (fv, z0: the lane's old value and z0); sol goes to
 * fn (unless the element's last item relaxes and stores it, LF) and
 * w = (1 - theta) fv + theta sol to the hand-off words, each rounded once to
 * R. Returns the lane's sol. */
template <typename R, bool LF = false>
__device__ inline double sweep_finish(ItemBuf &B, const double (&X)[3][36], R fv, double z0,
				      R *__restrict__ fn, u64 *ll, int ne, unsigned epoch,
				      const SweepTables &T)
{
	const int t = lane_id();
	const int iw = t / 10, i = t % 10;
	const int nslot = uni(B.nslot);
	if (t < 60) {
		double y = 0.0;
		for (int s = 0; s < nslot; s++) {
			const int f = uni(B.face[s]);
			double sm = 0.0;
#pragma unroll
			for (int b = 0; b < 6; b++)
				sm += T.ML[f][i][b] * X[s][6 * iw + b];
			y += B.cf[s] * sm;
		}
		B.Z[t] = z0 - y;
	}
	__syncthreads();
	double sol = 0.0;
	if (t < 60) {
		double sm = 0.0;
#pragma unroll
		for (int l = 0; l < 10; l++)
			sm += B.D[i][l] * B.Z[10 * iw + l];
		sol = sm - (double)fv;
		const int v = uni(B.v), e = uni(B.e);
		/* the hand-off word first: the downstream items wait for it */
		const R wr = (R)((1.0 - THETA) * (double)fv + THETA * sol);
		Handoff<R>::put(ll_item<R>(ll, v, e, ne), t, wr, epoch);
		/* no consumer reads it during the sweep: non-temporal (separate
		 * relaxation 2 % faster, fused unchanged) */
		if (!LF)
			__builtin_nontemporal_store((R)sol, fn + ((long)v * ne + e) * 60 + t);
	}
	__syncthreads();
	return sol;
}

/*
 * Relaxation by the element's last item (LF). The other items of the
 * element (its other velocities u) come earlier in the level order, so each
 * is done once its hand-off words show this sweep's parity; they hold
 * w_u = (1 - theta) f_u + theta x_u, and with f_u its old value (the
 * sweep's input, unchanged during the sweep) its transport output is
 * x_u = (w_u - (1 - theta) f_u) / theta, rounded to R as a stored value
 * would be (within three units in the last place of x_u, from the rounding
 * and the parity bit of w_u). With the item's own output x_v every velocity
 * of the element is relaxed, with the arithmetic of k_relax
 * (src/dg/dg.c:109-132), and stored. The wait is bounded like the others.
 */
template <typename R, int NV>
__device__ inline void relax_last(int v, int e, double sol, const R *__restrict__ fo,
				  R *__restrict__ fn, u64 *ll, int ne, unsigned epoch,
				  const KinPar &kp, int *status)
{
	const int t = lane_id();
	const int tc = min(t, 59), iw = tc / 10, i = tc % 10;
	typename Handoff<R>::W w[NV];
	R fu[NV];
#pragma unroll
	for (int u = 0; u < NV; u++) {
		w[u] = ll_load(ll_item<R>(ll, u, e, ne) + tc);
		fu[u] = fo[((long)u * ne + e) * 60 + tc];
	}
	unsigned pend = 0;
#pragma unroll
	for (int u = 0; u < NV; u++)
		if (u != v && !Handoff<R>::fresh(w[u], epoch))
			pend |= 1u << u;
	int spins = 0;
	while (__ballot(pend != 0)) {
		__builtin_amdgcn_s_sleep(1);
#pragma unroll
		for (int u = 0; u < NV; u++)
			if (pend & (1u << u)) {
				w[u] = ll_load(ll_item<R>(ll, u, e, ne) + tc);
				if (Handoff<R>::fresh(w[u], epoch))
					pend &= ~(1u << u);
			}
		if (spin_expired(++spins, status))
			break;
	}
	double x[NV];
#pragma unroll
	for (int u = 0; u < NV; u++)
		x[u] = u == v ? (double)(R)sol
			      : (double)(R)((Handoff<R>::value(w[u]) - (1.0 - THETA) * (double)fu[u]) /
					    THETA);
	double s = 0.0;
#pragma unroll
	for (int u = 0; u < NV; u++)
		s += x[u];
	double wn[6];
#pragma unroll
	for (int q = 0; q < 6; q++)
		wn[q] = __shfl(s, 10 * q + i, 64);
	const double ie = kp.ieps ? kp.ieps[e] : 1.0;
#pragma unroll
	for (int q = 0; q < 3; q++)
		wn[q] *= ie;
	double r[3];
	maxwell_q_row(wn, iw, r);
	const double om = kp.omega;
	if (t < 60) {
		R *p = fn + (long)e * 60 + t;
		const long vs = 60L * ne;
		if (kp.kind == GDH_KIN_D3Q4) {
			const double l2 = kp.lambda * kp.lambda;
#pragma unroll
			for (int u = 0; u < NV; u++) {
				const double *vv = kp.vi + 3 * u;
				const double qv = r[0] * vv[0] + r[1] * vv[1] + r[2] * vv[2];
				__builtin_nontemporal_store(
					(R)(om * ((s + qv / l2) / kp.nv) + (1. - om) * x[u]), p + u * vs);
			}
		} else {
#pragma unroll
			for (int u = 0; u < NV; u++) {
				const double a = s / 6., b = r[u % 3] / (2. * kp.lambda);
				__builtin_nontemporal_store(
					(R)(om * (u < 3 ? a + b : a - b) + (1. - om) * x[u]), p + u * vs);
			}
		}
	}
}

__device__ inline void load_tables(SweepTables &T)
{
	for (int k = threadIdx.x; k < 240; k += blockDim.x)
		(&T.ML[0][0][0])[k] = (&c_minv_lift[0][0][0])[k];
	__syncthreads();
}

/* One upwind level per launch, one item per 64-lane block; the inflow
 * traces are the hand-off words of earlier levels (a word without the
 * sweep's epoch is reported in *status). */
template <typename R, int NV, bool LB = false>
__global__ __launch_bounds__(64) void k_sweep_level(
	const ItemDesc *__restrict__ desc, int begin, const R *__restrict__ fo,
	R *__restrict__ fn, const R *__restrict__ xl, const double *__restrict__ dinv,
	int nv, int ne, int neq, u64 *ll, unsigned epoch_arg,
	const unsigned *__restrict__ epoch_dev, int *status, int relax_in, KinPar kp)
{
	__shared__ SweepTables T;
	__shared__ ItemBuf B;
	__shared__ ItemScratch<LB> W;
	__shared__ double X[3][36];
	load_tables(T);
	const unsigned epoch = epoch_dev ? *epoch_dev : epoch_arg;
	const int it = begin + blockIdx.x;
	const ItemDesc d = desc[it];
	PrepLoads<R, NV> P;
	prep_issue<R, NV, LB>(it, d, fo, dinv, ne, relax_in, kp, P);
	R fv;
	const double z0 =
		prep_compute<R, NV, LB>(d, P, fo, xl, ne, neq, nv, relax_in, kp, T, B, W, fv);
	const int t = threadIdx.x;
	const unsigned pending =
		ll_poll<R>(B, ll, ne, epoch, t < 36 ? (1u << B.nslot) - 1u : 0u, X);
	if (__ballot(pending != 0) && t == 0)
		__hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	__syncthreads();
	sweep_finish<R>(B, X, fv, z0, fn, ll, ne, epoch, T);
}

/* an ItemDesc held one dword per lane (lane k: dword k), made uniform */
/* descriptor pad bit 0 on the item last[e] of every element e */
__global__ void k_flag_last(ItemDesc *__restrict__ desc, const int *__restrict__ last, int ne)
{
	const int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e < ne && last[e] >= 0)
		desc[last[e]].pad = 1;
}

__device__ inline ItemDesc desc_from_lanes(int x)
{
	static_assert(sizeof(ItemDesc) == 80, "ItemDesc is 20 dwords");
	int w[20];
#pragma unroll
	for (int k = 0; k < 20; k++)
		w[k] = __builtin_amdgcn_readlane(x, k);
	ItemDesc d;
	__builtin_memcpy(&d, w, sizeof d);
	return d;
}

/*
 * The whole sweep in one launch. Items are in level order; block b takes
 * items b, b + G, b + 2G, ... and polls the values of its inflow neighbours
 * (always earlier items). All G blocks are co-resident, so the earliest
 * unfinished item always has its inputs ready and the sweep cannot
 * deadlock; a bounded spin reports a timeout in *status instead of hanging.
 * The loads that prepare the next item are issued before the current item's
 * poll (one round trip for both), and while the current item's inputs are
 * not ready the block prepares the next item. Item descriptors are read one
 * item further ahead.
 *
 * With disc != nullptr the launch only measures residency: every block
 * records a 100 us busy interval, and the largest number of overlapping
 * intervals is a set of blocks that were resident together.
 */
/* the next index of the sweep's item queue (lane 0's request; the caller
 * makes it uniform where it needs it) */
__device__ inline unsigned queue_take(unsigned *q)
{
	unsigned r = 0;
	if (threadIdx.x == 0)
		r = __hip_atomic_fetch_add(q, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	return r;
}

/*
 * With QU the blocks take their items from counters instead of in the
 * order b, b + G, ...: the items are dealt into S = min(QUEUE_SHARDS, G)
 * shards (item i to shard i mod S), block b serves shard b mod S, and the
 * blocks of a shard take its items in level order from the shard's counter
 * (one counter per cache line; slot `slot` of two, zeroed by the launch
 * before, which zeroes the other slot for the next one). A block that is
 * not resident takes nothing, and a block takes its items in increasing
 * order holding at most its current one and the two after it, so the
 * earliest unfinished item of a shard is a running block's current item
 * as long as one block of every shard runs: the first S blocks of the
 * grid suffice, not the whole grid. Each block knows two items ahead: it
 * prepares the next one, loads the descriptor of the one after, and asks
 * the counter for a third.
 */
#define QUEUE_SHARDS 64
#define QUEUE_STRIDE 32 /* unsigned words between counters: 128 B */
#define QUEUE_BYTES (2 * QUEUE_SHARDS * QUEUE_STRIDE * sizeof(unsigned))
template <typename R, bool FULL_POLL, int NV, bool LB = false, bool QU = false, bool LF = false>
__global__ __launch_bounds__(64, LB ? 4 : 5) void k_sweep_persistent(
	const ItemDesc *__restrict__ desc, int nitems, const R *__restrict__ fo,
	R *__restrict__ fn, const R *__restrict__ xl, const double *__restrict__ dinv,
	int nv, int ne, int neq, u64 *ll, unsigned epoch_arg,
	const unsigned *__restrict__ epoch_dev, int *status, u64 *disc, int relax_in,
	KinPar kp, unsigned *queue, int slot)
{
	__shared__ SweepTables T;
	__shared__ ItemBuf B[2];
	__shared__ ItemScratch<LB> W;
	__shared__ double X[3][36];
	const int t = threadIdx.x;
	if (disc) {
		if (t == 0) {
			const u64 t0 = __builtin_amdgcn_s_memrealtime(); /* 100 MHz */
			u64 t1 = t0;
			for (int n = 0; n < (1 << 26) && t1 - t0 < 10000; n++)
				t1 = __builtin_amdgcn_s_memrealtime();
			disc[2 * blockIdx.x] = t0;
			disc[2 * blockIdx.x + 1] = t1;
		}
		return;
	}
	load_tables(T);
	/* the sweep's epoch: an argument, or (captured graphs) device memory */
	const unsigned epoch = epoch_dev ? *epoch_dev : epoch_arg;
	const int G = gridDim.x;
	const int S = min(QUEUE_SHARDS, G), shard = blockIdx.x % S;
	unsigned *q = QU ? queue + (slot * QUEUE_SHARDS + shard) * QUEUE_STRIDE : nullptr;
	if (QU && blockIdx.x == 0 && t < QUEUE_SHARDS)
		__hip_atomic_store(queue + ((1 - slot) * QUEUE_SHARDS + t) * QUEUE_STRIDE, 0u,
				   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	/* the current item, the next one and the one after it */
	int it, next, nn;
	if (QU) {
		it = shard + S * (int)__builtin_amdgcn_readfirstlane(queue_take(q));
		next = it < nitems ? shard + S * (int)__builtin_amdgcn_readfirstlane(queue_take(q))
				   : nitems;
		nn = next < nitems ? shard + S * (int)__builtin_amdgcn_readfirstlane(queue_take(q))
				   : nitems;
	} else {
		it = blockIdx.x;
		next = it + G;
		nn = next + G;
	}
	int k = 0;
	if (it >= nitems)
		return;
	R fv; /* the lane's old value and z0 of the current item */
	double z0;
	{
		const ItemDesc d = desc[it];
		PrepLoads<R, NV> P;
		prep_issue<R, NV, LB>(it, d, fo, dinv, ne, relax_in, kp, P);
		z0 = prep_compute<R, NV, LB>(d, P, fo, xl, ne, neq, nv, relax_in, kp, T, B[0], W,
					     fv);
	}
	ItemDesc dn; /* descriptor of the next item */
	if (next < nitems)
		dn = desc[next];
	for (; it < nitems; k ^= 1) {
		const bool has_next = next < nitems;
		PrepLoads<R, NV> P;
		if (has_next)
			prep_issue<R, NV, LB>(next, dn, fo, dinv, ne, relax_in, kp, P);
		/* descriptor of the item after next: one dword per lane (t < 20);
		 * with the queue, the index of the one after that */
		int dnn = 0;
		const int tl = lane_id();
		if (nn < nitems && tl < (int)(sizeof(ItemDesc) / 4))
			dnn = reinterpret_cast<const int *>(desc + nn)[tl];
		const unsigned nnn_req = QU && nn < nitems ? queue_take(q) : 0u;
		double nz0 = 0.0;
		R nfv = (R)0;
		unsigned pending = t < 36 ? (1u << uni(B[k].nslot)) - 1u : 0u;
		pending = ll_poll<R>(B[k], ll, ne, epoch, pending, X);
		if (__ballot(pending != 0)) {
			/* prepare the next item, then watch one word per slot */
			if (has_next)
				nz0 = prep_compute<R, NV, LB>(dn, P, fo, xl, ne, neq, nv, relax_in,
							      kp, T, B[k ^ 1], W, nfv);
			int spins = 0;
			if (!FULL_POLL) {
				bool ready = ll_sentinels<R>(B[k], ll, ne, epoch);
				while (!ready) {
					__builtin_amdgcn_s_sleep(2);
					ready = ll_sentinels<R>(B[k], ll, ne, epoch) ||
						spin_expired(++spins, status);
				}
				pending = ll_poll<R>(B[k], ll, ne, epoch, pending, X);
			}
			while (__ballot(pending != 0)) {
				__builtin_amdgcn_s_sleep(1);
				pending = ll_poll<R>(B[k], ll, ne, epoch, pending, X);
				if (spin_expired(++spins, status))
					break;
			}
			/* the finish at the highest wave priority: downstream items
			 * wait for its hand-off words (32^3 binary64 fused 3 %
			 * faster) */
			__builtin_amdgcn_s_setprio(3);
			__syncthreads(); /* X complete */
			const double sol = sweep_finish<R, LF>(B[k], X, fv, z0, fn, ll, ne, epoch, T);
			__builtin_amdgcn_s_setprio(0);
			/* the element's last item relaxes it now (deferring it to the
			 * next wait delays that item's finish: 1.02 vs 0.88 ms at 32^3) */
			if (LF && uni(B[k].last))
				relax_last<R, NV>(uni(B[k].v), uni(B[k].e), sol, fo, fn, ll, ne, epoch, kp,
						  status);
		} else {
			__builtin_amdgcn_s_setprio(3);
			__syncthreads(); /* X complete */
			const double sol = sweep_finish<R, LF>(B[k], X, fv, z0, fn, ll, ne, epoch, T);
			__builtin_amdgcn_s_setprio(0);
			if (LF && uni(B[k].last))
				relax_last<R, NV>(uni(B[k].v), uni(B[k].e), sol, fo, fn, ll, ne, epoch, kp,
						  status);
			if (has_next)
				nz0 = prep_compute<R, NV, LB>(dn, P, fo, xl, ne, neq, nv, relax_in,
							      kp, T, B[k ^ 1], W, nfv);
		}
		if (has_next) {
			fv = nfv;
			z0 = nz0;
		}
		const int nnn = !QU ? nn + G
				   : nn < nitems
					   ? shard + S * (int)__builtin_amdgcn_readfirstlane(nnn_req)
					   : nitems;
		if (nn < nitems)
			dn = desc_from_lanes(dnn);
		it = next;
		next = nn;
		nn = nnn;
	}
}

/* ------------------------------------------------------------------ */
/* Cluster sweep                                                          */
/* ------------------------------------------------------------------ */

/* dword w (uniform, below 128) of a cluster's descriptors held one dword
 * per lane: lane l holds dwords l (a) and 64 + l (b) */
__device__ inline int cl_dword(int a, int b, int w)
{
	const int x = __builtin_amdgcn_readlane(a, w & 63);
	const int y = __builtin_amdgcn_readlane(b, w & 63);
	return w < 64 ? x : y;
}

/* descriptor of item j (a compile-time index once unrolled) */
__device__ inline ItemDesc cl_desc(int a, int b, int j)
{
	int w[20];
#pragma unroll
	for (int k = 0; k < 20; k++) {
		const int x = 20 * j + k;
		w[k] = x < 64 ? __builtin_amdgcn_readlane(a, x) : __builtin_amdgcn_readlane(b, x - 64);
	}
	ItemDesc d;
	__builtin_memcpy(&d, w, sizeof d);
	return d;
}

/* descriptor of item j of a cluster, read field by field where it is used
 * (one readlane each, j a compile-time index once unrolled): the fields an
 * item does not use are never read */
struct DescLanes {
	int a, b, j;
	__device__ int dw(int k) const
	{
		const int x = 20 * j + k;
		return x < 64 ? __builtin_amdgcn_readlane(a, x) : __builtin_amdgcn_readlane(b, x - 64);
	}
	__device__ double cf(int f) const
	{
		return __hiloint2double(dw(2 * f + 1), dw(2 * f));
	}
	__device__ int e() const { return dw(8); }
	__device__ int v() const { return dw(9); }
	__device__ int flags() const { return dw(10); }
	__device__ int pad() const { return dw(11); }
	__device__ int nb(int f) const { return dw(12 + f); }
	__device__ int map(int f) const { return dw(16 + f); }
};

/* the descriptors of the n items from item i0 on, one dword per lane */
__device__ inline void cl_load_desc(const ItemDesc *__restrict__ cdesc, int i0, int n, int &a,
				    int &b)
{
	const int t = lane_id();
	const int *p = reinterpret_cast<const int *>(cdesc + i0);
	a = t < 20 * n ? p[t] : 0;
	b = 64 + t < 20 * n ? p[64 + t] : 0;
}

/*
 * One pass over the inflow faces x of the cluster in um (uniform) whose
 * trace lane t = 6 q + b (t < 36) still misses (bit x of pend): a complete
 * value goes to Xc[x][t]. Face x is face f of item j, byte x of the
 * header's words z, w holding 4 j + f. All loads are issued before any is
 * decoded. Returns the faces still pending in this lane.
 */
template <typename R>
__device__ inline unsigned cl_poll(int hz, int hw, int da, int db, unsigned um, unsigned pend,
				   u64 *ll, int v, int ne, unsigned epoch, double (*Xc)[36])
{
	const int t = lane_id();
	const int q = t / 6, b = t % 6;
	typename Handoff<R>::W w[CL_XMAX];
#pragma unroll
	for (int x = 0; x < CL_XMAX; x++) {
		w[x] = 0;
		if (!(um & (1u << x)))
			continue;
		const int jf = uni(((x < 4 ? hz : hw) >> (8 * (x & 3))) & 255);
		const int j = jf >> 2, f = jf & 3;
		const int nb = cl_dword(da, db, 20 * j + 12 + f);
		const int map = cl_dword(da, db, 20 * j + 16 + f);
		if (t < 36 && (pend & (1u << x)))
			w[x] = ll_load(ll_item<R>(ll, v, nb, ne) + q * 10 + ((map >> (4 * b)) & 15));
	}
#pragma unroll
	for (int x = 0; x < CL_XMAX; x++) {
		if ((um & (1u << x)) && t < 36 && (pend & (1u << x)) &&
		    Handoff<R>::fresh(w[x], epoch)) {
			Xc[x][t] = Handoff<R>::value(w[x]);
			pend &= ~(1u << x);
		}
	}
	return pend;
}

/* faces with a bit of pend set in some lane (uniform) */
__device__ inline unsigned cl_any(unsigned pend)
{
	unsigned um = 0;
#pragma unroll
	for (int x = 0; x < CL_XMAX; x++)
		um |= __ballot((pend >> x) & 1u) ? 1u << x : 0u;
	return um;
}

/* every block records a 100 us busy interval (residency measurement) */
__device__ inline void residency_probe(u64 *disc)
{
	if (threadIdx.x == 0) {
		const u64 t0 = __builtin_amdgcn_s_memrealtime(); /* 100 MHz */
		u64 t1 = t0;
		for (int n = 0; n < (1 << 26) && t1 - t0 < 10000; n++)
			t1 = __builtin_amdgcn_s_memrealtime();
		disc[2 * blockIdx.x] = t0;
		disc[2 * blockIdx.x + 1] = t1;
	}
}

/*
 * Transport of one item of a cluster, in direct form once its inflow
 * traces are known (the item sweep instead forms D^-1 ML_s per slot before
 * its wait):
 *   sol = D^-1 (2 f - sum_f c_f ML_f u_f) - f
 * over every inflow face f in face order, u_f the ghost value, the lagged
 * value (1 - theta) f_nb + theta x_nb, the published trace of a producer
 * outside the cluster (Xc[x]) or of an earlier item of the cluster (Wc[j]).
 * Writes sol to fn, w = (1 - theta) f + theta sol (rounded to R) to wl and,
 * when a consumer lies in another cluster, to the hand-off words. Lanes
 * t = 10 iw + i (t < 60).
 */
template <typename R>
__device__ inline void cl_item(const DescLanes &D, const PrepLoads<R, 1> &P,
			       const R *__restrict__ fo, const R *__restrict__ xl,
			       R *__restrict__ fn, u64 *ll, int ne, int neq, int nv, unsigned epoch,
			       const SweepTables &T, PrepScratch &W, const double (*Wc)[60],
			       const double (*Xc)[36], double *wl)
{
	/* lanes 60..63 repeat lane 59's arithmetic (on its value, cl_issue
	 * loads it for them) instead of branching around it; only stores are
	 * limited to t < 60 */
	const int t = lane_id();
	const int tc = min(t, 59);
	const int iw = tc / 10, i = tc % 10;
	const R fv = P.fr[0];
	const int flags = D.flags();
	W.D[t / 10][t % 10] = P.d0;
	if (t < 36)
		W.D[(64 + t) / 10][(64 + t) % 10] = P.d1;
	double y = 0.0;
#pragma unroll
	for (int f = 0; f < 4; f++) {
		if (!(flags & (1 << f)))
			continue;
		const double cf = D.cf(f);
		double s = 0.0;
		if (flags & (0x110 << f)) {
			/* ghost or lagged: the trace staged through W.S */
			if (t < 36) {
				const int q = t / 6, b = t % 6, v = D.v(), nbf = D.nb(f);
				const long g = (flags & (16 << f))
						       ? 60L * nv * ne + 6L * v * (neq - 10L * ne) + 36L * nbf + t
						       : ((long)v * ne + nbf) * 60 + 10 * q +
								 ((D.map(f) >> (4 * b)) & 15);
				double u = (double)fo[g];
				if (!(flags & (16 << f))) {
					u = (1.0 - THETA) * u;
					if (xl)
						u += THETA * (double)xl[g];
				}
				W.S[t] = u;
			}
			__syncthreads();
#pragma unroll
			for (int b = 0; b < 6; b++)
				s += T.ML[f][i][b] * W.S[6 * iw + b];
			__syncthreads();
		} else {
			const int code = (flags >> (12 + 4 * f)) & 15;
			if (code >= 8) {
				const int m = D.map(f);
#pragma unroll
				for (int b = 0; b < 6; b++)
					s += T.ML[f][i][b] * Wc[code - 8][10 * iw + ((m >> (4 * b)) & 15)];
			} else {
#pragma unroll
				for (int b = 0; b < 6; b++)
					s += T.ML[f][i][b] * Xc[code][6 * iw + b];
			}
		}
		y += cf * s;
	}
	W.S[tc] = 2.0 * (double)fv - y; /* lanes 59..63 write the same value */
	__syncthreads();
	double s = 0.0;
#pragma unroll
	for (int l = 0; l < 10; l++)
		s += W.D[i][l] * W.S[10 * iw + l];
	const double sol = s - (double)fv;
	const R wr = (R)((1.0 - THETA) * (double)fv + THETA * sol);
	if (t < 60) {
		const int v = D.v(), e = D.e();
		__builtin_nontemporal_store((R)sol, fn + ((long)v * ne + e) * 60 + t);
		if (flags & (1 << 28))
			Handoff<R>::put(ll_item<R>(ll, v, e, ne), t, wr, epoch);
		/* in-cluster consumers take the value as the hand-off carries it */
		wl[t] = (double)published(wr, epoch);
	}
	__syncthreads();
}

/* the loads of the n items of a cluster whose descriptors are held one
 * dword per lane */
template <typename R>
__device__ inline void cl_issue(int da, int db, int n, const R *__restrict__ fo,
				const double *__restrict__ dinv, int ne, PrepLoads<R, 1> (&P)[CL_MAX])
{
	const int t = lane_id();
#pragma unroll
	for (int j = 0; j < CL_MAX; j++)
		if (j < n) {
			/* every lane loads (lanes 60..63 lane 59's value, lanes 36..63
			 * D^-1 entry 99 as their second entry), so that no load sits
			 * behind a branch; read once per sweep: non-temporal, which
			 * leaves the caches to the hand-off words */
			const DescLanes d{ da, db, j };
			const double *D = dinv + 100L * d.pad();
			P[j].fr[0] = __builtin_nontemporal_load(fo + ((long)d.v() * ne + d.e()) * 60 + min(t, 59));
			P[j].d0 = __builtin_nontemporal_load(D + t);
			P[j].d1 = __builtin_nontemporal_load(D + 64 + min(t, 35));
		}
}

/*
 * The sweep over clusters in one launch. Cluster k (header chead[k]: first
 * item, n | nx << 8, inflow faces from outside) holds n <= CL_MAX items of
 * one velocity in level order; block b takes clusters b, b + G, ... in a
 * topological order of the cluster graphs, so that the earliest unfinished
 * cluster always has its inputs from outside ready (all G blocks are
 * co-resident; bounded waits report a timeout in *status). The loads of a
 * cluster's items (PB) are issued while the block solves the cluster before
 * it (PA) and moved to PA when that one is done, when they have long
 * arrived. At the start of a cluster a first poll of its nx inflow faces
 * from outside goes out before the next cluster's loads, so that its wait
 * (loads complete in order) does not include them. The items are then
 * solved in order (cl_item), taking the traces of in-cluster producers from
 * LDS (Wc) and the others from Xc once complete. Descriptors are read two
 * clusters ahead, headers three.
 */
template <typename R>
__global__ __launch_bounds__(64, 4) void k_sweep_cluster(
	const int4 *__restrict__ chead, const ItemDesc *__restrict__ cdesc, int ncl,
	const R *__restrict__ fo, R *__restrict__ fn, const R *__restrict__ xl,
	const double *__restrict__ dinv, int nv, int ne, int neq, u64 *ll, unsigned epoch_arg,
	const unsigned *__restrict__ epoch_dev, int *status, u64 *disc, KinPar kp)
{
	__shared__ SweepTables T;
	__shared__ PrepScratch W;
	__shared__ double Wc[CL_MAX][60];
	__shared__ double Xc[CL_XMAX][36];
	if (disc) {
		residency_probe(disc);
		return;
	}
	load_tables(T);
	const unsigned epoch = epoch_dev ? *epoch_dev : epoch_arg;
	const int G = gridDim.x;
	int ci = blockIdx.x;
	if (ci >= ncl)
		return;
	const int t = threadIdx.x;
	PrepLoads<R, 1> PA[CL_MAX], PB[CL_MAX];
	int4 h = chead[ci];
	int da, db;
	cl_load_desc(cdesc, h.x, h.y & 255, da, db);
	cl_issue<R>(da, db, uni(h.y & 255), fo, dinv, ne, PA);
	int4 hn = ci + G < ncl ? chead[ci + G] : make_int4(0, 0, 0, 0);
	int nda = 0, ndb = 0;
	if (ci + G < ncl)
		cl_load_desc(cdesc, uni(hn.x), uni(hn.y & 255), nda, ndb);
	int4 hnn = ci + 2 * G < ncl ? chead[ci + 2 * G] : make_int4(0, 0, 0, 0);
	for (; ci < ncl; ci += G) {
		const int n = uni(h.y & 255), nx = uni((h.y >> 8) & 255);
		const int hz = uni(h.z), hw = uni(h.w);
		const int v = cl_dword(da, db, 9);
		const unsigned all = (1u << nx) - 1u;
		unsigned pend = cl_poll<R>(hz, hw, da, db, all, t < 36 ? all : 0u, ll, v, ne,
					   epoch, Xc);
		/* the next cluster's loads, its successor's descriptors, the header
		 * after that */
		if (ci + G < ncl)
			cl_issue<R>(nda, ndb, uni(hn.y & 255), fo, dinv, ne, PB);
		int nnda = 0, nndb = 0;
		if (ci + 2 * G < ncl)
			cl_load_desc(cdesc, uni(hnn.x), uni(hnn.y & 255), nnda, nndb);
		const int4 hnnn = ci + 3 * G < ncl ? chead[ci + 3 * G] : make_int4(0, 0, 0, 0);
#pragma unroll
		for (int j = 0; j < CL_MAX; j++) {
			if (j >= n)
				break;
			const DescLanes d{ da, db, j };
			/* inflow faces from outside that this item reads */
			unsigned need = 0;
			const int flags = d.flags();
			const int sf = flags & ~(flags >> 4) & ~(flags >> 8) & 15;
#pragma unroll
			for (int f = 0; f < 4; f++) {
				const int code = (flags >> (12 + 4 * f)) & 15;
				if ((sf & (1 << f)) && code < 8)
					need |= 1u << code;
			}
			int spins = 0;
			unsigned um = __ballot((pend & need) != 0u) ? cl_any(pend & need) : 0u;
			for (; um; um = cl_any(pend & need)) {
				__builtin_amdgcn_s_sleep(1);
				pend = cl_poll<R>(hz, hw, da, db, um, pend, ll, v, ne, epoch, Xc);
				if (spin_expired(++spins, status))
					break;
			}
			/* the item at the highest wave priority, the waits at the
			 * lowest: blocks with work go first (64^3 jittered sweep
			 * 4.27 -> 4.05 ms binary64, 4.09 -> 3.90 ms binary32) */
			__builtin_amdgcn_s_setprio(3);
			cl_item<R>(d, PA[j], fo, xl, fn, ll, ne, neq, nv, epoch, T, W, Wc, Xc,
				   &Wc[j][0]);
			__builtin_amdgcn_s_setprio(0);
		}
#pragma unroll
		for (int j = 0; j < CL_MAX; j++)
			PA[j] = PB[j];
		h = hn;
		hn = hnn;
		hnn = hnnn;
		da = nda;
		db = ndb;
		nda = nnda;
		ndb = nndb;
	}
}

/* largest number of the g0 intervals [h[2b], h[2b + 1]) that overlap */
static int max_overlap(const std::vector<u64> &h, int g0)
{
	std::vector<std::pair<u64, int>> ev;
	ev.reserve(2 * (size_t)g0);
	for (int b = 0; b < g0; b++) {
		ev.push_back({ h[2 * b], 1 });
		ev.push_back({ h[2 * b + 1], -1 });
	}
	std::sort(ev.begin(), ev.end(), [](const std::pair<u64, int> &x,
					   const std::pair<u64, int> &y) {
		return x.first < y.first || (x.first == y.first && x.second < y.second);
	});
	int cur = 0, best = 0;
	for (const auto &e : ev) {
		cur += e.second;
		best = std::max(best, cur);
	}
	return std::max(1, std::min(best, g0));
}

template <typename R> static int cluster_residency(gdh_ctx *c)
{
	int dev = 0, ncu = 0, per_cu = 0;
	HIPCHECK(hipGetDevice(&dev));
	HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
	const auto kern = k_sweep_cluster<R>;
	HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, 0));
	const int g0 = std::max(1, per_cu * ncu);
	u64 *disc = (u64 *)dmalloc(sizeof(u64) * 2 * (size_t)g0);
	kern<<<g0, 64, 0, c->stream>>>(nullptr, nullptr, 0, nullptr, nullptr, nullptr, nullptr, 0,
				       0, 0, nullptr, 0, nullptr, nullptr, disc, c->kp);
	HIPCHECK(hipGetLastError());
	std::vector<u64> h(2 * (size_t)g0);
	HIPCHECK(hipMemcpyAsync(h.data(), disc, sizeof(u64) * h.size(), hipMemcpyDeviceToHost,
				c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
	HIPCHECK(hipFree(disc));
	return max_overlap(h, g0);
}

/* cluster-order descriptors: the level-order descriptor of each item, its
 * level-order index in pad and the trace sources in the flags */
__global__ void k_cluster_desc(const ItemDesc *__restrict__ desc, const int *__restrict__ perm,
			       const int *__restrict__ code, int n, ItemDesc *__restrict__ cdesc)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n)
		return;
	ItemDesc d = desc[perm[k]];
	d.pad = perm[k];
	d.flags |= code[k] << 12;
	cdesc[k] = d;
}

/* the cluster plan on the device, and the blocks the cluster sweep holds
 * at once; no plan (ncl = 0) when the mesh gives none */
static void cluster_setup(gdh_ctx *c, const uint8_t *cls, const int *items_ve,
			  const int *level_ptr, int nlevels)
{
	c->ncl = 0;
	ClusterPlan P;
	if (!cluster_plan(c, cls, items_ve, level_ptr, nlevels, P) || P.head.empty())
		return;
	c->ncl = (int)P.head.size();
	c->cl_multi = P.multi;
	c->cl_levels = P.levels;
	c->chead = (int4 *)dmalloc(sizeof(int4) * P.head.size());
	HIPCHECK(hipMemcpy(c->chead, P.head.data(), sizeof(int4) * P.head.size(),
			   hipMemcpyHostToDevice));
	int *perm = (int *)dmalloc(sizeof(int) * P.perm.size());
	int *code = (int *)dmalloc(sizeof(int) * P.code.size());
	HIPCHECK(hipMemcpy(perm, P.perm.data(), sizeof(int) * P.perm.size(), hipMemcpyHostToDevice));
	HIPCHECK(hipMemcpy(code, P.code.data(), sizeof(int) * P.code.size(), hipMemcpyHostToDevice));
	c->cdesc = (ItemDesc *)dmalloc(sizeof(ItemDesc) * (size_t)c->nitems);
	k_cluster_desc<<<nblocks(c->nitems, 256), 256, 0, c->stream>>>(c->desc, perm, code,
								       c->nitems, c->cdesc);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipStreamSynchronize(c->stream));
	HIPCHECK(hipFree(perm));
	HIPCHECK(hipFree(code));
	c->cl_grid = grid_cap(c->fp32 ? cluster_residency<float>(c) : cluster_residency<double>(c));
}

/*
 * Blocks of k_sweep_persistent that the device holds at once, measured: the
 * occupancy query over-counts on gfx950 when LDS bounds the residency (it
 * reports 20 blocks per CU for 8000 B of LDS where 18 fit), and a
 * cooperative launch sized by it does not have all its blocks resident.
 */
template <typename R>
using SweepKernel = void (*)(const ItemDesc *, int, const R *, R *, const R *, const double *,
			     int, int, int, u64 *, unsigned, const unsigned *, int *, u64 *, int,
			     KinPar, unsigned *, int);

/* the persistent sweep of this context: precision, way of polling and
 * number of velocities ... */
template <typename R, bool LB, bool QU>
static SweepKernel<R> sweep_kernel_q(const gdh_ctx *c)
{
	/* the relaxation by the element's last item: whole-trace polling, the
	 * stored inverses and the static order only (relax_choose) */
	if (c->relax_fused == 2 && !LB && !QU && c->sweep_poll)
		return c->kp.nv == 4 ? k_sweep_persistent<R, true, 4, false, false, true>
				     : k_sweep_persistent<R, true, 6, false, false, true>;
	if (c->kp.nv == 4)
		return c->sweep_poll ? k_sweep_persistent<R, true, 4, LB, QU>
				     : k_sweep_persistent<R, false, 4, LB, QU>;
	if (c->kp.nv == 6)
		return c->sweep_poll ? k_sweep_persistent<R, true, 6, LB, QU>
				     : k_sweep_persistent<R, false, 6, LB, QU>;
	fprintf(stderr, "gdon3d-hip: no transport sweep for %d velocities\n", c->kp.nv);
	abort();
}

/* ... and the way its blocks take their items (static order or queue) */
template <typename R, bool LB> static SweepKernel<R> sweep_kernel_lb(const gdh_ctx *c)
{
	return c->sweep_queue ? sweep_kernel_q<R, LB, true>(c) : sweep_kernel_q<R, LB, false>(c);
}

/* the kernel with the learned refinement only where learned inverses are
 * enabled: the refinement's registers stay out of the exact sweep */
template <typename R> static SweepKernel<R> sweep_kernel(const gdh_ctx *c)
{
	return c->kp.lb_p ? sweep_kernel_lb<R, true>(c) : sweep_kernel_lb<R, false>(c);
}

template <typename R> static int sweep_residency(gdh_ctx *c)
{
	int dev = 0, ncu = 0, per_cu = 0;
	HIPCHECK(hipGetDevice(&dev));
	HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
	const SweepKernel<R> kern = sweep_kernel<R>(c);
	HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, 0));
	const int g0 = std::max(1, per_cu * ncu);
	u64 *disc = (u64 *)dmalloc(sizeof(u64) * 2 * (size_t)g0);
	kern<<<g0, 64, 0, c->stream>>>(nullptr, 0, nullptr, nullptr, nullptr, nullptr, 0, 0, 0,
				       nullptr, 0, nullptr, nullptr, disc, 0, c->kp, nullptr, 0);
	HIPCHECK(hipGetLastError());
	std::vector<u64> h(2 * (size_t)g0);
	HIPCHECK(hipMemcpyAsync(h.data(), disc, sizeof(u64) * h.size(),
				hipMemcpyDeviceToHost, c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
	HIPCHECK(hipFree(disc));
	return max_overlap(h, g0);
}

void sweep_setup(gdh_ctx *c)
{
	c->desc = (ItemDesc *)dmalloc(sizeof(ItemDesc) * (size_t)c->nitems);
	k_item_desc<<<nblocks(c->nitems, 64), 64, 0, c->stream>>>(
		c->items, c->nitems, c->geom, c->nbr, c->nbmap, c->cls, c->kp, c->dt,
		c->desc, c->bt);
	HIPCHECK(hipGetLastError());
	/* one word of the state precision per value */
	const size_t llb = c->rsz * 60 * (size_t)c->nitems;
	c->ll = (u64 *)dmalloc(llb);
	HIPCHECK(hipMemsetAsync(c->ll, 0, llb, c->stream));
	c->ll_partial = 0;
	c->status = (int *)dmalloc(sizeof(int));
	HIPCHECK(hipMemsetAsync(c->status, 0, sizeof(int), c->stream));
	c->queue = (unsigned *)dmalloc(QUEUE_BYTES);
	HIPCHECK(hipMemsetAsync(c->queue, 0, QUEUE_BYTES, c->stream));
	c->qslot = 0;
	/* GDN_SWEEP_QUEUE=static|dynamic: the persistent sweep's items in the
	 * order b, b + G, ... (needs every block resident: cooperative launch)
	 * or taken from a counter (any grid) */
	const char *qenv = getenv("GDN_SWEEP_QUEUE");
	c->sweep_queue = qenv ? !strcmp(qenv, "dynamic") : 0;
	HIPCHECK(hipStreamSynchronize(c->stream));
	c->epoch = 0;
	/* GDN_SWEEP=level|item (or persistent)|cluster fixes the form of the sweep; by default
	 * the item and cluster sweeps are timed at set-up (relax_choose) */
	const char *mode = getenv("GDN_SWEEP");
	c->sweep_mode = !mode ? -1 : !strcmp(mode, "level") ? 0
			: (!strcmp(mode, "item") || !strcmp(mode, "persistent")) ? 1
			: !strcmp(mode, "cluster") ? 2 : -1;
	/* waiting items poll the whole traces on every pass ("full", the
	 * default), or one word per inflow face and then the traces
	 * ("sentinel"); measured on MI355X, D3Q4 binary64 item sweep with one
	 * hand-off word per value, full vs sentinel: 32^3 fused 0.876 vs
	 * 0.882 ms, separate 0.768 vs 0.796 ms, 64^3 5.67 vs 5.72 ms (with
	 * two words per binary64 value the sentinel had won, 0.910 vs 0.925);
	 * binary32 0.738 vs 0.760 ms. GDN_SWEEP_POLL=sentinel|full overrides. */
	const char *poll = getenv("GDN_SWEEP_POLL");
	c->sweep_poll = poll ? strcmp(poll, "sentinel") != 0 : 1;
	c->sweep_grid = c->fp32 ? sweep_residency<float>(c) : sweep_residency<double>(c);
	c->sweep_grid = grid_cap(c->sweep_grid);
}

/* GDN_SWEEP_GRID=n: at most n workgroups in the one-launch sweeps (fewer
 * than the resident count keeps every wait satisfiable) */
static int grid_cap(int g)
{
	const char *env = getenv("GDN_SWEEP_GRID");
	const int cap = env ? atoi(env) : 0;
	return cap > 0 ? std::min(g, cap) : g;
}

/* zero the hand-off words and restart the epochs at 1 (set-up only: not
 * in a captured stream) */
static void ll_restart(gdh_ctx *c)
{
	hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
	HIPCHECK(hipStreamIsCapturing(c->stream, &cs));
	if (cs != hipStreamCaptureStatusNone) {
		fprintf(stderr, "gdon3d-hip: the sweep form changed inside a captured graph\n");
		abort();
	}
	HIPCHECK(hipMemsetAsync(c->ll, 0, c->rsz * 60 * (size_t)c->nitems, c->stream));
	HIPCHECK(hipMemsetAsync(c->queue, 0, QUEUE_BYTES, c->stream));
	c->qslot = 0;
	if (c->epoch_dev)
		HIPCHECK(hipMemsetAsync(c->epoch_dev, 0, sizeof(unsigned), c->stream));
	c->epoch = 0;
	c->ll_partial = 0;
}

__global__ void k_epoch_bump(unsigned *e)
{
	const unsigned n = *e + 1u;
	*e = n == 0u ? 2u : n; /* wrapped: keep the parity alternating */
}

/*
 * A cooperative launch the runtime refuses (seen when another process holds
 * the device): the sweeps go on with ordinary launches, the item sweep
 * taking its items from the queue, which needs only QUEUE_SHARDS resident
 * workgroups. Any other error stops as every HIP error does.
 */
static bool coop_refused(gdh_ctx *c, hipError_t e, const char *where)
{
	if (e == hipSuccess)
		return false;
	if (e != hipErrorIllegalState && e != hipErrorNotSupported &&
	    e != hipErrorCooperativeLaunchTooLarge) {
		fprintf(stderr, "gdon3d-hip: %s failed: %s\n", where, hipGetErrorString(e));
		abort();
	}
	(void)hipGetLastError();
	if (!c->coop_refused)
		fprintf(stderr, "gdon3d-hip: the runtime refused a cooperative launch (%s); the "
				"transport sweep continues as the item sweep over its queue\n",
			hipGetErrorString(e));
	c->coop_refused = 1;
	return true;
}

template <typename R>
static void sweep_step_t(gdh_ctx *c, R *fo, R *fn, const R *xl, int relax_in)
{
	/* once a cooperative launch was refused: the item sweep over its queue */
	if (c->coop_refused && c->sweep_mode != 0) {
		c->sweep_mode = 1;
		c->sweep_queue = 1;
	}
	const unsigned epoch0 = c->epoch;
	/* a sweep that reads every item after one in which some items did not
	 * publish (cluster sweep): zero the words and restart the epochs, so
	 * that no word left from two sweeps back passes for this one */
	if (c->ll_partial && c->sweep_mode != 2)
		ll_restart(c);
	/* the sweep's epoch: on the host, or (batches, captured in graphs) in
	 * device memory advanced by a one-thread node; only its parity reaches
	 * the words */
	if (c->epoch_dev)
		k_epoch_bump<<<1, 1, 0, c->stream>>>(c->epoch_dev);
	else if (++c->epoch == 0)
		c->epoch = 2; /* wrapped: 2^32 is even, keep the alternation */
	if (c->sweep_mode == 2)
		c->ll_partial = 1;
	const ItemDesc *desc = c->desc;
	int nitems = c->nitems, nv = c->kp.nv, ne = c->ne, neq = c->neq;
	const double *dinv = c->dinv;
	u64 *ll = c->ll, *disc = nullptr;
	int *status = c->status;
	unsigned epoch = c->epoch;
	const unsigned *edev = c->epoch_dev;
	KinPar kp = c->kp;
	if (c->sweep_mode == 0) {
		for (int L = 0; L < c->nlevels; L++) {
			const int b = c->level_ptr[L], n = c->level_ptr[L + 1] - b;
			if (!n)
				continue;
			c->prof_launches++;
			const bool lb = kp.lb_p != nullptr;
			auto lk = nv == 4 ? (lb ? k_sweep_level<R, 4, true> : k_sweep_level<R, 4, false>)
					  : (lb ? k_sweep_level<R, 6, true> : k_sweep_level<R, 6, false>);
			lk<<<n, 64, 0, c->stream>>>(desc, b, fo, fn, xl, dinv, nv, ne, neq, ll, epoch,
						    edev, status, relax_in, kp);
		}
		HIPCHECK(hipGetLastError());
		return;
	}
	c->prof_launches++;
	const R *fo_c = fo;
	if (c->sweep_mode == 2) {
		if (relax_in) {
			fprintf(stderr, "gdon3d-hip: the cluster sweep takes a relaxed state\n");
			abort();
		}
		const int4 *chead = c->chead;
		const ItemDesc *cdesc = c->cdesc;
		int ncl = c->ncl;
		const int cgrid = std::min(c->cl_grid, ncl);
		const auto ck = k_sweep_cluster<R>;
		if (edev) {
			ck<<<cgrid, 64, 0, c->stream>>>(chead, cdesc, ncl, fo_c, fn, xl, dinv, nv, ne,
							neq, ll, 0u, edev, status, disc, kp);
			HIPCHECK(hipGetLastError());
			return;
		}
		void *cargs[] = { &chead, &cdesc, &ncl, &fo_c, &fn, &xl, &dinv, &nv, &ne, &neq,
				  &ll, &epoch, &edev, &status, &disc, &kp };
		if (coop_refused(c, hipLaunchCooperativeKernel((const void *)ck, dim3(cgrid), dim3(64),
							       cargs, 0, c->stream),
				 "the cluster sweep's cooperative launch")) {
			c->epoch = epoch0;
			sweep_step_t<R>(c, fo, fn, xl, relax_in);
		}
		return;
	}
	const int grid = std::min(c->relax_fused == 2 ? c->sweep_grid_lf : c->sweep_grid, nitems);
	const SweepKernel<R> kern = sweep_kernel<R>(c);
	/* items from the queue (slots alternate between launches) or in the
	 * static order */
	unsigned *queue = c->sweep_queue ? c->queue : nullptr;
	int slot = c->qslot;
	if (queue)
		c->qslot = 1 - c->qslot;
	if (edev || queue) {
		/* an ordinary launch with the measured resident grid: a captured
		 * stream takes no cooperative launch, and the queue needs no
		 * co-residency */
		kern<<<grid, 64, 0, c->stream>>>(desc, nitems, fo_c, fn, xl, dinv, nv, ne, neq, ll,
						 epoch, edev, status, disc, relax_in, kp, queue, slot);
		HIPCHECK(hipGetLastError());
		return;
	}
	void *args[] = { &desc, &nitems, &fo_c, &fn, &xl, &dinv, &nv, &ne, &neq, &ll,
			 &epoch, &edev, &status, &disc, &relax_in, &kp, &queue, &slot };
	if (coop_refused(c, hipLaunchCooperativeKernel((const void *)kern, dim3(grid), dim3(64),
						       args, 0, c->stream),
			 "the item sweep's cooperative launch")) {
		c->epoch = epoch0;
		sweep_step_t<R>(c, fo, fn, xl, relax_in);
	}
}

void sweep_step(gdh_ctx *c, void *fo, void *fn, const void *xl, int relax_in)
{
	if (c->fp32)
		sweep_step_t<float>(c, (float *)fo, (float *)fn, (const float *)xl, relax_in);
	else
		sweep_step_t<double>(c, (double *)fo, (double *)fn, (const double *)xl, relax_in);
}

/* f <- omega f_eq(w(f)) + (1 - omega) f at DoFs [ipg0, n)
 * (src/dg/dg.c:109-132, src/simulation/simulation.c:42-54), one DoF per
 * thread (the ghost DoFs); NV = kp.nv. */
template <int NV, typename R>
__global__ __launch_bounds__(256) void k_relax_dof(R *__restrict__ f, KinPar kp,
						   long ipg0, long n, long ne, long neq)
{
	const long ipg = ipg0 + (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (ipg >= n)
		return;
	double fo[6][NV];
	double w[6];
#pragma unroll
	for (int iw = 0; iw < 6; iw++) {
		double s = 0.0;
#pragma unroll
		for (int v = 0; v < NV; v++) {
			const double x = (double)f[kin_at(ne, neq, NV, ipg, iw, v)];
			fo[iw][v] = x;
			s += x;
		}
		w[iw] = s;
	}
	/* the flux of the equilibrium on (D / eps, B) */
	const double ie = kin_ieps(kp, ipg, ne);
	const double wt[6] = { w[0] * ie, w[1] * ie, w[2] * ie, w[3], w[4], w[5] };
	double q[6][3];
	maxwell_qM(wt, q);
	const double om = kp.omega;
#pragma unroll
	for (int iw = 0; iw < 6; iw++)
#pragma unroll
		for (int v = 0; v < NV; v++)
			f[kin_at(ne, neq, NV, ipg, iw, v)] =
				(R)(om * feq_comp(kp, w, q, iw, v) + (1. - om) * fo[iw][v]);
}

/*
 * The same relaxation over the element DoFs, one 64-lane wave per element:
 * lane t = 10 iw + i (t < 60) holds value (iw, i) of every velocity, so each
 * velocity's 60 values are one contiguous load of the item-major state.
 * The node's six w components come from the lanes i, 10 + i, ... by
 * shuffles. BS / 64 elements per block.
 */
template <int NV, typename R, int BS>
__global__ __launch_bounds__(BS) void k_relax(R *__restrict__ f, KinPar kp, long ne)
{
	const long e = (long)blockIdx.x * (BS / 64) + threadIdx.x / 64;
	if (e >= ne)
		return; /* whole waves */
	const int t = threadIdx.x % 64;
	const int tt = t < 60 ? t : 59, iw = tt / 10, i = tt % 10;
	R *p = f + e * 60 + tt;
	const long vs = 60 * ne;
	double fo[NV];
	double s = 0.0;
#pragma unroll
	for (int v = 0; v < NV; v++) {
		fo[v] = (double)p[v * vs];
		s += fo[v];
	}
	double w[6];
#pragma unroll
	for (int q = 0; q < 6; q++)
		w[q] = __shfl(s, 10 * q + i, 64);
	/* the flux of the equilibrium on (D / eps, B) */
	const double ie = kp.ieps ? kp.ieps[e] : 1.0;
#pragma unroll
	for (int q = 0; q < 3; q++)
		w[q] *= ie;
	double r[3];
	maxwell_q_row(w, iw, r);
	const double om = kp.omega;
	double out[NV];
	if (kp.kind == GDH_KIN_D3Q4) {
		/* feq_comp, D3Q4 branch */
		const double l2 = kp.lambda * kp.lambda;
#pragma unroll
		for (int v = 0; v < NV; v++) {
			const double *vv = kp.vi + 3 * v;
			const double qv = r[0] * vv[0] + r[1] * vv[1] + r[2] * vv[2];
			out[v] = om * ((s + qv / l2) / kp.nv) + (1. - om) * fo[v];
		}
	} else {
		/* feq_comp, D3Q6 branch: v = +-lambda e_k, index k or 3 + k */
#pragma unroll
		for (int v = 0; v < NV; v++) {
			const double a = s / 6., b = r[v % 3] / (2. * kp.lambda);
			out[v] = om * (v < 3 ? a + b : a - b) + (1. - om) * fo[v];
		}
	}
	if (t < 60)
#pragma unroll
		for (int v = 0; v < NV; v++)
			p[v * vs] = (R)out[v];
}

/* two values of the state precision moved as one access */
template <typename R> struct Pair2;
template <> struct Pair2<double> {
	using T = double2;
};
template <> struct Pair2<float> {
	using T = float2;
};

/*
 * The same relaxation with two elements per wave and two values per lane:
 * each velocity's 120 values of elements e0, e0 + 1 (e0 even) are one
 * contiguous run, read and written by lanes 0..59 with one 16-byte (binary64)
 * or 8-byte (binary32) access each; lane l holds values 2 l and 2 l + 1 of
 * the pair, nodes i0 and i0 + 1 (i0 even) of field iw of element el. The
 * node sums w come by shuffles from the lane holding the same two nodes of
 * field q; the arithmetic per value is that of k_relax, bit for bit. Half
 * as many waves as k_relax and twice the bytes per access.
 */
template <int NV, typename R, int BS>
__global__ __launch_bounds__(BS) void k_relax_pair(R *__restrict__ f, KinPar kp, long ne)
{
	using R2 = typename Pair2<R>::T;
	const long e0 = 2 * ((long)blockIdx.x * (BS / 64) + threadIdx.x / 64);
	if (e0 >= ne)
		return; /* whole waves */
	const int t = threadIdx.x % 64;
	const int tt = t < 60 ? t : 59;
	const int el = tt / 30, loc = 2 * tt - 60 * el, iw = loc / 10, i0 = loc % 10;
	/* an odd number of elements leaves the last wave one of them */
	const bool live = t < 60 && (el == 0 || e0 + 1 < ne);
	R2 *p = reinterpret_cast<R2 *>(f + e0 * 60) + tt;
	const long vs = 30 * ne; /* one velocity's values, in pairs */
	double fa[NV], fb[NV];
	double sa = 0.0, sb = 0.0;
#pragma unroll
	for (int v = 0; v < NV; v++) {
		R2 x = live ? p[v * vs] : R2{ 0, 0 };
		fa[v] = (double)x.x;
		fb[v] = (double)x.y;
		sa += fa[v];
		sb += fb[v];
	}
	/* lane holding nodes i0, i0 + 1 of field q of the same element */
	double wa[6], wb[6];
#pragma unroll
	for (int q = 0; q < 6; q++) {
		const int src = 30 * el + 5 * q + i0 / 2;
		wa[q] = __shfl(sa, src, 64);
		wb[q] = __shfl(sb, src, 64);
	}
	const double ie = kp.ieps && live ? kp.ieps[e0 + el] : 1.0;
#pragma unroll
	for (int q = 0; q < 3; q++) {
		wa[q] *= ie;
		wb[q] *= ie;
	}
	double ra[3], rb[3];
	maxwell_q_row(wa, iw, ra);
	maxwell_q_row(wb, iw, rb);
	const double om = kp.omega;
	R2 out[NV];
	if (kp.kind == GDH_KIN_D3Q4) {
		const double l2 = kp.lambda * kp.lambda;
#pragma unroll
		for (int v = 0; v < NV; v++) {
			const double *vv = kp.vi + 3 * v;
			const double qa = ra[0] * vv[0] + ra[1] * vv[1] + ra[2] * vv[2];
			const double qb = rb[0] * vv[0] + rb[1] * vv[1] + rb[2] * vv[2];
			out[v].x = (R)(om * ((sa + qa / l2) / kp.nv) + (1. - om) * fa[v]);
			out[v].y = (R)(om * ((sb + qb / l2) / kp.nv) + (1. - om) * fb[v]);
		}
	} else {
#pragma unroll
		for (int v = 0; v < NV; v++) {
			const double aa = sa / 6., ba = ra[v % 3] / (2. * kp.lambda);
			const double ab = sb / 6., bb = rb[v % 3] / (2. * kp.lambda);
			out[v].x = (R)(om * (v < 3 ? aa + ba : aa - ba) + (1. - om) * fa[v]);
			out[v].y = (R)(om * (v < 3 ? ab + bb : ab - bb) + (1. - om) * fb[v]);
		}
	}
	if (live)
#pragma unroll
		for (int v = 0; v < NV; v++)
			p[v * vs] = out[v];
}

template <int NV, typename R, int BS> static void relax_launch_nv(gdh_ctx *c, R *f, long n)
{
	if (c->relax_pair)
		k_relax_pair<NV, R, BS><<<nblocks((c->ne + 1) / 2, BS / 64), BS, 0, c->stream>>>(
			f, c->kp, c->ne);
	else
		k_relax<NV, R, BS><<<nblocks(c->ne, BS / 64), BS, 0, c->stream>>>(f, c->kp, c->ne);
	const long g0 = 10L * c->ne;
	if (n > g0)
		k_relax_dof<NV, R><<<nblocks(n - g0, 256), 256, 0, c->stream>>>(f, c->kp, g0, n,
										  c->ne, c->neq);
}

template <typename R, int BS> static void relax_launch_bs(gdh_ctx *c, R *f, long n)
{
	if (n < 10L * c->ne) {
		fprintf(stderr, "gdon3d-hip: relaxation of part of the elements\n");
		abort();
	}
	if (c->kp.nv == 4)
		relax_launch_nv<4, R, BS>(c, f, n);
	else if (c->kp.nv == 6)
		relax_launch_nv<6, R, BS>(c, f, n);
	else {
		fprintf(stderr, "gdon3d-hip: no relaxation kernel for %d velocities\n",
			c->kp.nv);
		abort();
	}
	HIPCHECK(hipGetLastError());
}

template <typename R> static void relax_launch_t(gdh_ctx *c, R *f, long n)
{
	switch (c->relax_bs) {
	case 128: relax_launch_bs<R, 128>(c, f, n); break;
	case 512: relax_launch_bs<R, 512>(c, f, n); break;
	default: relax_launch_bs<R, 256>(c, f, n); break;
	}
}

void relax_launch(gdh_ctx *c, void *f, long n)
{
	if (c->fp32)
		relax_launch_t<float>(c, (float *)f, n);
	else
		relax_launch_t<double>(c, (double *)f, n);
}

/*
 * Exact transport over upwind cycles (capacitance form). The sweep with
 * lagged inflow values z is affine in (f, z): x = sweep(f, z) = x0 + S z with
 * x0 = sweep(f, 0). The lagged values are the new solution at the lagged
 * producers' face nodes, z = T x, so (I - T S) z = T x0: a small dense system
 * per velocity (q = distinct lagged nodes), the same for the 6 fields. S is
 * built once at set-up from sweeps with unit lagged inputs (6 columns per
 * sweep, one per field slot, written on the device), and (I - T S) is
 * inverted there in binary64 by Gauss-Jordan elimination with partial
 * pivoting, in place: the result is X = (P (I - T S))^-1 for the row
 * permutation P of the pivots, so z = X (P t). Each step: x0 = sweep(f, 0);
 * t = T x0; z = X (P t), a dense product for the 6 fields at once;
 * x = sweep(f, z) — two sweeps, exact up to rounding.
 */
template <typename R>
__global__ void k_cap_gather(const R *__restrict__ x, const int *__restrict__ ipg,
			     const int *__restrict__ vel, int q, int nv, long ne, long neq,
			     double *__restrict__ t)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= 6 * q)
		return;
	const int iw = k / q, r = k % q;
	t[k] = (double)x[kin_at(ne, neq, nv, ipg[r], iw, vel[r])];
}

#define CAP_JC 256 /* columns of X per workgroup of the product */

/* partial products of z_v = X_v (P t)_v over the columns [c CAP_JC,
 * (c + 1) CAP_JC) of X_v, the 6 fields at once; one row per lane, X read
 * column by column (coalesced); grid (rows / 256, column chunks, nv) */
__global__ __launch_bounds__(256) void k_cap_product(
	const double *__restrict__ X, const long *__restrict__ luoff, const int *__restrict__ off,
	const int *__restrict__ piv, const double *__restrict__ t, int qtot, int qmax,
	double *__restrict__ part)
{
	__shared__ double y[6][CAP_JC];
	const int v = blockIdx.z, c = blockIdx.y;
	const int o = off[v], q = off[v + 1] - o;
	const int j0 = c * CAP_JC;
	if (j0 >= q)
		return;
	const int nj = min(q - j0, CAP_JC);
	for (int j = threadIdx.x; j < nj; j += 256) {
		const int pj = piv[o + j0 + j];
#pragma unroll
		for (int iw = 0; iw < 6; iw++)
			y[iw][j] = t[(long)iw * qtot + o + pj];
	}
	__syncthreads();
	const int i = blockIdx.x * 256 + threadIdx.x;
	if (i >= q)
		return;
	const double *A = X + luoff[v] + (long)j0 * q + i;
	double acc[6] = { 0, 0, 0, 0, 0, 0 };
	for (int j = 0; j < nj; j++) {
		const double a = A[(long)j * q];
#pragma unroll
		for (int iw = 0; iw < 6; iw++)
			acc[iw] += a * y[iw][j];
	}
	const int nch = (qmax + CAP_JC - 1) / CAP_JC;
#pragma unroll
	for (int iw = 0; iw < 6; iw++)
		part[(((long)v * nch + c) * 6 + iw) * qmax + i] = acc[iw];
}

/* z: the sum of the partial products (fixed order), scattered into the
 * lagged-value state buffer; grid (6 qmax / 256, nv) */
template <typename R>
__global__ void k_cap_reduce(const double *__restrict__ part, const int *__restrict__ off,
			     const int *__restrict__ ipg, int qmax, int nv, long ne, long neq,
			     R *__restrict__ wz)
{
	const int v = blockIdx.y;
	const int o = off[v], q = off[v + 1] - o;
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= 6 * q)
		return;
	const int iw = k / q, i = k % q;
	const int nc = (q + CAP_JC - 1) / CAP_JC, nch = (qmax + CAP_JC - 1) / CAP_JC;
	double s = 0;
	for (int c = 0; c < nc; c++)
		s += part[(((long)v * nch + c) * 6 + iw) * qmax + i];
	wz[kin_at(ne, neq, nv, ipg[o + i], iw, v)] = (R)s;
}

/* unit lagged inputs of set-up batch c: field slot iw carries column
 * 6 c + iw of every velocity */
template <typename R>
__global__ void k_cap_unit(R *__restrict__ wz, const int *__restrict__ ipg,
			   const int *__restrict__ off, int batch, int nv, long ne, long neq)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= 6 * nv)
		return;
	const int v = k / 6, iw = k % 6, r = 6 * batch + iw;
	if (r < off[v + 1] - off[v])
		wz[kin_at(ne, neq, nv, ipg[off[v] + r], iw, v)] = (R)1;
}

/* columns 6 batch + iw of (I - T S)_v from the gathered T S e (column-major);
 * grid (6 qmax / 256, nv) */
__global__ void k_cap_column(const double *__restrict__ t, const int *__restrict__ off,
			     const long *__restrict__ luoff, int batch, int qtot,
			     double *__restrict__ X)
{
	const int v = blockIdx.y;
	const int o = off[v], q = off[v + 1] - o;
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= 6 * q)
		return;
	const int iw = k / q, i = k % q, col = 6 * batch + iw;
	if (col < q)
		X[luoff[v] + (long)col * q + i] =
			(i == col ? 1.0 : 0.0) - t[(long)iw * qtot + o + i];
}

/* Step k of the in-place Gauss-Jordan inversion of every X_v (one
 * workgroup per velocity): the row p >= k of the largest |X_v(p, k)| is
 * swapped with row k (piv records it), then row k and column k are saved
 * (rk: row k, then column k, 2 qmax per velocity) for the update; a zero
 * or non-finite pivot sets *sing. */
__global__ __launch_bounds__(1024) void k_gj_pivot(double *__restrict__ X,
						   const long *__restrict__ luoff,
						   const int *__restrict__ off, int *__restrict__ piv,
						   double *__restrict__ rk, int qmax, int k,
						   int *__restrict__ sing)
{
	__shared__ double bv[1024];
	__shared__ int bi[1024];
	const int v = blockIdx.x, tid = threadIdx.x;
	const int o = off[v], q = off[v + 1] - o;
	if (k >= q)
		return;
	double *A = X + luoff[v];
	double best = -1.0;
	int bidx = k;
	for (int i = k + tid; i < q; i += 1024) {
		const double a = fabs(A[(long)k * q + i]);
		if (a > best) {
			best = a;
			bidx = i;
		}
	}
	bv[tid] = best;
	bi[tid] = bidx;
	__syncthreads();
	for (int s = 512; s > 0; s >>= 1) {
		if (tid < s && (bv[tid + s] > bv[tid] ||
				(bv[tid + s] == bv[tid] && bi[tid + s] < bi[tid]))) {
			bv[tid] = bv[tid + s];
			bi[tid] = bi[tid + s];
		}
		__syncthreads();
	}
	const int p = bi[0];
	if (p != k) {
		for (int j = tid; j < q; j += 1024) {
			const double a = A[(long)j * q + k];
			A[(long)j * q + k] = A[(long)j * q + p];
			A[(long)j * q + p] = a;
		}
		if (tid == 0) {
			const int s = piv[o + k];
			piv[o + k] = piv[o + p];
			piv[o + p] = s;
		}
	}
	__syncthreads();
	double *r = rk + (long)v * 2 * qmax;
	for (int j = tid; j < q; j += 1024) {
		r[j] = A[(long)j * q + k];
		r[qmax + j] = A[(long)k * q + j];
	}
	if (tid == 0) {
		const double pv = A[(long)k * q + k];
		if (pv == 0.0 || !isfinite(pv))
			*sing = 1;
	}
}

#define GJ_JC 16 /* columns per lane in the Gauss-Jordan update */

/* the rest of step k: every entry of X_v from its old value and the saved
 * row k (r) and column k (cc): X(k, k) = 1 / p, X(k, j) = r_j / p,
 * X(i, k) = -cc_i / p, X(i, j) -= cc_i r_j / p; grid (rows / 256,
 * columns / GJ_JC, nv) */
__global__ __launch_bounds__(256) void k_gj_update(double *__restrict__ X,
						   const long *__restrict__ luoff,
						   const int *__restrict__ off,
						   const double *__restrict__ rk, int qmax, int k)
{
	const int v = blockIdx.z;
	const int o = off[v], q = off[v + 1] - o;
	const int i = blockIdx.x * 256 + threadIdx.x;
	const int j0 = blockIdx.y * GJ_JC;
	if (k >= q || i >= q || j0 >= q)
		return;
	const int j1 = min(q, j0 + GJ_JC);
	const double *r = rk + (long)v * 2 * qmax, *cc = r + qmax;
	const double inv = 1.0 / cc[k], ci = cc[i];
	double *A = X + luoff[v];
	for (int j = j0; j < j1; j++) {
		const double rj = r[j] * inv;
		double a;
		if (i == k)
			a = j == k ? inv : rj;
		else if (j == k)
			a = -ci * inv;
		else
			a = A[(long)j * q + i] - ci * rj;
		A[(long)j * q + i] = a;
	}
}

static void cap_gather(gdh_ctx *c, const void *x)
{
	const int q = c->cap_qtot;
	if (c->fp32)
		k_cap_gather<float><<<nblocks(6 * q, 256), 256, 0, c->stream>>>(
			(const float *)x, c->cap_ipg, c->cap_vel, q, c->kp.nv, c->ne, c->neq, c->cap_t);
	else
		k_cap_gather<double><<<nblocks(6 * q, 256), 256, 0, c->stream>>>(
			(const double *)x, c->cap_ipg, c->cap_vel, q, c->kp.nv, c->ne, c->neq, c->cap_t);
	HIPCHECK(hipGetLastError());
}

/* Set-up of the capacitance form: positions, (I - T S) by unit sweeps, its
 * inverse (or, above the node limit, the matrix-free form's start). */
static void cap_setup(gdh_ctx *c)
{
	const int nv = c->kp.nv;
	/* distinct lagged producer nodes per velocity, from the item descriptors */
	std::vector<ItemDesc> desc(c->nitems);
	HIPCHECK(hipMemcpy(desc.data(), c->desc, sizeof(ItemDesc) * desc.size(),
			   hipMemcpyDeviceToHost));
	std::vector<std::vector<int>> pos(nv);
	for (const ItemDesc &d : desc)
		for (int f = 0; f < 4; f++)
			if (d.flags & (256 << f))
				for (int b = 0; b < 6; b++)
					pos[d.v].push_back(10 * d.nb[f] + ((d.map[f] >> (4 * b)) & 15));
	std::vector<int> off(nv + 1, 0), ipg, vel;
	for (int v = 0; v < nv; v++) {
		std::sort(pos[v].begin(), pos[v].end());
		pos[v].erase(std::unique(pos[v].begin(), pos[v].end()), pos[v].end());
		off[v + 1] = off[v] + (int)pos[v].size();
		for (int x : pos[v]) {
			ipg.push_back(x);
			vel.push_back(v);
		}
	}
	const int qtot = off[nv];
	int qmax = 0;
	for (int v = 0; v < nv; v++)
		qmax = std::max(qmax, off[v + 1] - off[v]);
	int qcap = CAP_MAX_NODES;
	if (const char *env = getenv("GDN_CAP_MAX_NODES"))
		qcap = std::min(std::max(atoi(env), 0), CAP_MAX_DENSE);
	c->cap_gmres = qmax > qcap;
	c->cap_qtot = qtot;
	c->cap_off_h = off;
	c->cap_ipg = (int *)dmalloc(sizeof(int) * qtot);
	c->cap_vel = (int *)dmalloc(sizeof(int) * qtot);
	c->cap_off = (int *)dmalloc(sizeof(int) * (nv + 1));
	HIPCHECK(hipMemcpy(c->cap_ipg, ipg.data(), sizeof(int) * qtot, hipMemcpyHostToDevice));
	HIPCHECK(hipMemcpy(c->cap_vel, vel.data(), sizeof(int) * qtot, hipMemcpyHostToDevice));
	HIPCHECK(hipMemcpy(c->cap_off, off.data(), sizeof(int) * (nv + 1), hipMemcpyHostToDevice));
	c->cap_t = (double *)dmalloc(sizeof(double) * 6 * (qtot > 0 ? qtot : 1));

	/* S columns: sweeps of zero data with unit lagged inputs */
	const size_t blen = (size_t)c->m * c->neq;
	c->wz = dmalloc(c->rsz * blen);
	void *zero = c->wx, *out = c->w[1 - c->cur];
	HIPCHECK(hipMemsetAsync(zero, 0, c->rsz * blen, c->stream));
	if (c->cap_gmres) {
		HIPCHECK(hipMemsetAsync(c->wz, 0, c->rsz * blen, c->stream));
		c->cap_zh.assign(6 * (size_t)qtot, 0.0);
		HIPCHECK(hipStreamSynchronize(c->stream));
		return;
	}
	std::vector<long> luoff(nv + 1, 0);
	std::vector<int> piv(qtot > 0 ? qtot : 1);
	for (int v = 0; v < nv; v++) {
		const int q = off[v + 1] - off[v];
		luoff[v + 1] = luoff[v] + (long)q * q;
		for (int r = 0; r < q; r++)
			piv[off[v] + r] = r;
	}
	c->cap_lu = (double *)dmalloc(sizeof(double) * (luoff[nv] > 0 ? luoff[nv] : 1));
	c->cap_luoff = (long *)dmalloc(sizeof(long) * (nv + 1));
	HIPCHECK(hipMemcpy(c->cap_luoff, luoff.data(), sizeof(long) * (nv + 1),
			   hipMemcpyHostToDevice));
	c->cap_piv = (int *)dmalloc(sizeof(int) * piv.size());
	HIPCHECK(hipMemcpy(c->cap_piv, piv.data(), sizeof(int) * piv.size(),
			   hipMemcpyHostToDevice));
	const int nbatch = (qmax + 5) / 6;
	for (int bt = 0; bt < nbatch; bt++) {
		HIPCHECK(hipMemsetAsync(c->wz, 0, c->rsz * blen, c->stream));
		if (c->fp32)
			k_cap_unit<float><<<nblocks(6 * nv, 64), 64, 0, c->stream>>>(
				(float *)c->wz, c->cap_ipg, c->cap_off, bt, nv, c->ne, c->neq);
		else
			k_cap_unit<double><<<nblocks(6 * nv, 64), 64, 0, c->stream>>>(
				(double *)c->wz, c->cap_ipg, c->cap_off, bt, nv, c->ne, c->neq);
		HIPCHECK(hipGetLastError());
		sweep_step(c, zero, out, c->wz, 0);
		cap_gather(c, out);
		k_cap_column<<<dim3(nblocks(6 * qmax, 256), nv), 256, 0, c->stream>>>(
			c->cap_t, c->cap_off, c->cap_luoff, bt, qtot, c->cap_lu);
		HIPCHECK(hipGetLastError());
	}
	HIPCHECK(hipMemsetAsync(c->wz, 0, c->rsz * blen, c->stream));
	/* X = (P (I - T S))^-1 in place */
	c->cap_rk = (double *)dmalloc(sizeof(double) * 2 * (size_t)qmax * nv);
	int *sing = (int *)dmalloc(sizeof(int));
	HIPCHECK(hipMemsetAsync(sing, 0, sizeof(int), c->stream));
	const dim3 ug(nblocks(qmax, 256), (qmax + GJ_JC - 1) / GJ_JC, nv);
	for (int k = 0; k < qmax; k++) {
		k_gj_pivot<<<nv, 1024, 0, c->stream>>>(c->cap_lu, c->cap_luoff, c->cap_off,
							c->cap_piv, c->cap_rk, qmax, k, sing);
		k_gj_update<<<ug, 256, 0, c->stream>>>(c->cap_lu, c->cap_luoff, c->cap_off,
						       c->cap_rk, qmax, k);
	}
	HIPCHECK(hipGetLastError());
	int hsing = 0;
	HIPCHECK(hipMemcpyAsync(&hsing, sing, sizeof(int), hipMemcpyDeviceToHost, c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
	HIPCHECK(hipFree(sing));
	if (hsing) {
		fprintf(stderr, "gdon3d-hip: singular capacitance matrix of the "
				"transport over upwind cycles\n");
		abort();
	}
	const int nch = (qmax + CAP_JC - 1) / CAP_JC;
	c->cap_part = (double *)dmalloc(sizeof(double) * (size_t)nv * nch * 6 * qmax);
}

/*
 * Matrix-free form of the same solve, for velocities with more lagged nodes
 * than the dense inverse takes: restarted GMRES (Saad and Schultz, 1986; modified
 * Gram-Schmidt, Givens rotations) on (I - T S) z = T x0. The operator is
 * block diagonal, one system per (velocity, field) as above, so the systems
 * run in lockstep, each with its own Krylov basis and Hessenberg matrix,
 * and one product S z for all of them is one sweep of zero data with
 * lagged inputs z. The Krylov vectors (6 q values each) live on the host;
 * each iteration moves them through the device once. Started from the last
 * step's lagged values; stops when every system's true residual is at most
 * tol |(T x0)_k| + 1e-16 max_k |(T x0)_k| (tol 1e-14 in binary64, 2e-6 in
 * binary32), when a restart cycle no longer lowers it (rounding level), or
 * after GMRES_MAX_SWEEPS products.
 */
#define GMRES_RESTART 30
#define GMRES_MAX_SWEEPS 1200

template <typename R>
__global__ void k_cap_scatter(const double *__restrict__ z, const int *__restrict__ ipg,
			      const int *__restrict__ vel, int q, int nv, long ne, long neq,
			      R *__restrict__ wz)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= 6 * q)
		return;
	const int iw = k / q, r = k % q;
	wz[kin_at(ne, neq, nv, ipg[r], iw, vel[r])] = (R)z[k];
}

/* z (host, 6 q values, field-major) into the lagged-value buffer */
static void cap_put(gdh_ctx *c, const double *z)
{
	const int q = c->cap_qtot;
	HIPCHECK(hipMemcpyAsync(c->cap_t, z, sizeof(double) * 6 * q, hipMemcpyHostToDevice,
				c->stream));
	if (c->fp32)
		k_cap_scatter<float><<<nblocks(6 * q, 256), 256, 0, c->stream>>>(
			c->cap_t, c->cap_ipg, c->cap_vel, q, c->kp.nv, c->ne, c->neq, (float *)c->wz);
	else
		k_cap_scatter<double><<<nblocks(6 * q, 256), 256, 0, c->stream>>>(
			c->cap_t, c->cap_ipg, c->cap_vel, q, c->kp.nv, c->ne, c->neq, (double *)c->wz);
	HIPCHECK(hipGetLastError());
}

/* y = (I - T S) z for every system; out: a state buffer to sweep into */
static void cap_apply(gdh_ctx *c, const double *z, double *y, void *out)
{
	const size_t n = 6 * (size_t)c->cap_qtot;
	cap_put(c, z);
	sweep_step(c, c->wx, out, c->wz, 0);
	cap_gather(c, out);
	HIPCHECK(hipMemcpyAsync(y, c->cap_t, sizeof(double) * n, hipMemcpyDeviceToHost,
				c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
	for (size_t k = 0; k < n; k++)
		y[k] = z[k] - y[k];
}

/* The lagged values from T x0 (in cap_t), left in wz; returns the number
 * of sweeps made. out: a state buffer to sweep into. */
static int cap_gmres(gdh_ctx *c, void *out)
{
	const int nv = c->kp.nv, q = c->cap_qtot, nb = 6 * nv, M = GMRES_RESTART;
	const size_t n = 6 * (size_t)q;
	const std::vector<int> &off = c->cap_off_h;
	/* the Krylov basis is kept by the context (no zeroing per call: every
	 * vector is written before it is read) */
	std::vector<double> &V = c->cap_krylov;
	if (V.size() < (size_t)(M + 1) * n)
		V.resize((size_t)(M + 1) * n);
	std::vector<double> b(n), r(n), w(n);
	HIPCHECK(hipMemcpyAsync(b.data(), c->cap_t, sizeof(double) * n, hipMemcpyDeviceToHost,
				c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
	std::vector<double> &x = c->cap_zh;
	/* system k = iw nv + v: values [iw q + off[v], iw q + off[v + 1]) */
	auto lo = [&](int k) { return (size_t)(k / nv) * q + off[k % nv]; };
	auto hi = [&](int k) { return (size_t)(k / nv) * q + off[k % nv + 1]; };
	auto dot = [&](const double *a, const double *d, int k) {
		double s = 0;
		for (size_t i = lo(k); i < hi(k); i++)
			s += a[i] * d[i];
		return s;
	};
	const double tol = c->fp32 ? 2e-6 : 1e-14;
	std::vector<double> tolk(nb), beta(nb), H((size_t)nb * (M + 1) * M), cs((size_t)nb * M),
		sn((size_t)nb * M), g((size_t)nb * (M + 1));
	std::vector<int> ncol(nb);
	std::vector<char> act(nb);
	double bmax = 0;
	for (int k = 0; k < nb; k++) {
		tolk[k] = std::sqrt(dot(b.data(), b.data(), k));
		bmax = std::max(bmax, tolk[k]);
	}
	if (bmax == 0) { /* T x0 = 0: z = 0 */
		std::fill(x.begin(), x.end(), 0.0);
		cap_put(c, x.data());
		return 0;
	}
	for (int k = 0; k < nb; k++)
		tolk[k] = tol * tolk[k] + 1e-16 * bmax;
	int sweeps = 0;
	double prev = HUGE_VAL;
	for (;;) {
		cap_apply(c, x.data(), r.data(), out);
		sweeps++;
		int nact = 0;
		double worst = 0; /* largest residual relative to its tolerance */
		for (int k = 0; k < nb; k++) {
			for (size_t i = lo(k); i < hi(k); i++)
				r[i] = b[i] - r[i];
			beta[k] = std::sqrt(dot(r.data(), r.data(), k));
			act[k] = beta[k] > tolk[k];
			nact += act[k];
			if (tolk[k] > 0)
				worst = std::max(worst, beta[k] / tolk[k]);
		}
		if (!nact)
			break;
		if (sweeps >= GMRES_MAX_SWEEPS || worst > 0.9 * prev) {
			if (!c->cap_warned)
				fprintf(stderr, "gdon3d-hip: GMRES over the upwind cycles stopped "
						"at %.2g times its tolerance after %d sweeps\n",
					worst, sweeps);
			c->cap_warned = 1;
			break;
		}
		prev = worst;
		std::fill(V.begin(), V.begin() + n, 0.0);
		for (int k = 0; k < nb; k++) {
			ncol[k] = 0;
			std::fill(g.begin() + (size_t)k * (M + 1), g.begin() + (size_t)(k + 1) * (M + 1),
				  0.0);
			if (!act[k])
				continue;
			g[(size_t)k * (M + 1)] = beta[k];
			for (size_t i = lo(k); i < hi(k); i++)
				V[i] = r[i] / beta[k];
		}
		for (int j = 0; j < M && nact; j++) {
			double *vj = V.data() + (size_t)j * n, *vn = vj + n;
			cap_apply(c, vj, w.data(), out);
			sweeps++;
			std::fill(vn, vn + n, 0.0);
			for (int k = 0; k < nb; k++) {
				if (!act[k])
					continue;
				double *h = H.data() + (size_t)k * (M + 1) * M; /* h[i M + j] */
				for (int i = 0; i <= j; i++) {
					const double *vi = V.data() + (size_t)i * n;
					const double hij = dot(w.data(), vi, k);
					h[i * M + j] = hij;
					for (size_t l = lo(k); l < hi(k); l++)
						w[l] -= hij * vi[l];
				}
				const double hn = std::sqrt(dot(w.data(), w.data(), k));
				h[(j + 1) * M + j] = hn;
				if (hn > 0)
					for (size_t l = lo(k); l < hi(k); l++)
						vn[l] = w[l] / hn;
				double *ck = cs.data() + (size_t)k * M, *sk = sn.data() + (size_t)k * M;
				double *gk = g.data() + (size_t)k * (M + 1);
				for (int i = 0; i < j; i++) {
					const double a = h[i * M + j], d = h[(i + 1) * M + j];
					h[i * M + j] = ck[i] * a + sk[i] * d;
					h[(i + 1) * M + j] = -sk[i] * a + ck[i] * d;
				}
				const double a = h[j * M + j], d = h[(j + 1) * M + j];
				const double rho = std::hypot(a, d);
				ck[j] = rho > 0 ? a / rho : 1.0;
				sk[j] = rho > 0 ? d / rho : 0.0;
				h[j * M + j] = rho;
				h[(j + 1) * M + j] = 0;
				gk[j + 1] = -sk[j] * gk[j];
				gk[j] = ck[j] * gk[j];
				ncol[k] = j + 1;
				if (std::fabs(gk[j + 1]) <= tolk[k] || hn == 0) {
					act[k] = 0;
					nact--;
				}
			}
		}
		/* x_k += V_k y_k with H_k y_k = g_k (upper triangular) */
		for (int k = 0; k < nb; k++) {
			const int nc = ncol[k];
			if (!nc)
				continue;
			const double *h = H.data() + (size_t)k * (M + 1) * M;
			double y[GMRES_RESTART];
			for (int i = nc - 1; i >= 0; i--) {
				double s = g[(size_t)k * (M + 1) + i];
				for (int l = i + 1; l < nc; l++)
					s -= h[i * M + l] * y[l];
				y[i] = s / h[i * M + i];
			}
			for (int i = 0; i < nc; i++) {
				const double *vi = V.data() + (size_t)i * n;
				for (size_t l = lo(k); l < hi(k); l++)
					x[l] += y[i] * vi[l];
			}
		}
	}
	cap_put(c, x.data());
	return sweeps;
}

static void cap_solve(gdh_ctx *c)
{
	const int nv = c->kp.nv, qmax = c->cap_qmax();
	const int nch = (qmax + CAP_JC - 1) / CAP_JC;
	k_cap_product<<<dim3(nblocks(qmax, 256), nch, nv), 256, 0, c->stream>>>(
		c->cap_lu, c->cap_luoff, c->cap_off, c->cap_piv, c->cap_t, c->cap_qtot, qmax,
		c->cap_part);
	HIPCHECK(hipGetLastError());
	const dim3 rg(nblocks(6 * qmax, 256), nv);
	if (c->fp32)
		k_cap_reduce<float><<<rg, 256, 0, c->stream>>>(c->cap_part, c->cap_off, c->cap_ipg,
							      qmax, nv, c->ne, c->neq,
							      (float *)c->wz);
	else
		k_cap_reduce<double><<<rg, 256, 0, c->stream>>>(c->cap_part, c->cap_off,
							       c->cap_ipg, qmax, nv, c->ne,
							       c->neq, (double *)c->wz);
	HIPCHECK(hipGetLastError());
}

/*
 * One step of the kinetic scheme: transport, then the node-local relaxation
 * (src/dg/relaxation.c:130-160). The relaxation is applied by the next
 * step's sweep as it reads its input (each item relaxes its own values
 * from those of every velocity at its nodes; consumers see the published
 * averaged traces, never a neighbour's stored values); after a step the
 * state holds the transport's output and relax_pending is set.
 * relax_materialize applies it in place wherever the state itself is needed
 * (download, L2 error, relaxation of every DoF, and before the sweeps over
 * upwind cycles, whose lagged faces read neighbours' stored values).
 */
extern "C" void gdh_relax_step(gdh_ctx *c)
{
	/* unless fused (see relax_choose), and over upwind cycles (lagged
	 * faces read the neighbour's stored values), the relaxation of the last
	 * step is applied before the sweep */
	if (c->relax_fused != 1)
		relax_materialize(c);
	void *fo = c->w[c->cur], *fn = c->w[1 - c->cur];
	const int rin = c->relax_pending;
	/* Ghost DoFs are identity rows of M1 and M2; the sweep reads them from
	 * fo and nothing reads the ghost region of fn before the next ghost
	 * update overwrites it. */
	prof_begin(c, 0);
	if (!c->nlagged) {
		sweep_step(c, fo, fn, nullptr, rin);
	} else {
		/* exact over the upwind cycles: two sweeps around a solve for
		 * the lagged values (dense inverse, or GMRES with sweeps as products) */
		sweep_step(c, fo, fn, nullptr, rin);
		cap_gather(c, fn);
		int extra = 0;
		if (c->cap_gmres)
			extra = cap_gmres(c, fn);
		else
			cap_solve(c);
		sweep_step(c, fo, fn, c->wz, rin);
		c->lag_last = 2 + extra;
		c->lag_max = std::max(c->lag_max, c->lag_last);
	}
	prof_end(c, 0);
	/* relaxed by the elements' last items, or pending */
	c->relax_pending = c->relax_fused != 2;
	c->cur = 1 - c->cur;
	if (!c->relax_fused)
		relax_materialize(c);
}

/*
 * Form of the sweep and of the relaxation, by measurement on this mesh.
 * The item sweep (one item per workgroup, every value handed over through
 * memory) can apply the relaxation as it reads its input (fused): that
 * saves the relaxation kernel's pass over the state but gives every sweep
 * item the loads of all velocities at its nodes and the relaxation
 * arithmetic. Or each element's last item in the level order relaxes the
 * whole element once its other velocities are done (LF, binary64, only on
 * request, GDN_RELAX_FUSED=2): the other items store no state, the last one
 * reads their published traces and old values and stores every velocity
 * relaxed, and there is no pass; its extra work sits on the item-turnover
 * path and it was slower than the best other form at every size measured
 * (item sweep per step: 8^3 0.120 vs 0.108 ms fused, 16^3 0.247 vs 0.227
 * fused, 24^3 0.503 vs 0.479 fused and 0.511 separate, 32^3 0.883 vs 0.849
 * separate, 64^3 6.58 vs 5.95 separate). The cluster sweep takes a relaxed
 * state. Each candidate (item sweep with the relaxation separate or fused,
 * cluster sweep with it separate) is timed twice, five steps after one each
 * time, from the current state into the other buffer, and its better time
 * counts; nothing of the state changes. GDN_SWEEP and GDN_RELAX_FUSED=0|1|2
 * restrict the candidates.
 */
static void relax_choose(gdh_ctx *c)
{
	const char *env = getenv("GDN_RELAX_FUSED");
	/* lagged faces and the projector boundary state read stored element
	 * values, which must hold the relaxation (lagged faces: the transport
	 * output): no fused or LF form; LF also needs binary64 (its recovery of
	 * the other velocities' output from the published traces rounds at the
	 * state precision), the stored inverses, whole-trace polling and the
	 * static item order */
	const int fixed = env ? atoi(env) : -1;
	const bool any_fused = !(c->nlagged || c->kp.fr_kind == GDH_FR_PROJECTOR);
	const bool lf_ok = any_fused && !c->fp32 && !c->kp.lb_p && c->sweep_poll && !c->sweep_queue;
	if (fixed == 2 && !lf_ok) {
		fprintf(stderr, "gdon3d-hip: GDN_RELAX_FUSED=2 needs binary64, the stored inverses, "
				"whole-trace polling, the static order and no upwind cycles or "
				"projector boundary\n");
		abort();
	}
	int cand[4][2], nc = 0; /* (sweep mode, relaxation form) */
	for (int mode = 0; mode <= 2; mode++) {
		if (c->sweep_mode >= 0 ? mode != c->sweep_mode : mode == 0)
			continue;
		if (mode == 2 && !c->ncl)
			continue;
		for (int form = 0; form < 3; form++) {
			if (fixed >= 0 && form != fixed)
				continue;
			if (form && (mode != 1 || !any_fused))
				continue;
			if (form == 2 && (!lf_ok || fixed != 2))
				continue;
			cand[nc][0] = mode;
			cand[nc][1] = form;
			nc++;
		}
	}
	if (nc == 0) {
		fprintf(stderr, "gdon3d-hip: no form of the sweep fits GDN_SWEEP / GDN_RELAX_FUSED\n");
		abort();
	}
	/* the LF kernel's own residency (its registers differ) */
	for (int k = 0; k < nc; k++)
		if (cand[k][1] == 2 && !c->sweep_grid_lf) {
			c->relax_fused = 2;
			c->sweep_grid_lf = grid_cap(std::min(c->sweep_grid, sweep_residency<double>(c)));
			c->relax_fused = 0;
		}
	int best = 0;
	if (nc > 1) {
		void *fo = c->w[c->cur], *fn = c->w[1 - c->cur];
		const int prof = c->prof;
		c->prof = 0;
		/* two rounds over the candidates, the best time of each kept: the
		 * first launches after set-up run slower, which would otherwise
		 * count against whichever candidate is timed first */
		float tms[4] = { 1e30f, 1e30f, 1e30f, 1e30f };
		for (int round = 0; round < 2; round++)
			for (int k = 0; k < nc; k++) {
				c->sweep_mode = cand[k][0];
				c->relax_fused = cand[k][1];
				for (int r = 0; r < 6; r++) {
					if (r == 1)
						HIPCHECK(hipEventRecord(c->ev0, c->stream));
					sweep_step(c, fo, fn, nullptr, cand[k][1] == 1);
					if (!cand[k][1])
						relax_launch(c, fn, 10L * c->ne);
				}
				HIPCHECK(hipEventRecord(c->ev1, c->stream));
				HIPCHECK(hipEventSynchronize(c->ev1));
				float ms = 0;
				HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
				if (getenv("GDN_VERBOSE"))
					fprintf(stderr, "gdon3d-hip: sweep form %d, relaxation form %d: "
							"%.3f ms per step\n",
						cand[k][0], cand[k][1], ms / 5);
				tms[k] = std::min(tms[k], ms);
			}
		for (int k = 1; k < nc; k++)
			if (tms[k] < tms[best])
				best = k;
		c->prof = prof;
	}
	c->sweep_mode = cand[best][0];
	c->relax_fused = cand[best][1];
	/* the words are made consistent for the chosen form now, before any
	 * step can be captured */
	if (c->ll_partial && c->sweep_mode != 2)
		ll_restart(c);
}

__global__ void k_flag_learned(ItemDesc *__restrict__ desc, const int *__restrict__ ok,
			       int nitems)
{
	const int it = blockIdx.x * blockDim.x + threadIdx.x;
	if (it >= nitems)
		return;
	int fl = desc[it].flags & ~(LB_FLAG | (3 << 30));
	if (ok && ok[it] >= 2) /* refinement steps, 2 .. 5 */
		fl |= LB_FLAG | ((ok[it] - 2) << 30);
	desc[it].flags = fl;
}

/* descriptor flags of the items on the learned inverse, ok[it] their
 * refinement steps (0 or null: the stored inverse);
 * the learned inverses are refined by the item and level sweeps, so a
 * context on the cluster sweep moves to the item sweep (the form of the
 * relaxation chosen again by timing) */
void sweep_set_learned(gdh_ctx *c, const int *ok)
{
	k_flag_learned<<<nblocks(c->nitems, 256), 256, 0, c->stream>>>(c->desc, ok, c->nitems);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipStreamSynchronize(c->stream));
	/* the kernel of the item sweep changes with kp.lb_p: its residency is
	 * measured again */
	c->sweep_grid = grid_cap(c->fp32 ? sweep_residency<float>(c) : sweep_residency<double>(c));
	if (ok && c->sweep_mode == 2)
		c->sweep_mode = 1;
	if (c->sweep_mode != 0)
		relax_choose(c);
}

extern "C" int gdh_relax_fused(gdh_ctx *c)
{
	return c->relax_fused;
}

void relax_materialize(gdh_ctx *c)
{
	if (!c->scheme || !c->relax_pending)
		return;
	prof_begin(c, 1);
	relax_launch(c, c->w[c->cur], 10L * c->ne);
	prof_end(c, 1);
	c->relax_pending = 0;
}

/* relaxation of every DoF, ghosts included (src/dg/dg.c:109-147) */
extern "C" void gdh_apply_relaxation(gdh_ctx *c)
{
	relax_materialize(c);
	const long n = c->neq;
	relax_launch(c, c->w[c->cur], n);
}


extern "C" int gdh_sweep_grid(gdh_ctx *c)
{
	return c->sweep_mode == 2 ? c->cl_grid : c->sweep_grid;
}

extern "C" void gdh_sweep_info(gdh_ctx *c, int out[5])
{
	out[0] = c->sweep_mode;
	out[1] = c->ncl;
	out[2] = c->cl_levels;
	out[3] = (int)std::min(c->cl_multi, (long)2147483647);
	out[4] = c->sweep_mode == 2 ? c->cl_grid : c->sweep_grid;
}

extern "C" void gdh_transport_stats(gdh_ctx *c, int out[3])
{
	out[0] = c->nlagged;
	out[1] = c->lag_last;
	out[2] = c->lag_max;
}

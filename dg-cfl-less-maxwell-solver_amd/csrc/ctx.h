/*
 * Device context of the GPU layer and helpers shared by its translation
 * units (gdh_core.hip, rk3.hip, sweep.hip).
 */
#ifndef GDH_CTX_H
#define GDH_CTX_H

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

/* Instances of a batch (several meshes as one disjoint union): per mesh
 * element and per boundary face the instance, per instance the exact-
 * solution parameters, time step and clock; RK3 also per storage column
 * the instance and 1/eps. ninst == 0: not a batch. Passed by value. */
struct BatchTab {
	int ninst;
	const int *einst, *binst, *einst_col;
	const double *par; /* GDH_IMPOSED_NPAR per instance */
	const double *dt;
	double *time;
	const double *ieps_col; /* RK3: 1 / eps per storage column, or null */
};

struct gdh_ctx {
	int ne, nb, neq, m, scheme;
	int fp32; /* state stored (and fluxes computed) in binary32 */
	size_t rsz; /* bytes per state value: 8, or 4 with fp32 */
	gdh_model_desc md;
	ImposedPar ip;
	KinPar kp;
	hipStream_t stream;
	hipEvent_t ev0, ev1;

	double *geom;
	int4 *nbr, *nbmap;
	int2 *bface; /* kinetic: per boundary face, its element and local face */
	double *ieps; /* 1/eps per element (kinetic: mesh order; RK3: by storage
		       * column), or null */
	double *lb_x; /* learned block inverses: 8 numbers per sweep item */
	unsigned *lb_p; /* and the bfloat16 predictions, 50 words per item */
	double *xnode, *xghost;
	/* RK3: w = structure-of-arrays interior state W[q][i][e], gh = ghosts
	 * GH[q][6 bf + a]; kinetic: w = item-major state (kin_at), ghosts
	 * included.
	 * Elements are double, or float with fp32. */
	void *w[2];
	int cur;
	void *k2;
	void *gh;
	double *geom_soa; /* G[k][c], k < 10, by storage column c */
	int *elem_pos; /* RK3: storage column of each mesh element */
	double *stage; /* array-of-structures staging for state transfers */
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
	struct ItemDesc *desc; /* per item: element, velocity, inflow faces */
	unsigned long long *ll; /* per item: 60 hand-off words, one per value in the
				 * state precision, epoch parity in the lowest bit */
	int ll_partial; /* the last sweep left some items' words unwritten */
	unsigned *queue; /* item counters of the persistent sweep (two slots) */
	int qslot; /* the slot of the next launch */
	int sweep_queue; /* 1: items taken from the counter, 0: static order */
	int coop_refused; /* the runtime refused a cooperative launch */
	int *status; /* nonzero: a dependency wait timed out */
	unsigned epoch;
	int sweep_grid; /* resident blocks of the persistent sweep */
	int sweep_grid_lf; /* the same with the relaxation by the element's last
			    * item (0: not measured) */
	unsigned *epoch_dev; /* batches: sweep epoch in device memory (graphs) */

	/* batch of instances; graphs of two steps (by state parity) */
	BatchTab bt;
	int *bt_mem[4];
	double *bt_dmem[4];
	std::vector<double> bt_dt, bt_par;
	std::vector<int> bt_e0, bt_nb0; /* first element / boundary face per instance */
	hipGraphExec_t graph[2];
	int rk3_bs, relax_bs; /* tile sizes (autotuned for batches) */
	int relax_pair; /* relaxation pass: two elements per wave (k_relax_pair)
			 * or one (k_relax) */
	int rk3_tile; /* RK3 stage: 0 global state, 1 LDS tile, 2 LDS tile with
		       * two lanes per element */
	int rk3_ntk; /* RK3 stage: k2 with non-temporal loads and stores (1, the
		      * default; batches time both) or ordinary ones (0) */
	long nlagged; /* inflow faces lagged to break upwind cycles */
	void *wx; /* zero state (set-up of the capacitance matrices) */
	void *wz; /* lagged values z, scattered into a state-shaped buffer */
	int lag_last, lag_max; /* sweeps per step: last step, largest */
	/* capacitance form over the cycles: distinct lagged nodes (ipg and
	 * velocity), per-velocity offsets, the inverse of the row-permuted
	 * (I - T S) column-major with its row permutation, partial products */
	int cap_qtot;
	std::vector<int> cap_off_h;
	int *cap_ipg, *cap_vel, *cap_off, *cap_piv;
	double *cap_lu, *cap_t, *cap_part, *cap_rk;
	long *cap_luoff;
	/* matrix-free form (more lagged nodes than the dense inverse takes): GMRES
	 * on the same systems, started from the last step's lagged values */
	int cap_gmres;
	std::vector<double> cap_zh, cap_krylov;
	int cap_warned; /* GMRES stopped above its tolerance (reported once) */
	int cap_qmax() const
	{
		int m = 0;
		for (size_t v = 0; v + 1 < cap_off_h.size(); v++)
			m = std::max(m, cap_off_h[v + 1] - cap_off_h[v]);
		return m;
	}
	int sweep_mode; /* 2 clusters, 1 persistent items, 0 one launch per level */
	int sweep_poll; /* 1: waiting items poll whole traces */
	int relax_pending; /* the state awaits the relaxation of its last step */
	int relax_fused; /* 1: the sweep applies the relaxation as it reads its
			  * input; 2: each element's last item relaxes the
			  * element's output (LF); 0: a pass after the sweep */
	/* cluster sweep: host copies of the element vertices (binary32, 12 per
	 * element) and neighbours (mesh order), taken at creation; per
	 * cluster (v, elements) a header {first item, n | nx << 8, external
	 * faces}, item descriptors in cluster order */
	std::vector<float> vtx;
	std::vector<int> nbr_h;
	int ncl;
	int4 *chead;
	struct ItemDesc *cdesc;
	int cl_grid;
	long cl_multi; /* items in clusters of more than one element */
	int cl_levels;

	/* optional event timing of the main kernels: pairs of events recorded
	 * on the stream around each kernel group, resolved when read */
	int prof;
	std::vector<hipEvent_t> ev_pool;
	std::vector<int> ev_slot; /* slot of each recorded pair */
	std::vector<hipEvent_t> ev_pairs; /* begin, end, begin, end, ... */
	double prof_s[3];
	long prof_launches;
};

static inline hipEvent_t ev_get(gdh_ctx *c)
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

static inline void prof_begin(gdh_ctx *c, int slot)
{
	if (!c->prof)
		return;
	hipEvent_t e = ev_get(c);
	HIPCHECK(hipEventRecord(e, c->stream));
	c->ev_pairs.push_back(e);
	c->ev_slot.push_back(slot);
}

static inline void prof_end(gdh_ctx *c, int slot)
{
	(void)slot;
	if (!c->prof)
		return;
	hipEvent_t e = ev_get(c);
	HIPCHECK(hipEventRecord(e, c->stream));
	c->ev_pairs.push_back(e);
}

/* wait for the recorded pairs and accumulate their durations */
static inline void prof_resolve(gdh_ctx *c)
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

static inline void *dmalloc(size_t bytes)
{
	void *p = nullptr;
	HIPCHECK(hipMalloc(&p, bytes ? bytes : 16));
	return p;
}

static inline int nblocks(long n, int bs)
{
	return (int)((n + bs - 1) / bs);
}

/* Where the macro state of DoF ipg (component q) lives, RK3 layout. */
/* RK3 state: element DoFs in SoA, element e of the mesh stored at column
 * pos[e] (a locality order, see elem_order in gdh_core.hip), ghosts apart;
 * T = double or float */
template <typename T> struct StateRef {
	T *w, *gh;
	const int *pos;
	long ne, ngh;
};

template <typename T>
__device__ inline T *macro_ptr(const StateRef<T> &s, long ipg, int q)
{
	if (ipg < 10 * s.ne)
		return s.w + (q * 10 + ipg % 10) * s.ne + s.pos[ipg / 10];
	return s.gh + q * s.ngh + (ipg - 10 * s.ne);
}

/*
 * Kinetic state on the device, item-major: the 60 values of item (v, e)
 * (6 fields x 10 nodes) are contiguous, K[(v ne + e) 60 + 10 iw + i], and
 * the ghost values follow, K[60 nv ne + (v nb + b) 36 + 6 iw + a]. The
 * transport sweep reads and writes whole items and face traces of one
 * element, which this layout keeps within 480 contiguous bytes. The
 * reference layout f[(iw nv + v) neq + ipg] (src/simulation/simulation.c:22-40,
 * src/dg/relaxation.c:106) is restored at host transfers. neq = 10 ne + 6 nb.
 */
__host__ __device__ inline long kin_at(long ne, long neq, int nv, long ipg, int iw, int v)
{
	const long ne10 = 10 * ne;
	if (ipg < ne10)
		return ((long)v * ne + ipg / 10) * 60 + 10 * iw + ipg % 10;
	const long g = ipg - ne10, a = g % 6;
	return 60L * nv * ne + 6L * v * (neq - ne10) + 6 * (g - a) + 6 * iw + a;
}

/* Cluster plan of the transport sweep (cluster.hip): clusters of at most
 * CL_MAX elements of one velocity with at most CL_XMAX inflow faces from
 * outside, in a topological order of the cluster graph. */
#define CL_MAX 6
#define CL_XMAX 8
struct ClusterPlan {
	std::vector<int4> head; /* { first item, n | nx << 8, external faces
				 * (byte x = 4 j + f: face f of item j) } */
	std::vector<int> perm; /* cluster-order item -> level-order item */
	std::vector<int> code; /* per item: 4 bits per face, external face x
				* (0..7) or 8 + in-cluster producer j, and bit
				* 16 set when a downstream neighbour lies in
				* another cluster; at bit 12 of the descriptor
				* flags */
	long multi; /* items in clusters of more than one element */
	int levels; /* longest path of the cluster graphs */
};
struct gdh_ctx;
bool cluster_plan(const gdh_ctx *c, const uint8_t *cls, const int *items_ve,
		  const int *level_ptr, int nlevels, ClusterPlan &P);

/* launchers implemented in rk3.hip and sweep.hip */
void sweep_setup(gdh_ctx *c);
void sweep_step(gdh_ctx *c, void *fo, void *fn, const void *xl, int relax_in);
void relax_materialize(gdh_ctx *c);
void relax_launch(gdh_ctx *c, void *f, long n);
void block_inverse_launch(gdh_ctx *c);
void set_permittivity(gdh_ctx *c, const double *eps);
void check_sweep_status(gdh_ctx *c);
void sweep_set_learned(gdh_ctx *c, const int *ok);

#endif

/*
 * Batches: several independent instances (meshes of any size, each with its
 * own exact solution, time step and, for RK3, permittivity per element) in
 * one context whose mesh is their disjoint union, so that every kernel of a
 * step covers all instances and all velocities in one launch. Steps are
 * captured once into a HIP graph (two steps, so that the state buffers are
 * back in place) and replayed; the tile sizes of the RK3 stage and of the
 * relaxation are chosen by timing the candidates on the batch itself.
 */
#include "ctx.h"

using namespace gdh;

void ghost_fill_batch(gdh_ctx *c, double cstage);
double error_range(gdh_ctx *c, int e0, int e1, double t, const ImposedPar &ip,
		   const double *hv);
void rk3_trial(gdh_ctx *c, int first);

__global__ void k_time_bump(double *__restrict__ t, const double *__restrict__ dt, int n)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k < n)
		t[k] += dt[k];
}

extern "C" void gdh_batch_setup(gdh_ctx *c, int ninst, const int *elem_first,
				const int *bface_first, const double *par,
				const double *dt, const double *eps)
{
	if (ninst < 1 || elem_first[ninst] != c->ne || bface_first[ninst] != c->nb) {
		fprintf(stderr, "gdon3d-hip: batch ranges do not cover the mesh\n");
		abort();
	}
	if (c->ip.kind == GDH_IMPOSED_HOST) {
		fprintf(stderr, "gdon3d-hip: batches need closed-form exact data\n");
		abort();
	}

	const int ne = c->ne, nb = c->nb;
	std::vector<int> einst(ne), binst(nb > 0 ? nb : 1, 0), ecol(ne);
	for (int i = 0; i < ninst; i++) {
		for (int e = elem_first[i]; e < elem_first[i + 1]; e++)
			einst[e] = i;
		for (int b = bface_first[i]; b < bface_first[i + 1]; b++)
			binst[b] = i;
	}
	if (!c->scheme) {
		std::vector<int> pos(ne);
		HIPCHECK(hipMemcpy(pos.data(), c->elem_pos, sizeof(int) * ne,
				   hipMemcpyDeviceToHost));
		for (int e = 0; e < ne; e++)
			ecol[pos[e]] = einst[e];
	}
	c->bt_mem[0] = (int *)dmalloc(sizeof(int) * ne);
	c->bt_mem[1] = (int *)dmalloc(sizeof(int) * binst.size());
	HIPCHECK(hipMemcpy(c->bt_mem[0], einst.data(), sizeof(int) * ne, hipMemcpyHostToDevice));
	HIPCHECK(hipMemcpy(c->bt_mem[1], binst.data(), sizeof(int) * binst.size(),
			   hipMemcpyHostToDevice));
	if (!c->scheme) {
		c->bt_mem[2] = (int *)dmalloc(sizeof(int) * ne);
		HIPCHECK(hipMemcpy(c->bt_mem[2], ecol.data(), sizeof(int) * ne,
				   hipMemcpyHostToDevice));
	}
	c->bt_par.assign(par, par + (size_t)GDH_IMPOSED_NPAR * ninst);
	c->bt_dt.assign(dt, dt + ninst);
	c->bt_dmem[0] = (double *)dmalloc(sizeof(double) * c->bt_par.size());
	c->bt_dmem[1] = (double *)dmalloc(sizeof(double) * ninst);
	c->bt_dmem[2] = (double *)dmalloc(sizeof(double) * ninst);
	HIPCHECK(hipMemcpy(c->bt_dmem[0], par, sizeof(double) * c->bt_par.size(),
			   hipMemcpyHostToDevice));
	HIPCHECK(hipMemcpy(c->bt_dmem[1], dt, sizeof(double) * ninst, hipMemcpyHostToDevice));
	HIPCHECK(hipMemset(c->bt_dmem[2], 0, sizeof(double) * ninst));
	c->bt.ninst = ninst;
	c->bt.einst = c->bt_mem[0];
	c->bt.binst = c->bt_mem[1];
	c->bt.einst_col = c->bt_mem[2];
	c->bt.par = c->bt_dmem[0];
	c->bt.dt = c->bt_dmem[1];
	c->bt.time = c->bt_dmem[2];
	/* permittivity per element: RK3 per storage column, kinetic per
	 * element (the equilibrium's flux on (D / eps, B)) */
	if (eps)
		set_permittivity(c, eps);
	c->bt_e0.assign(elem_first, elem_first + ninst + 1);
	c->bt_nb0.assign(bface_first, bface_first + ninst + 1);
	if (c->scheme) {
		c->epoch_dev = (unsigned *)dmalloc(sizeof(unsigned));
		HIPCHECK(hipMemset(c->epoch_dev, 0, sizeof(unsigned)));
	}
}

/* mean device time of `reps` launches of f after one warm-up, ms */
template <typename F> static float time_launches(gdh_ctx *c, int reps, F f)
{
	f();
	HIPCHECK(hipEventRecord(c->ev0, c->stream));
	for (int r = 0; r < reps; r++)
		f();
	HIPCHECK(hipEventRecord(c->ev1, c->stream));
	HIPCHECK(hipEventSynchronize(c->ev1));
	float ms = 0;
	HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
	return ms / reps;
}

/*
 * Tile sizes by measurement on the batch: the RK3 stage with its state
 * staged in LDS, two lanes per element (32 or 128 elements per block) or
 * one (64 or 128), or with 128 or 256 elements per block
 * over the global state (first-stage launches from the
 * current state into the other buffer, before the run starts), the
 * relaxation with 128, 256 or 512 nodes per block (on the other buffer).
 * The candidates are timed in two rounds and each keeps its faster round:
 * the first launches after an idle spell run at lower clocks, which would
 * count against whichever candidate comes first. out = { rk3 elements per
 * block, relax nodes per block }.
 */
extern "C" void gdh_batch_autotune(gdh_ctx *c, int out[2])
{
	const int prof = c->prof;
	c->prof = 0;
	if (!c->scheme) {
		/* { elements per block, kernel }: pair with one or four waves per
		 * tile, tile with one or two, soa */
		static const int cand[6][2] = { { 32, 2 }, { 128, 2 }, { 64, 1 }, { 128, 1 },
						{ 128, 0 }, { 256, 0 } };
		const int nc = 6;
		/* GDN_RK3=pair|tile|soa keeps the candidates of that kernel */
		const char *env = getenv("GDN_RK3");
		const int only = !env ? -1 : !strcmp(env, "soa") ? 0 : !strcmp(env, "tile") ? 1
				: !strcmp(env, "pair") ? 2 : -1;
		auto usable = [&](int k) { return only < 0 || cand[k][1] == only; };
		float ms[6] = { 1e30f, 1e30f, 1e30f, 1e30f, 1e30f, 1e30f };
		for (int round = 0; round < 2; round++)
			for (int k = 0; k < nc; k++) {
				if (!usable(k))
					continue;
				c->rk3_bs = cand[k][0];
				c->rk3_tile = cand[k][1];
				ms[k] = std::min(ms[k], time_launches(c, 5, [&] { rk3_trial(c, 1); }));
			}
		int pick = -1;
		for (int k = 0; k < nc; k++)
			if (usable(k) && (pick < 0 || ms[k] < ms[pick]))
				pick = k;
		c->rk3_bs = cand[pick][0];
		c->rk3_tile = cand[pick][1];
		/* the k2 stream with non-temporal or ordinary loads and stores,
		 * timed on later-stage launches (the ones that read k2), unless
		 * GDN_RK3_K2 fixes it */
		float mt[2] = { 1e30f, 1e30f };
		const int fixed = getenv("GDN_RK3_K2") != nullptr;
		for (int round = 0; round < 2 && !fixed; round++)
			for (int k = 0; k < 2; k++) {
				c->rk3_ntk = 1 - k;
				mt[k] = std::min(mt[k], time_launches(c, 5, [&] { rk3_trial(c, 0); }));
			}
		if (!fixed)
			c->rk3_ntk = mt[0] <= mt[1];
		if (getenv("GDN_VERBOSE") && !fixed)
			fprintf(stderr, "[gdh] batch RK3: %s kernel, %d elements per block; k2 stream "
				"non-temporal %.3f ms, ordinary %.3f ms per stage: %s\n",
				c->rk3_tile == 2 ? "pair" : c->rk3_tile ? "tile" : "soa", c->rk3_bs,
				mt[0], mt[1], c->rk3_ntk ? "non-temporal" : "ordinary");
	} else {
		static const int cand[3] = { 128, 256, 512 };
		float ms[3] = { 1e30f, 1e30f, 1e30f };
		for (int round = 0; round < 2; round++)
			for (int k = 0; k < 3; k++) {
				c->relax_bs = cand[k];
				ms[k] = std::min(ms[k], time_launches(c, 5, [&] {
							 relax_launch(c, c->w[1 - c->cur], 10L * c->ne);
						 }));
			}
		int pick = 1;
		for (int k = 0; k < 3; k++)
			if (ms[k] < ms[pick])
				pick = k;
		c->relax_bs = cand[pick];
	}
	c->prof = prof;
	out[0] = c->rk3_bs;
	out[1] = c->relax_bs;
}

/* one time step of every instance, launches only (capturable) */
static void batch_step(gdh_ctx *c)
{
	static const double C[3] = { 0., 1. / 3., 3. / 4. };
	if (!c->scheme) {
		for (int s = 0; s < 3; s++) {
			ghost_fill_batch(c, C[s]);
			gdh_rk3_stage(c, s, 0.0);
		}
	} else {
		ghost_fill_batch(c, 0.5);
		gdh_relax_step(c);
	}
	k_time_bump<<<nblocks(c->bt.ninst, 64), 64, 0, c->stream>>>(c->bt.time, c->bt.dt,
								    c->bt.ninst);
	HIPCHECK(hipGetLastError());
}

/*
 * nsteps steps of every instance: pairs of steps replay a captured graph
 * (one per parity of the state buffers), an odd last step runs directly.
 */
extern "C" void gdh_batch_advance(gdh_ctx *c, int nsteps)
{
	if (!c->bt.ninst) {
		fprintf(stderr, "gdon3d-hip: not a batch context\n");
		abort();
	}
	if (c->nlagged) {
		fprintf(stderr, "gdon3d-hip: batches of meshes with upwind cycles are "
				"not captured\n");
		abort();
	}
	const int prof = c->prof;
	c->prof = 0; /* no event records inside captures */
	/* kinetic: a captured step relaxes its input (the state awaits the
	 * relaxation of the step before); a state that does not (fresh, or
	 * materialized by a download) takes one direct step first */
	if (c->scheme && c->relax_fused == 1 && !c->relax_pending && nsteps > 0) {
		batch_step(c);
		nsteps--;
	}
	const int pairs = nsteps / 2;
	if (pairs > 0 && !c->graph[c->cur]) {
		const int cur0 = c->cur;
		hipGraph_t g;
		HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
		batch_step(c);
		batch_step(c);
		HIPCHECK(hipStreamEndCapture(c->stream, &g));
		HIPCHECK(hipGraphInstantiate(&c->graph[cur0], g, nullptr, nullptr, 0));
		HIPCHECK(hipGraphDestroy(g));
		if (c->cur != cur0) {
			fprintf(stderr, "gdon3d-hip: captured steps changed the state parity\n");
			abort();
		}
	}
	for (int k = 0; k < pairs; k++)
		HIPCHECK(hipGraphLaunch(c->graph[c->cur], c->stream));
	if (nsteps % 2)
		batch_step(c);
	c->prof = prof;
}

extern "C" void gdh_batch_times(gdh_ctx *c, double *t)
{
	HIPCHECK(hipMemcpyAsync(t, c->bt.time, sizeof(double) * c->bt.ninst,
				hipMemcpyDeviceToHost, c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
}

/* ghosts at every instance's clock (the final update of the kinetic loop) */
extern "C" void gdh_batch_ghosts(gdh_ctx *c)
{
	ghost_fill_batch(c, 0.0);
}

/* L2 error of each instance against its exact solution at its clock */
extern "C" void gdh_batch_error_l2(gdh_ctx *c, double *err)
{
	std::vector<double> t(c->bt.ninst);
	gdh_batch_times(c, t.data());
	for (int i = 0; i < c->bt.ninst; i++) {
		ImposedPar ip = c->ip;
		for (int j = 0; j < GDH_IMPOSED_NPAR; j++)
			ip.p[j] = c->bt_par[(size_t)GDH_IMPOSED_NPAR * i + j];
		err[i] = error_range(c, c->bt_e0[i], c->bt_e0[i + 1], t[i], ip, nullptr);
	}
}

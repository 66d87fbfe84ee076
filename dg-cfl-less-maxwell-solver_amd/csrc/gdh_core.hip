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
 *  - kinetic state item-major (kin_at in ctx.h), the reference layout
 *    f[(iw * nv + iv) * neq + ipg] restored at host transfers.
 */
#include "ctx.h"

extern "C" int gdh_device_count(void)
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess)
		return 0;
	return n;
}

/*
 * Storage order of the RK3 state: elements sorted along a Morton (Z-order)
 * curve of their vertex centroids (21 bits per axis over the bounding box,
 * ties by mesh index). Face neighbours then sit close together in memory,
 * and the blocks that one XCD runs (a contiguous tile range) cover a compact
 * region whose neighbour traces stay in that XCD's L2; in mesh order the
 * neighbours across z-layers are 6 n^2 elements apart and their traces are
 * re-fetched from HBM. Returns order[c] = mesh element stored at column c.
 */
static std::vector<int> elem_order(const gdh_mesh_desc *mesh)
{
	const long ne = mesh->nelem;
	std::vector<double> cen(3 * ne);
	double lo[3] = { 1e300, 1e300, 1e300 }, hi[3] = { -1e300, -1e300, -1e300 };
	for (long e = 0; e < ne; e++)
		for (int d = 0; d < 3; d++) {
			double s = 0.0;
			for (int i = 0; i < 4; i++)
				s += mesh->xnode[30 * e + 3 * i + d];
			cen[3 * e + d] = 0.25 * s;
			lo[d] = std::min(lo[d], cen[3 * e + d]);
			hi[d] = std::max(hi[d], cen[3 * e + d]);
		}
	auto spread = [](uint64_t x) { /* 21 bits -> every third bit */
		x &= 0x1fffff;
		x = (x | x << 32) & 0x1f00000000ffffull;
		x = (x | x << 16) & 0x1f0000ff0000ffull;
		x = (x | x << 8) & 0x100f00f00f00f00full;
		x = (x | x << 4) & 0x10c30c30c30c30c3ull;
		x = (x | x << 2) & 0x1249249249249249ull;
		return x;
	};
	std::vector<std::pair<uint64_t, int>> key(ne);
	for (long e = 0; e < ne; e++) {
		uint64_t code = 0;
		for (int d = 0; d < 3; d++) {
			const double ext = hi[d] - lo[d];
			const double u = ext > 0 ? (cen[3 * e + d] - lo[d]) / ext : 0.0;
			const uint64_t q = (uint64_t)std::min(2097151.0, std::max(0.0, u * 2097151.0));
			code |= spread(q) << d;
		}
		key[e] = { code, (int)e };
	}
	std::sort(key.begin(), key.end());
	std::vector<int> order(ne);
	for (long c = 0; c < ne; c++)
		order[c] = key[c].second;
	return order;
}

/*
 * One stream for every context of the process: the persistent transport
 * sweep needs all its blocks resident at once, which holds only while no
 * other kernel shares the device, so contexts do not run concurrently with
 * each other (each context's kernels fill the device anyway). Created on
 * first use, never destroyed.
 */
static hipStream_t shared_stream()
{
	static hipStream_t s = nullptr;
	if (!s)
		HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
	return s;
}

extern "C" gdh_ctx *gdh_create(const gdh_mesh_desc *mesh,
			       const gdh_model_desc *model, int scheme)
{
	return gdh_create_ex(mesh, model, scheme, GDH_FP64);
}

extern "C" gdh_ctx *gdh_create_ex(const gdh_mesh_desc *mesh,
				  const gdh_model_desc *model, int scheme,
				  int precision)
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
	if (precision != GDH_FP64 && precision != GDH_FP32) {
		fprintf(stderr, "gdon3d-hip: unknown precision %d\n", precision);
		abort();
	}
	c->fp32 = precision == GDH_FP32;
	c->rsz = c->fp32 ? sizeof(float) : sizeof(double);
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
	c->kp.fr_kind = model->fr_kind;
	memcpy(c->kp.proj_pos, model->proj_pos, sizeof(c->kp.proj_pos));
	memcpy(c->kp.proj_neg, model->proj_neg, sizeof(c->kp.proj_neg));
	if (scheme && c->kp.kind == GDH_KIN_NONE) {
		fprintf(stderr, "gdon3d-hip: kinetic scheme without velocity set\n");
		abort();
	}
	c->stream = shared_stream();
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
	HIPCHECK(hipMemcpy(c->xnode, mesh->xnode, sizeof(double) * 30 * ne,
			   hipMemcpyHostToDevice));
	if (nb)
		HIPCHECK(hipMemcpy(c->xghost, mesh->xghost, sizeof(double) * 18 * nb,
				   hipMemcpyHostToDevice));

	const size_t wlen = (size_t)c->m * c->neq;
	/* RK3: interior SoA buffers; kinetic: full item-major buffers */
	const size_t blen = scheme ? wlen : 60 * ne;
	for (int b = 0; b < 2; b++) {
		c->w[b] = dmalloc(c->rsz * blen);
		HIPCHECK(hipMemset(c->w[b], 0, c->rsz * blen));
	}
	c->cur = 0;
	c->k2 = nullptr;
	c->gh = nullptr;
	c->geom_soa = nullptr;
	c->elem_pos = nullptr;
	c->stage = nullptr;
	c->bface = nullptr;
	c->ieps = nullptr;
	c->lb_x = nullptr;
	c->lb_p = nullptr;
	if (scheme) {
		/* the sweep reads neighbours in mesh order */
		HIPCHECK(hipMemcpy(c->nbr, mesh->nbr, sizeof(int4) * ne,
				   hipMemcpyHostToDevice));
		HIPCHECK(hipMemcpy(c->nbmap, mesh->nbmap, sizeof(int4) * ne,
				   hipMemcpyHostToDevice));
		/* element and local face of each boundary face (projector
		 * boundary state) */
		std::vector<int2> bf(nb ? nb : 1, make_int2(0, 0));
		for (size_t e = 0; e < ne; e++)
			for (int f = 0; f < 4; f++) {
				const int x = mesh->nbr[4 * e + f];
				if (x < 0 && -1 - x < (int)nb)
					bf[-1 - x] = make_int2((int)e, f);
			}
		c->bface = (int2 *)dmalloc(sizeof(int2) * bf.size());
		HIPCHECK(hipMemcpy(c->bface, bf.data(), sizeof(int2) * bf.size(),
				   hipMemcpyHostToDevice));
		c->kp.bface = c->bface;
	} else {
		/* RK3 works on storage columns: geometry, neighbours (as columns)
		 * and face maps permuted, ghosts unchanged */
		const std::vector<int> order = elem_order(mesh);
		std::vector<int> pos(ne);
		for (size_t col = 0; col < ne; col++)
			pos[order[col]] = (int)col;
		std::vector<double> gs(10 * ne);
		std::vector<int4> nb4(ne), nm4(ne);
		for (size_t col = 0; col < ne; col++) {
			const size_t e = order[col];
			for (int k = 0; k < 10; k++)
				gs[k * ne + col] = mesh->geom[10 * e + k];
			int nn[4];
			for (int f = 0; f < 4; f++) {
				const int x = mesh->nbr[4 * e + f];
				nn[f] = x >= 0 ? pos[x] : x;
			}
			nb4[col] = make_int4(nn[0], nn[1], nn[2], nn[3]);
			nm4[col] = make_int4(mesh->nbmap[4 * e], mesh->nbmap[4 * e + 1],
					     mesh->nbmap[4 * e + 2], mesh->nbmap[4 * e + 3]);
		}
		HIPCHECK(hipMemcpy(c->nbr, nb4.data(), sizeof(int4) * ne,
				   hipMemcpyHostToDevice));
		HIPCHECK(hipMemcpy(c->nbmap, nm4.data(), sizeof(int4) * ne,
				   hipMemcpyHostToDevice));
		c->elem_pos = (int *)dmalloc(sizeof(int) * ne);
		HIPCHECK(hipMemcpy(c->elem_pos, pos.data(), sizeof(int) * ne,
				   hipMemcpyHostToDevice));
		c->k2 = dmalloc(c->rsz * blen);
		HIPCHECK(hipMemset(c->k2, 0, c->rsz * blen));
		c->gh = dmalloc(c->rsz * 36 * (nb ? nb : 1));
		HIPCHECK(hipMemset(c->gh, 0, c->rsz * 36 * (nb ? nb : 1)));
		c->geom_soa = (double *)dmalloc(sizeof(double) * 10 * ne);
		HIPCHECK(hipMemcpy(c->geom_soa, gs.data(), sizeof(double) * 10 * ne,
				   hipMemcpyHostToDevice));
		c->stage = (double *)dmalloc(sizeof(double) * wlen);
	}
	if (scheme)
		c->stage = (double *)dmalloc(sizeof(double) * wlen);
	c->hostvals = nullptr;
	c->hostvals_len = 0;
	c->npartials = 1024;
	c->partials = (double *)dmalloc(sizeof(double) * 8 * (c->npartials + 1));
	c->items = nullptr;
	c->dinv = nullptr;
	c->cls = nullptr;
	c->desc = nullptr;
	c->ll = nullptr;
	c->wx = nullptr;
	c->wz = nullptr;
	c->epoch_dev = nullptr;
	memset(&c->bt, 0, sizeof(c->bt));
	for (int k = 0; k < 4; k++) {
		c->bt_mem[k] = nullptr;
		c->bt_dmem[k] = nullptr;
	}
	c->graph[0] = c->graph[1] = nullptr;
	{
		/* RK3 stage k2 stream: non-temporal (default) or ordinary loads and
		 * stores; GDN_RK3_K2=nt|plain overrides (batches then skip timing it) */
		const char *k = getenv("GDN_RK3_K2");
		c->rk3_ntk = !(k && !strcmp(k, "plain"));
	}
	{
		/* RK3 stage kernel, all three over the same arithmetic: the tile's
		 * state staged in LDS with two lanes per element ("pair", the binary64
		 * default: 1.17 vs 1.34 ms per stage at 64^3) or one ("tile", the
		 * binary32 default: 0.68 vs 0.71 ms), or one element per thread over
		 * the global state ("soa"); GDN_RK3=pair|tile|soa overrides */
		const char *k = getenv("GDN_RK3");
		c->rk3_tile = c->fp32 ? 1 : 2;
		if (k && !strcmp(k, "soa"))
			c->rk3_tile = 0;
		else if (k && !strcmp(k, "tile"))
			c->rk3_tile = 1;
		else if (k && !strcmp(k, "pair"))
			c->rk3_tile = 2;
		/* elements per block: the pair kernel stages 128 columns per block
		 * of four waves (64^3 binary64 stage 1.072 ms; 1.099 with 64, 1.130
		 * with 32), the tile kernel 128 per block of two (binary32 0.617 ms;
		 * 0.629 with 64, 0.664 with 256: LDS then caps the waves per CU),
		 * the global-state kernel 128 threads; GDN_RK3_BLOCK overrides with
		 * a size the kernel has */
		c->rk3_bs = 128;
		const char *b = getenv("GDN_RK3_BLOCK");
		if (b) {
			const int v = atoi(b);
			const bool ok = c->rk3_tile == 2 ? (v == 32 || v == 64 || v == 128)
					: c->rk3_tile ? (v == 64 || v == 128 || (v == 256 && c->fp32))
						      : (v == 64 || v == 128 || v == 256);
			if (!ok) {
				fprintf(stderr, "gdon3d-hip: GDN_RK3_BLOCK=%s is not a block size of "
						"the %s kernel\n", b,
					c->rk3_tile == 2 ? "pair" : c->rk3_tile ? "tile" : "soa");
				abort();
			}
			c->rk3_bs = v;
		}
	}
	c->relax_bs = 256;
	{
		/* the relaxation pass with two elements per wave and 16-byte
		 * accesses (binary64 default: 32^3 0.122 -> 0.119 ms, 64^3 1.233 ->
		 * 1.210 ms) or one element per wave (binary32 default: 0.087 vs
		 * 0.092 ms with 8-byte pairs); GDN_RELAX_PASS=pair|wave overrides */
		const char *k = getenv("GDN_RELAX_PASS");
		c->relax_pair = k ? !strcmp(k, "pair") : !c->fp32;
	}
	c->cap_qtot = 0;
	c->cap_ipg = c->cap_vel = c->cap_off = c->cap_piv = nullptr;
	c->cap_lu = c->cap_t = c->cap_part = c->cap_rk = nullptr;
	c->cap_luoff = nullptr;
	c->cap_gmres = 0;
	c->queue = nullptr;
	c->qslot = 0;
	c->sweep_queue = 0;
	c->coop_refused = 0;
	c->cap_warned = 0;
	c->nlagged = 0;
	c->lag_last = c->lag_max = 0;
	c->status = nullptr;
	c->nlevels = 0;
	c->nitems = 0;
	c->relax_pending = 0;
	c->relax_fused = 0;
	c->sweep_grid_lf = 0;
	c->ncl = 0;
	c->chead = nullptr;
	c->cdesc = nullptr;
	c->cl_grid = 0;
	c->cl_multi = 0;
	c->cl_levels = 0;
	if (scheme) {
		/* vertices and neighbours for the cluster plan of the sweep */
		c->vtx.resize(12 * ne);
		for (size_t e = 0; e < ne; e++)
			for (int k = 0; k < 12; k++)
				c->vtx[12 * e + k] = (float)mesh->xnode[30 * e + k];
		c->nbr_h.assign(mesh->nbr, mesh->nbr + 4 * ne);
	}
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
			 c->cls,   c->desc,  c->ll, c->status, c->gh, c->geom_soa, c->elem_pos,
			 c->wx, c->wz, c->cap_ipg, c->cap_vel, c->cap_off, c->cap_piv,
			 c->cap_lu, c->cap_t, c->cap_part, c->cap_rk, c->cap_luoff, c->queue, c->epoch_dev, c->chead, c->cdesc,
			 c->bt_mem[0], c->bt_mem[1], c->bt_mem[2], c->bt_mem[3],
			 c->bt_dmem[0], c->bt_dmem[1], c->bt_dmem[2], c->bt_dmem[3],
			 c->stage, c->bface, c->ieps, c->lb_x, c->lb_p };
	for (void *p : ptrs)
		if (p)
			HIPCHECK(hipFree(p));
	for (int k = 0; k < 2; k++)
		if (c->graph[k])
			HIPCHECK(hipGraphExecDestroy(c->graph[k]));
	HIPCHECK(hipEventDestroy(c->ev0));
	HIPCHECK(hipEventDestroy(c->ev1));
	prof_resolve(c);
	for (hipEvent_t e : c->ev_pool)
		HIPCHECK(hipEventDestroy(e));
	delete c;
}

/* RK3 layout conversions between the reference AoS wn[ipg * 6 + q] and
 * the device SoA state */
template <typename T>
__global__ void k_aos_to_soa(const double *__restrict__ a, StateRef<T> s, long n)
{
	const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n)
		return;
	const long ipg = k / 6;
	*macro_ptr(s, ipg, (int)(k % 6)) = (T)a[k];
}

template <typename T>
__global__ void k_soa_to_aos(double *__restrict__ a, StateRef<T> s, long n)
{
	const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n)
		return;
	a[k] = (double)*macro_ptr(s, k / 6, (int)(k % 6));
}

/* kinetic state: reference layout a[(iw nv + v) neq + ipg] (binary64) <->
 * item-major device state (kin_at), T = double or float */
template <typename T>
__global__ void k_ref_to_kin(const double *__restrict__ a, T *__restrict__ b, long ne,
			     long neq, int nv, long n)
{
	const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n)
		return;
	const long row = k / neq, ipg = k - row * neq;
	b[kin_at(ne, neq, nv, ipg, (int)(row / nv), (int)(row % nv))] = (T)a[k];
}

template <typename T>
__global__ void k_kin_to_ref(const T *__restrict__ b, double *__restrict__ a, long ne,
			     long neq, int nv, long n)
{
	const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n)
		return;
	const long row = k / neq, ipg = k - row * neq;
	a[k] = (double)b[kin_at(ne, neq, nv, ipg, (int)(row / nv), (int)(row % nv))];
}

template <typename T> static StateRef<T> state_ref(gdh_ctx *c)
{
	StateRef<T> s;
	s.w = (T *)c->w[c->cur];
	s.gh = (T *)c->gh;
	s.pos = c->elem_pos;
	s.ne = c->ne;
	s.ngh = 6L * c->nb;
	return s;
}

extern "C" void gdh_upload_state(gdh_ctx *c, const double *wn)
{
	const size_t len = (size_t)c->m * c->neq;
	c->relax_pending = 0;
	HIPCHECK(hipMemcpyAsync(c->stage, wn, sizeof(double) * len, hipMemcpyHostToDevice,
				c->stream));
	const int nb = nblocks(len, 256);
	if (c->scheme && c->fp32)
		k_ref_to_kin<float><<<nb, 256, 0, c->stream>>>(
			c->stage, (float *)c->w[c->cur], c->ne, c->neq, c->kp.nv, (long)len);
	else if (c->scheme)
		k_ref_to_kin<double><<<nb, 256, 0, c->stream>>>(
			c->stage, (double *)c->w[c->cur], c->ne, c->neq, c->kp.nv, (long)len);
	else if (c->fp32)
		k_aos_to_soa<float><<<nb, 256, 0, c->stream>>>(c->stage, state_ref<float>(c),
							       (long)len);
	else
		k_aos_to_soa<double><<<nb, 256, 0, c->stream>>>(c->stage, state_ref<double>(c),
								(long)len);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipStreamSynchronize(c->stream));
}

extern "C" void gdh_download_state(gdh_ctx *c, double *wn)
{
	relax_materialize(c);
	const size_t len = (size_t)c->m * c->neq;
	const int nb = nblocks(len, 256);
	if (c->scheme && c->fp32)
		k_kin_to_ref<float><<<nb, 256, 0, c->stream>>>(
			(const float *)c->w[c->cur], c->stage, c->ne, c->neq, c->kp.nv, (long)len);
	else if (c->scheme)
		k_kin_to_ref<double><<<nb, 256, 0, c->stream>>>(
			(const double *)c->w[c->cur], c->stage, c->ne, c->neq, c->kp.nv, (long)len);
	else if (c->fp32)
		k_soa_to_aos<float><<<nb, 256, 0, c->stream>>>(c->stage, state_ref<float>(c),
							       (long)len);
	else
		k_soa_to_aos<double><<<nb, 256, 0, c->stream>>>(c->stage, state_ref<double>(c),
								(long)len);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipMemcpyAsync(wn, c->stage, sizeof(double) * len, hipMemcpyDeviceToHost,
				c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
	check_sweep_status(c);
}

/* after a stream synchronisation: abort if a sweep's bounded wait expired
 * (its results are then not a solution) */
void check_sweep_status(gdh_ctx *c)
{
	if (c->status) {
		int st = 0;
		HIPCHECK(hipMemcpy(&st, c->status, sizeof(int), hipMemcpyDeviceToHost));
		if (st) {
			fprintf(stderr, "gdon3d-hip: the transport sweep timed out "
					"waiting for an upwind dependency\n");
			abort();
		}
	}
}

extern "C" void gdh_synchronize(gdh_ctx *c)
{
	HIPCHECK(hipStreamSynchronize(c->stream));
	check_sweep_status(c);
}

/* 1/eps per mesh element into the context: RK3 by storage column
 * (bt.ieps_col, read by the stage kernels), kinetic in mesh order
 * (kp.ieps) */
void set_permittivity(gdh_ctx *c, const double *eps)
{
	const long ne = c->ne;
	std::vector<double> ie(ne > 0 ? ne : 1, 1.0);
	std::vector<int> pos;
	if (!c->scheme) {
		pos.resize(ne);
		HIPCHECK(hipMemcpy(pos.data(), c->elem_pos, sizeof(int) * ne, hipMemcpyDeviceToHost));
	}
	for (long e = 0; e < ne; e++) {
		if (!(eps[e] > 0.0) || !std::isfinite(eps[e])) {
			fprintf(stderr, "gdon3d-hip: permittivity of element %ld is %g, not positive\n",
				e, eps[e]);
			abort();
		}
		ie[c->scheme ? e : pos[e]] = 1.0 / eps[e];
	}
	if (c->scheme && c->kp.fr_kind == GDH_FR_PROJECTOR) {
		fprintf(stderr, "gdon3d-hip: the projector boundary state takes eps = 1 only\n");
		abort();
	}
	if (!c->ieps)
		c->ieps = (double *)dmalloc(sizeof(double) * ie.size());
	HIPCHECK(hipMemcpy(c->ieps, ie.data(), sizeof(double) * ie.size(), hipMemcpyHostToDevice));
	if (c->scheme)
		c->kp.ieps = c->ieps;
	else
		c->bt.ieps_col = c->ieps;
}

extern "C" void gdh_set_permittivity(gdh_ctx *c, const double *eps)
{
	set_permittivity(c, eps);
}

extern "C" void gdh_rk3_info(gdh_ctx *c, int out[3])
{
	out[0] = c->rk3_tile;
	out[1] = c->rk3_ntk;
	out[2] = c->rk3_bs;
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


/* ===================================================================== */
/*                   initial condition and ghost data                    */
/* ===================================================================== */

/* Write the macro state wv = (E, H) (6 values) of DoF ipg into the state,
 * as w itself (RK3, SoA) or as f_eq of the conserved fields (eps E, H)
 * (kinetic, item-major; the flux of the equilibrium is that of (E, H)). */
template <typename T>
__device__ inline void store_dof(const StateRef<T> &s, T *kin, int scheme,
				 const KinPar &kp, int neq, long ipg,
				 const double wv[6])
{
	if (!scheme) {
#pragma unroll
		for (int q = 0; q < 6; q++)
			*macro_ptr(s, ipg, q) = (T)wv[q];
		return;
	}
	double qm[6][3];
	maxwell_qM(wv, qm);
	double wc[6] = { wv[0], wv[1], wv[2], wv[3], wv[4], wv[5] };
	if (kp.ieps) {
		const double ie = kin_ieps(kp, ipg, s.ne);
		for (int q = 0; q < 3; q++)
			wc[q] = wv[q] / ie;
	}
	for (int iw = 0; iw < 6; iw++)
		for (int v = 0; v < kp.nv; v++)
			kin[kin_at(s.ne, neq, kp.nv, ipg, iw, v)] =
				(T)feq_comp(kp, wc, qm, iw, v);
}

/* t + c dt rounded as the host loops round it (no fused multiply-add) */
__device__ inline double stage_time(double t, double c, double dt)
{
#pragma clang fp contract(off)
	return t + c * dt;
}

/* n DoFs starting at ipg0 with coordinates x[3 k]; exact values from the
 * closed form of ip, or from hv[6 k] when hv != NULL */
template <typename T>
__global__ void k_fill_exact(StateRef<T> s, T *kin, int scheme, KinPar kp,
			     ImposedPar ip, int neq, long ipg0, long n,
			     const double *__restrict__ x,
			     const double *__restrict__ hv, double t, BatchTab bt,
			     double cstage)
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
		if (bt.ninst) {
			/* this DoF's instance: its parameters, clock and step */
			const long ipg = ipg0 + k, ne10 = 10L * s.ne;
			const int in = ipg < ne10 ? bt.einst[ipg / 10] : bt.binst[(ipg - ne10) / 6];
			ImposedPar q = ip;
			for (int j = 0; j < GDH_IMPOSED_NPAR; j++)
				q.p[j] = bt.par[GDH_IMPOSED_NPAR * in + j];
			imposed(q, xx, stage_time(bt.time[in], cstage, bt.dt[in]), wv);
		} else {
			imposed(ip, xx, t, wv);
		}
	}
	store_dof(s, kin, scheme, kp, neq, ipg0 + k, wv);
}

static void fill_launch(gdh_ctx *c, long ipg0, long n, const double *x,
			const double *hv, double t, double cstage)
{
	if (c->fp32)
		k_fill_exact<float><<<nblocks(n, 256), 256, 0, c->stream>>>(
			state_ref<float>(c), (float *)c->w[c->cur], c->scheme, c->kp, c->ip,
			c->neq, ipg0, n, x, hv, t, c->bt, cstage);
	else
		k_fill_exact<double><<<nblocks(n, 256), 256, 0, c->stream>>>(
			state_ref<double>(c), (double *)c->w[c->cur], c->scheme, c->kp,
			c->ip, c->neq, ipg0, n, x, hv, t, c->bt, cstage);
	HIPCHECK(hipGetLastError());
}

extern "C" void gdh_init_solution(gdh_ctx *c, const double *wexact)
{
	const long n = 10L * c->ne;
	if (c->ip.kind == GDH_IMPOSED_HOST && !wexact) {
		fprintf(stderr, "gdon3d-hip: host imposed data missing\n");
		abort();
	}
	const double *hv = stage_hostvals(c, wexact, 6 * n);
	c->relax_pending = 0;
	fill_launch(c, 0, n, c->xnode, hv, 0.0, 0.0);
}

/*
 * Ghost DoFs of the projector boundary state (the reference's get_fR with
 * use_feq = false, src/dg/dg.c:75-100): ghost k = 6 bf + a takes
 * fR = fr_projector(fL, vn, wI) with fL the state at the L element's node
 * of face node a, vn the outward normal of that face and wI the imposed
 * data at the node (closed form, or hv[6 k] from the host). The projectors
 * depend on vn / |vn| only, so the operator's normal codtau n_ref serves.
 */
template <typename T, int NV>
__global__ void k_ghost_projector(T *__restrict__ kin, KinPar kp, ImposedPar ip, long ne,
				  long neq, long n, const int2 *__restrict__ bface,
				  const double *__restrict__ geom, const double *__restrict__ x,
				  const double *__restrict__ hv, double t)
{
	const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n)
		return;
	const int2 ef = bface[k / 6];
	const long ipgL = 10L * ef.x + c_face2node[ef.y][k % 6];
	double fL[6][NV], fR[6][NV], wI[6], ct[9], vn[3];
	for (int iw = 0; iw < 6; iw++)
		for (int v = 0; v < NV; v++)
			fL[iw][v] = (double)kin[kin_at(ne, neq, NV, ipgL, iw, v)];
	for (int j = 0; j < 9; j++)
		ct[j] = geom[10L * ef.x + j];
	face_normal(ct, ef.y, vn);
	if (hv) {
		for (int q = 0; q < 6; q++)
			wI[q] = hv[6 * k + q];
	} else {
		const double xx[3] = { x[3 * k], x[3 * k + 1], x[3 * k + 2] };
		imposed(ip, xx, t, wI);
	}
	fr_projector<NV>(kp, fL, vn, wI, fR);
	for (int iw = 0; iw < 6; iw++)
		for (int v = 0; v < NV; v++)
			kin[kin_at(ne, neq, NV, 10L * ne + k, iw, v)] = (T)fR[iw][v];
}

template <typename T> static void ghost_projector_launch(gdh_ctx *c, const double *hv, double t)
{
	const long n = 6L * c->nb;
	T *kin = (T *)c->w[c->cur];
	if (c->kp.nv == 4)
		k_ghost_projector<T, 4><<<nblocks(n, 128), 128, 0, c->stream>>>(
			kin, c->kp, c->ip, c->ne, c->neq, n, c->bface, c->geom, c->xghost, hv, t);
	else if (c->kp.nv == 6)
		k_ghost_projector<T, 6><<<nblocks(n, 128), 128, 0, c->stream>>>(
			kin, c->kp, c->ip, c->ne, c->neq, n, c->bface, c->geom, c->xghost, hv, t);
	else {
		fprintf(stderr, "gdon3d-hip: no projector boundary state for %d velocities\n",
			c->kp.nv);
		abort();
	}
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
	if (c->scheme && c->kp.fr_kind == GDH_FR_PROJECTOR) {
		/* the projector state reads the element values next to the face,
		 * which must hold the last step's relaxation */
		relax_materialize(c);
		prof_begin(c, 2);
		if (c->fp32)
			ghost_projector_launch<float>(c, hv, t);
		else
			ghost_projector_launch<double>(c, hv, t);
		prof_end(c, 2);
		return;
	}
	prof_begin(c, 2);
	fill_launch(c, 10L * c->ne, n, c->xghost, hv, t, 0.0);
	prof_end(c, 2);
}

/* batches: ghosts at each instance's clock + cstage * its time step */
void ghost_fill_batch(gdh_ctx *c, double cstage)
{
	if (c->scheme && c->kp.fr_kind == GDH_FR_PROJECTOR) {
		fprintf(stderr, "gdon3d-hip: batches take the f_eq boundary state only\n");
		abort();
	}
	const long n = 6L * c->nb;
	if (n)
		fill_launch(c, 10L * c->ne, n, c->xghost, nullptr, 0.0, cstage);
}

/* ===================================================================== */
/*                              L2 error                                 */
/* ===================================================================== */

#define ERR_BS 256

/* per-block partial sums of (w - w_exact)^2 wpg_i det_e per component
 * (src/simulation/simulation.c:624-650) */
template <typename T>
__global__ __launch_bounds__(ERR_BS) void k_err_partial(
	const T *__restrict__ st, StateRef<T> sr, int scheme, KinPar kp,
	ImposedPar ip,
	const double *__restrict__ geom, const double *__restrict__ xnode,
	const double *__restrict__ hv, int e0, int e1, int neq, double t,
	double *__restrict__ partials)
{
	__shared__ double red[6][ERR_BS];
	double acc[6] = { 0, 0, 0, 0, 0, 0 };
	for (int e = e0 + blockIdx.x * ERR_BS + threadIdx.x; e < e1;
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
					wl[q] = (double)*macro_ptr(sr, ipg, q);
			} else {
				for (int q = 0; q < 6; q++) {
					double s = 0.0;
					for (int v = 0; v < kp.nv; v++)
						s += (double)st[kin_at(sr.ne, neq, kp.nv, ipg, q, v)];
					wl[q] = s;
				}
				/* E = D / eps against the exact E */
				if (kp.ieps)
					for (int q = 0; q < 3; q++)
						wl[q] *= kp.ieps[e];
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

/* fixed-order final sum: mean over components of sqrt|sum|; 64 lanes take
 * the partials in a fixed stride, then a fixed-order tree */
__global__ __launch_bounds__(64) void k_err_final(const double *__restrict__ partials,
						  int np, double *__restrict__ out)
{
	__shared__ double red[6][64];
	const int t = threadIdx.x;
	for (int q = 0; q < 6; q++) {
		double s = 0.0;
		for (int b = t; b < np; b += 64)
			s += partials[6 * b + q];
		red[q][t] = s;
	}
	__syncthreads();
	for (int h = 32; h > 0; h >>= 1) {
		if (t < h)
			for (int q = 0; q < 6; q++)
				red[q][t] += red[q][t + h];
		__syncthreads();
	}
	if (t == 0) {
		double err = 0.0;
		for (int q = 0; q < 6; q++)
			err += sqrt(fabs(red[q][0])) / 6;
		out[0] = err;
	}
}

/* L2 error over the elements [e0, e1) against ip at time t (hv: host
 * exact values of those elements' nodes, or null) */
double error_range(gdh_ctx *c, int e0, int e1, double t, const ImposedPar &ip,
		   const double *hv)
{
	relax_materialize(c);
	const int np = std::min(c->npartials, std::max(1, nblocks(e1 - e0, ERR_BS)));
	if (c->fp32)
		k_err_partial<float><<<np, ERR_BS, 0, c->stream>>>(
			(const float *)c->w[c->cur], state_ref<float>(c), c->scheme, c->kp,
			ip, c->geom, c->xnode, hv, e0, e1, c->neq, t, c->partials);
	else
		k_err_partial<double><<<np, ERR_BS, 0, c->stream>>>(
			(const double *)c->w[c->cur], state_ref<double>(c), c->scheme,
			c->kp, ip, c->geom, c->xnode, hv, e0, e1, c->neq, t, c->partials);
	HIPCHECK(hipGetLastError());
	k_err_final<<<1, 64, 0, c->stream>>>(c->partials, np,
					     c->partials + 6L * c->npartials);
	HIPCHECK(hipGetLastError());
	double err = 0;
	HIPCHECK(hipMemcpyAsync(&err, c->partials + 6L * c->npartials,
				sizeof(double), hipMemcpyDeviceToHost, c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
	check_sweep_status(c);
	return err;
}

extern "C" double gdh_error_l2(gdh_ctx *c, double t, const double *wexact)
{
	if (c->ip.kind == GDH_IMPOSED_HOST && !wexact) {
		fprintf(stderr, "gdon3d-hip: host exact values missing\n");
		abort();
	}
	const double *hv = stage_hostvals(c, wexact, 60L * c->ne);
	return error_range(c, 0, c->ne, t, c->ip, hv);
}

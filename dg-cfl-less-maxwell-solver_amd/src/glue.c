/*
 * Host-side preparation of the GPU layer: mesh and model descriptors, the
 * exact upwind classification, host evaluation of unknown imposed-data
 * callbacks, and context creation.
 */
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mesh/t10.h>
#include <mesh/tetrageometry.h>
#include <models/d3q4.h>
#include <models/d3q6.h>
#include <models/maxwell.h>

#include "host_internal.h"

void gdn_fatal(const char *msg)
{
	fprintf(stderr, "gdon3d: %s\n", msg);
	abort();
}

static void *xmalloc(size_t n)
{
	void *p = malloc(n ? n : 1);
	if (!p)
		gdn_fatal("out of memory");
	return p;
}

void gdn_mesh_describe(const gdn_mesh *msh, gdh_mesh_desc *out)
{
	if (!msh->is_built)
		gdn_fatal("the mesh must be built (mesh_build_connectivity)");
	const int ne = msh->nbelems, nb = msh->nboundaryfaces;
	double *geom = xmalloc(sizeof(double) * 10 * (size_t)ne);
	int *nbr = xmalloc(sizeof(int) * 4 * (size_t)ne);
	int *nbmap = xmalloc(sizeof(int) * 4 * (size_t)ne);
	double *xnode = xmalloc(sizeof(double) * 30 * (size_t)ne);
	double *xghost = xmalloc(sizeof(double) * 18 * (size_t)(nb ? nb : 1));

	/* elements on the host's threads: each writes only its own entries */
#pragma omp parallel for schedule(static)
	for (int e = 0; e < ne; e++) {
		gdn_real codtau[3][3], det;
		mesh_get_elem_codtau(msh, e, codtau);
		mesh_get_elem_det(msh, e, &det);
		for (int r = 0; r < 3; r++)
			for (int c = 0; c < 3; c++)
				geom[10 * e + 3 * r + c] = codtau[r][c];
		geom[10 * e + 9] = det;
		for (int i = 0; i < NQN; i++)
			for (int d = 0; d < 3; d++)
				xnode[30 * e + 3 * i + d] =
					msh->node[3 * msh->elem2qnode[NQN * e + i] + d];
		for (int f = 0; f < NFA; f++) {
			const int e2 = msh->elem2elem[NFA * e + f];
			nbr[4 * e + f] = e2;
			int code = 0;
			if (e2 >= 0) {
				for (int a = 0; a < NQV; a++) {
					const int q = msh->elem2qnode
						[NQN * e + t10_get_ref_face2node_entry(f, a)];
					int j = 0;
					while (j < NQN && msh->elem2qnode[NQN * e2 + j] != q)
						j++;
					if (j == NQN)
						gdn_fatal("inconsistent face nodes");
					code |= j << (4 * a);
				}
			}
			nbmap[4 * e + f] = code;
		}
	}
	for (int bf = 0; bf < nb; bf++) {
		const int F = msh->boundaryface[bf];
		const int eL = msh->face2elem[E2F * F], fL = msh->face2elem[E2F * F + 1];
		nbr[4 * eL + fL] = -1 - bf;
		for (int a = 0; a < NQV; a++) {
			const int q = msh->elem2qnode[NQN * ne + NQV * bf + a];
			for (int d = 0; d < 3; d++)
				xghost[18 * bf + 3 * a + d] = msh->node[3 * q + d];
		}
	}
	out->nelem = ne;
	out->nbface = nb;
	out->geom = geom;
	out->nbr = nbr;
	out->nbmap = nbmap;
	out->xnode = xnode;
	out->xghost = xghost;
}

void gdn_mesh_desc_release(gdh_mesh_desc *d)
{
	free((void *)d->geom);
	free((void *)d->nbr);
	free((void *)d->nbmap);
	free((void *)d->xnode);
	free((void *)d->xghost);
	memset(d, 0, sizeof(*d));
}

static int flux_kind(gdn_get_num_flux f)
{
	if (f == maxwell_num_flux_upwind)
		return GDH_FLUX_MAXWELL_UPWIND;
	if (f == (gdn_get_num_flux)maxwell_num_flux_boundary_metal)
		return GDH_FLUX_MAXWELL_METAL;
	return 0;
}

/* the library's Maxwell projectors as P = c[0] I + c[1] A^2 + c[2] A
 * (src/maxwell.c); 0 for any other function */
static int proj_coefs(gdn_get_proj_neg f, double c[3])
{
	static const struct {
		gdn_get_proj_neg f;
		double c[3];
	} known[] = {
		{ maxwell_get_proj_neg_prod_w, { 0.0, 0.5, -0.5 } },
		{ maxwell_get_proj_neg_null_prod_w, { 1.0, -0.5, -0.5 } },
		{ maxwell_get_proj_pos_prod_w, { 0.0, 0.5, 0.5 } },
		{ maxwell_get_proj_pos_null_prod_w, { 1.0, -0.5, 0.5 } },
	};
	for (size_t k = 0; k < sizeof(known) / sizeof(known[0]); k++)
		if (f == known[k].f) {
			memcpy(c, known[k].c, sizeof(known[k].c));
			return 1;
		}
	return 0;
}

void gdn_model_describe(const gdn_model *mdl, int kinetic, gdh_model_desc *out)
{
	memset(out, 0, sizeof(*out));
	if (mdl->nb_w != 6)
		gdn_fatal("the GPU solver implements Maxwell (nb_w = 6) only");
	out->flux_kind = flux_kind(mdl->get_num_flux);
	if (out->flux_kind != GDH_FLUX_MAXWELL_UPWIND)
		gdn_fatal("get_num_flux must be maxwell_num_flux_upwind");
	out->bflux_kind = flux_kind((gdn_get_num_flux)mdl->get_num_flux_boundary);
	if (!out->bflux_kind)
		gdn_fatal("get_num_flux_boundary must be maxwell_num_flux_upwind "
			  "or maxwell_num_flux_boundary_metal");

	const gdn_get_imposed_data g = mdl->get_imposed_data;
	const struct gdn_imposed_params *p = gdn_imposed_params();
	if (!g)
		gdn_fatal("get_imposed_data is not set");
	if (g == maxwell_imposed_macro_cos) {
		out->imposed_kind = GDH_IMPOSED_MACRO_COS;
	} else if (g == maxwell_imposed_macro_cavity) {
		out->imposed_kind = GDH_IMPOSED_CAVITY_TM;
	} else if (g == maxwell_imposed_stationary) {
		out->imposed_kind = GDH_IMPOSED_STATIONARY;
	} else if (g == maxwell_imposed_plane_wave) {
		out->imposed_kind = GDH_IMPOSED_PLANE_WAVE;
		for (int d = 0; d < 3; d++) {
			out->imposed_par[d] = p->pw_k[d];
			out->imposed_par[3 + d] = p->pw_e[d];
			out->imposed_par[6 + d] = p->pw_h[d];
		}
		out->imposed_par[9] = p->pw_omega;
		out->imposed_par[10] = p->pw_phase;
	} else if (g == maxwell_imposed_cavity_mode) {
		out->imposed_kind = GDH_IMPOSED_CAVITY_MODE;
		for (int d = 0; d < 3; d++) {
			out->imposed_par[d] = p->cm_k[d];
			out->imposed_par[3 + d] = p->cm_a[d];
		}
		out->imposed_par[6] = p->cm_omega;
	} else if (g == maxwell_imposed_gaussian_pulse) {
		out->imposed_kind = GDH_IMPOSED_GAUSSIAN_PULSE;
		for (int d = 0; d < 3; d++)
			out->imposed_par[d] = p->gp_c[d];
		out->imposed_par[3] = p->gp_sigma;
		out->imposed_par[4] = p->gp_amp;
	} else {
		out->imposed_kind = GDH_IMPOSED_HOST;
	}

	out->kinetic_kind = GDH_KIN_NONE;
	if (kinetic) {
		if (mdl->get_w == d3q4_get_w && mdl->get_feq == d3q4_get_feq &&
		    (mdl->get_fR == d3q4_get_fR || mdl->get_fR == d3q4_get_fR_projector))
			out->kinetic_kind = GDH_KIN_D3Q4;
		else if (mdl->get_w == d3q6_get_w && mdl->get_feq == d3q6_get_feq &&
			 (mdl->get_fR == d3q6_get_fR || mdl->get_fR == d3q6_get_fR_projector))
			out->kinetic_kind = GDH_KIN_D3Q6;
		else
			gdn_fatal("kinetic callbacks must be the d3q4_* or d3q6_* set");
		out->fr_kind = (mdl->get_fR == d3q4_get_fR_projector ||
				mdl->get_fR == d3q6_get_fR_projector)
				       ? GDH_FR_PROJECTOR
				       : GDH_FR_FEQ;
		if (out->fr_kind == GDH_FR_PROJECTOR) {
			if (!proj_coefs(mdl->get_proj_neg, out->proj_neg) ||
			    (out->kinetic_kind == GDH_KIN_D3Q6 &&
			     !proj_coefs((gdn_get_proj_neg)mdl->get_proj_pos, out->proj_pos)))
				gdn_fatal("the projector boundary state needs get_proj_pos / "
					  "get_proj_neg among the maxwell_get_proj_*_prod_w");
		}
		if (mdl->nb_v > GDH_MAX_V || !mdl->vi)
			gdn_fatal("invalid velocity set");
		out->nb_v = mdl->nb_v;
		memcpy(out->vi, mdl->vi, sizeof(double) * 3 * mdl->nb_v);
		out->lambda = mdl->lambda;
		out->omega = mdl->omega;
	}
}

uint8_t *gdn_upwind_classes(const gdn_mesh *msh, const gdn_model *mdl)
{
	const int ne = msh->nbelems, nv = mdl->nb_v;
	uint8_t *cls = calloc((size_t)ne * NFA * nv, 1);
	if (!cls)
		gdn_fatal("out of memory");
	/* faces on the host's threads: each face writes the classes of its
	 * own two (element, face) slots */
#pragma omp parallel for schedule(static)
	for (int F = 0; F < msh->nbfaces; F++) {
		const int ieL = msh->face2elem[E2F * F + 0];
		const int ifL = msh->face2elem[E2F * F + 1];
		const int ieR = msh->face2elem[E2F * F + 2];
		const int ifR = msh->face2elem[E2F * F + 3];
		/* same computation as src/simulation/simulation.c:511-537 */
		gdn_real physnode[NNO][3], xref[3] = { 0, 0, 0 }, dtau[3][3],
			codtau[3][3], vn[3];
		mesh_get_elem_node(msh, ieL, physnode);
		gdn_tetra_ref2phy(physnode, xref, NULL, ifL, NULL, dtau, codtau,
				  NULL, vn);
		const gdn_real cr = ifL == 3 ? sqrt(3) : 1;
		for (int d = 0; d < 3; d++)
			vn[d] *= cr;
		for (int v = 0; v < nv; v++) {
			const gdn_real *vi = mdl->vi + 3 * v;
			const gdn_real s = vi[0] * vn[0] + vi[1] * vn[1] + vi[2] * vn[2];
			uint8_t cL = 0, cR = 0;
			if (s > 0) {
				cL = 1;
				cR = 2;
			} else if (s < 0) {
				cL = 2;
				cR = 1;
			}
			cls[((size_t)ieL * NFA + ifL) * nv + v] = cL;
			if (ieR >= 0)
				cls[((size_t)ieR * NFA + ifR) * nv + v] = cR;
		}
	}
	return cls;
}

/* Kahn's algorithm over the inflow (class 2) edges of velocity v: lev[e]
 * = longest upwind path; returns the number of elements placed (nelem when
 * the graph is acyclic). */
static int kahn_levels(int nelem, int nv, int v, const int *nbr,
		       const uint8_t *cls, int *indeg, int *queue, int *lev)
{
	for (int e = 0; e < nelem; e++) {
		int d = 0;
		for (int f = 0; f < 4; f++)
			d += cls[((size_t)e * 4 + f) * nv + v] == 2 && nbr[4 * e + f] >= 0;
		indeg[e] = d;
		lev[e] = 0;
	}
	int head = 0, tail = 0;
	for (int e = 0; e < nelem; e++)
		if (!indeg[e])
			queue[tail++] = e;
	while (head < tail) {
		const int e = queue[head++];
		for (int f = 0; f < 4; f++) {
			const int nb = nbr[4 * e + f];
			if (nb < 0 || cls[((size_t)e * 4 + f) * nv + v] != 1)
				continue;
			/* the consumer sees e through an inflow face unless lagged */
			int g = 0;
			while (g < 4 && nbr[4 * nb + g] != e)
				g++;
			if (g == 4 || cls[((size_t)nb * 4 + g) * nv + v] != 2)
				continue;
			if (lev[e] + 1 > lev[nb])
				lev[nb] = lev[e] + 1;
			if (--indeg[nb] == 0)
				queue[tail++] = nb;
		}
	}
	return tail;
}

/* Depth-first search over the elements Kahn could not place (the cycles
 * and what lies downstream of them), following upwind -> downwind edges;
 * every edge into an element on the current DFS path closes a cycle and is
 * lagged: the consumer's face class becomes 3. A graph without its DFS
 * back edges is acyclic. Returns the number of lagged edges. */
static int lag_back_edges(int nelem, int nv, int v, const int *nbr, uint8_t *cls,
			  const int *indeg, int *color, int *stack)
{
	int nlag = 0;
	for (int e = 0; e < nelem; e++)
		color[e] = indeg[e] > 0 ? 0 : 2; /* 0 unvisited, 1 on path, 2 done */
	for (int r = 0; r < nelem; r++) {
		if (color[r])
			continue;
		int top = 0;
		stack[2 * top] = r;
		stack[2 * top + 1] = 0;
		color[r] = 1;
		while (top >= 0) {
			const int e = stack[2 * top];
			const int f = stack[2 * top + 1]++;
			if (f == 4) {
				color[e] = 2;
				top--;
				continue;
			}
			const int nb = nbr[4 * e + f];
			if (nb < 0 || cls[((size_t)e * 4 + f) * nv + v] != 1)
				continue;
			int g = 0;
			while (g < 4 && nbr[4 * nb + g] != e)
				g++;
			if (g == 4 || cls[((size_t)nb * 4 + g) * nv + v] != 2)
				continue;
			if (color[nb] == 1) {
				cls[((size_t)nb * 4 + g) * nv + v] = 3;
				nlag++;
			} else if (color[nb] == 0) {
				color[nb] = 1;
				top++;
				stack[2 * top] = nb;
				stack[2 * top + 1] = 0;
			}
		}
	}
	return nlag;
}

int gdn_sweep_schedule(int nelem, int nv, const int *nbr, uint8_t *cls,
		       int *items, int *level_ptr, int *nlevels)
{
	const size_t n1 = nelem ? nelem : 1;
	int *level = malloc(sizeof(int) * (size_t)nv * n1);
	if (!level)
		gdn_fatal("out of memory");
	int maxlev = -1, nlag = 0;
	/* the velocities on the host's threads: velocity v reads and lags only
	 * the classes of v and writes only its own levels */
#pragma omp parallel for schedule(static, 1) reduction(max : maxlev) reduction(+ : nlag)
	for (int v = 0; v < nv; v++) {
		int *indeg = malloc(sizeof(int) * n1);
		int *queue = malloc(sizeof(int) * n1);
		int *color = malloc(sizeof(int) * n1);
		int *stack = malloc(sizeof(int) * 2 * n1);
		if (!indeg || !queue || !color || !stack)
			gdn_fatal("out of memory");
		int *lev = level + (size_t)v * nelem;
		if (kahn_levels(nelem, nv, v, nbr, cls, indeg, queue, lev) != nelem) {
			nlag += lag_back_edges(nelem, nv, v, nbr, cls, indeg, color, stack);
			if (kahn_levels(nelem, nv, v, nbr, cls, indeg, queue, lev) != nelem)
				gdn_fatal("upwind graph still cyclic after lagging back edges");
		}
		for (int e = 0; e < nelem; e++)
			if (lev[e] > maxlev)
				maxlev = lev[e];
		free(indeg);
		free(queue);
		free(color);
		free(stack);
	}
	/* counting sort by level; (v, e) increasing inside a level */
	const int nl = maxlev >= 0 ? maxlev + 1 : 0;
	/* The items' order: each velocity's levels stretched over all nl
	 * levels, level L of a velocity with nl_v levels to the levels
	 * [a, b) = [floor(L nl / nl_v), floor((L + 1) nl / nl_v)), its items
	 * dealt over them by element index (every item of L before every item
	 * of L + 1, so still a topological order of each velocity's graph; the
	 * velocities do not depend on each other). A velocity with a shorter
	 * chain then shares every level evenly with the others instead of
	 * crowding the first ones: on the Kuhn 32^3 cube one D3Q4 velocity has
	 * 94 levels, the others 220 (persistent sweep, relaxation fused /
	 * separate: 0.823 / 0.715 ms unstretched, 0.806 / 0.716 stretched
	 * without dealing, 0.795 / 0.700 ms dealt; latest levels or the
	 * midpoint of earliest and latest measured no better).
	 * GDN_SWEEP_ORDER=level keeps the earliest levels. */
	{
		const char *o = getenv("GDN_SWEEP_ORDER");
		if (!(o && !strcmp(o, "level")) && nl > 0) {
			for (int v = 0; v < nv; v++) {
				int *lev = level + (size_t)v * nelem;
				int lv = 0;
				for (int e = 0; e < nelem; e++)
					if (lev[e] + 1 > lv)
						lv = lev[e] + 1;
				if (lv <= 0 || lv == nl)
					continue;
				for (int e = 0; e < nelem; e++) {
					const long L = lev[e];
					const int a = (int)(L * nl / lv), b = (int)((L + 1) * nl / lv);
					lev[e] = a + e % (b - a);
				}
			}
		}
	}
	for (int L = 0; L <= nl; L++)
		level_ptr[L] = 0;
	for (int v = 0; v < nv; v++)
		for (int e = 0; e < nelem; e++)
			level_ptr[level[(size_t)v * nelem + e] + 1]++;
	for (int L = 0; L < nl; L++)
		level_ptr[L + 1] += level_ptr[L];
	int *pos = malloc(sizeof(int) * (size_t)(nl > 0 ? nl : 1));
	if (!pos)
		gdn_fatal("out of memory");
	for (int L = 0; L < nl; L++)
		pos[L] = level_ptr[L];
	for (int v = 0; v < nv; v++)
		for (int e = 0; e < nelem; e++)
			items[pos[level[(size_t)v * nelem + e]]++] = v * nelem + e;
	free(pos);
	*nlevels = nl;
	free(level);
	return nlag;
}

double *gdn_eval_imposed_elements(const gdn_simulation *simu, double t)
{
	const gdn_mesh *msh = simu->tt;
	const int nw = simu->mdl->nb_w;
	double *w = xmalloc(sizeof(double) * (size_t)NQN * msh->nbelems * nw);
	for (int ipg = 0; ipg < NQN * msh->nbelems; ipg++)
		simu->mdl->get_imposed_data(msh->node + 3 * msh->elem2qnode[ipg], t,
					    w + (size_t)ipg * nw);
	return w;
}

double *gdn_eval_imposed_ghosts(const gdn_simulation *simu, double t)
{
	const gdn_mesh *msh = simu->tt;
	const int nw = simu->mdl->nb_w, nb = msh->nboundaryfaces;
	double *w = xmalloc(sizeof(double) * (size_t)NQV * (nb ? nb : 1) * nw);
	for (int k = 0; k < NQV * nb; k++) {
		const int q = msh->elem2qnode[NQN * msh->nbelems + k];
		simu->mdl->get_imposed_data(msh->node + 3 * q, t,
					    w + (size_t)k * nw);
	}
	return w;
}

/*
 * One device context kept for the standalone entry points
 * (dg_init_sol_macro, dg_update_boundary, dg_apply_relaxation), which a
 * reference-style driver may call every stage: the context of the last
 * simulation they served, reused while the simulation, its mesh, state
 * length, scheme and model (including the exact-solution parameters) are
 * the same; each call uploads simu->wn and downloads it again.
 * simulation_free drops it.
 */
static struct {
	const gdn_simulation *simu;
	const gdn_mesh *mesh;
	/* the mesh's sizes and arrays: a struct reused for another mesh
	 * without simulation_free does not match */
	int nbelems, nbnodes, nboundaryfaces;
	const void *node, *elem2node, *face2elem;
	int wlen, kinetic;
	gdh_model_desc md;
	gdh_ctx *ctx;
} g_ctx_cache;

static int ctx_cache_matches(const gdn_simulation *simu, const gdh_model_desc *md)
{
	const gdn_mesh *m = simu->tt;
	return g_ctx_cache.ctx && g_ctx_cache.simu == simu && g_ctx_cache.mesh == m &&
	       g_ctx_cache.nbelems == m->nbelems && g_ctx_cache.nbnodes == m->nbnodes &&
	       g_ctx_cache.nboundaryfaces == m->nboundaryfaces &&
	       g_ctx_cache.node == (const void *)m->node &&
	       g_ctx_cache.elem2node == (const void *)m->elem2node &&
	       g_ctx_cache.face2elem == (const void *)m->face2elem &&
	       g_ctx_cache.wlen == simu->wlen &&
	       g_ctx_cache.kinetic == (simu->use_kinetic_scheme ? 1 : 0) &&
	       !memcmp(&g_ctx_cache.md, md, sizeof(*md));
}

gdh_ctx *gdn_ctx_cached(gdn_simulation *simu, gdh_model_desc *md)
{
	if (gdh_device_count() < 1)
		gdn_fatal("no HIP device is visible; the GPU solver cannot run");
	gdn_model_describe(simu->mdl, simu->use_kinetic_scheme, md);
	if (ctx_cache_matches(simu, md)) {
		gdh_upload_state(g_ctx_cache.ctx, simu->wn);
		return g_ctx_cache.ctx;
	}
	gdn_ctx_cache_drop(NULL);
	gdh_model_desc m2;
	g_ctx_cache.ctx = gdn_ctx_open(simu, &m2);
	g_ctx_cache.simu = simu;
	g_ctx_cache.mesh = simu->tt;
	g_ctx_cache.nbelems = simu->tt->nbelems;
	g_ctx_cache.nbnodes = simu->tt->nbnodes;
	g_ctx_cache.nboundaryfaces = simu->tt->nboundaryfaces;
	g_ctx_cache.node = simu->tt->node;
	g_ctx_cache.elem2node = simu->tt->elem2node;
	g_ctx_cache.face2elem = simu->tt->face2elem;
	g_ctx_cache.wlen = simu->wlen;
	g_ctx_cache.kinetic = simu->use_kinetic_scheme ? 1 : 0;
	g_ctx_cache.md = *md;
	return g_ctx_cache.ctx;
}

void gdn_ctx_cache_drop(const gdn_simulation *simu)
{
	if (g_ctx_cache.ctx && (!simu || g_ctx_cache.simu == simu)) {
		gdh_destroy(g_ctx_cache.ctx);
		memset(&g_ctx_cache, 0, sizeof(g_ctx_cache));
	}
}

gdh_ctx *gdn_ctx_open(gdn_simulation *simu, gdh_model_desc *md)
{
	return gdn_ctx_open_ex(simu, md, GDH_FP64);
}

gdh_ctx *gdn_ctx_open_ex(gdn_simulation *simu, gdh_model_desc *md, int precision)
{
	/* GDN_VERBOSE: the wall time of each phase */
	const int verbose = getenv("GDN_VERBOSE") != NULL;
	double t0 = omp_get_wtime();
#define OPEN_MARK(what)                                                                  \
	do {                                                                             \
		if (verbose) {                                                           \
			const double t1 = omp_get_wtime();                               \
			fprintf(stderr, "gdon3d-hip: device context, %s: %.3f s\n", what, \
				t1 - t0);                                                \
			t0 = t1;                                                         \
		}                                                                        \
	} while (0)
	if (gdh_device_count() < 1)
		gdn_fatal("no HIP device is visible; the GPU solver cannot run");
	OPEN_MARK("device query");
	gdn_model_describe(simu->mdl, simu->use_kinetic_scheme, md);
	gdh_mesh_desc d;
	gdn_mesh_describe(simu->tt, &d);
	OPEN_MARK("mesh arrays");
	gdh_ctx *ctx = gdh_create_ex(&d, md, simu->use_kinetic_scheme ? 1 : 0, precision);
	gdn_mesh_desc_release(&d);
	OPEN_MARK("device buffers and uploads");
	gdh_upload_state(ctx, simu->wn);
	OPEN_MARK("state upload");
#undef OPEN_MARK
	return ctx;
}

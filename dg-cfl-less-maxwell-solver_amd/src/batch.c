/*
 * Batched solves (include/dg/batch.h): the instances' meshes are joined
 * into one disjoint union (elements, then boundary faces, instance after
 * instance), uploaded as one device context with per-instance tables, and
 * advanced with captured-graph steps.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <dg/batch.h>
#include <dg/stepper.h>

#include "host_internal.h"

struct gdn_batch {
	int count, kinetic, m;
	gdn_simulation **simus;
	gdh_ctx *ctx;
	int *e0, *b0; /* first element / boundary face of each instance */
	int tiles[2];
	int iter;
};

static void *xcalloc(size_t n, size_t s)
{
	void *p = calloc(n ? n : 1, s);
	if (!p)
		gdn_fatal("out of memory");
	return p;
}

gdn_batch *dg_batch_create(gdn_simulation **simus, int count, const double *par,
			   const double *eps, int precision)
{
	if (count < 1)
		gdn_fatal("empty batch");
	gdn_batch *b = xcalloc(1, sizeof(*b));
	b->count = count;
	b->simus = xcalloc(count, sizeof(*simus));
	memcpy(b->simus, simus, sizeof(*simus) * count);
	b->kinetic = simus[0]->use_kinetic_scheme ? 1 : 0;
	b->e0 = xcalloc(count + 1, sizeof(int));
	b->b0 = xcalloc(count + 1, sizeof(int));
	gdh_model_desc md;
	gdn_model_describe(simus[0]->mdl, b->kinetic, &md);
	if (md.imposed_kind != GDH_IMPOSED_PLANE_WAVE &&
	    md.imposed_kind != GDH_IMPOSED_CAVITY_MODE)
		gdn_fatal("batches use maxwell_imposed_plane_wave or "
			  "maxwell_imposed_cavity_mode");
	gdh_mesh_desc *d = xcalloc(count, sizeof(*d));
	for (int i = 0; i < count; i++) {
		gdh_model_desc mi;
		if ((simus[i]->use_kinetic_scheme ? 1 : 0) != b->kinetic)
			gdn_fatal("a batch mixes RK3 and kinetic simulations");
		gdn_model_describe(simus[i]->mdl, b->kinetic, &mi);
		if (mi.imposed_kind != md.imposed_kind || mi.kinetic_kind != md.kinetic_kind ||
		    mi.flux_kind != md.flux_kind || mi.bflux_kind != md.bflux_kind)
			gdn_fatal("the simulations of a batch need the same model");
		gdn_mesh_describe(simus[i]->tt, &d[i]);
		b->e0[i + 1] = b->e0[i] + d[i].nelem;
		b->b0[i + 1] = b->b0[i] + d[i].nbface;
	}
	const int ne = b->e0[count], nb = b->b0[count];
	double *geom = xcalloc((size_t)10 * ne, sizeof(double));
	double *xnode = xcalloc((size_t)30 * ne, sizeof(double));
	double *xghost = xcalloc((size_t)18 * nb, sizeof(double));
	int *nbr = xcalloc((size_t)4 * ne, sizeof(int));
	int *nbmap = xcalloc((size_t)4 * ne, sizeof(int));
	for (int i = 0; i < count; i++) {
		const int eo = b->e0[i], bo = b->b0[i], n = d[i].nelem;
		memcpy(geom + 10L * eo, d[i].geom, sizeof(double) * 10 * n);
		memcpy(xnode + 30L * eo, d[i].xnode, sizeof(double) * 30 * n);
		memcpy(xghost + 18L * bo, d[i].xghost, sizeof(double) * 18 * d[i].nbface);
		memcpy(nbmap + 4L * eo, d[i].nbmap, sizeof(int) * 4 * n);
		for (long k = 0; k < 4L * n; k++) {
			const int x = d[i].nbr[k];
			nbr[4L * eo + k] = x >= 0 ? x + eo : x - bo;
		}
	}
	gdh_mesh_desc u = { .nelem = ne, .nbface = nb, .geom = geom, .nbr = nbr,
			    .nbmap = nbmap, .xnode = xnode, .xghost = xghost };
	b->ctx = gdh_create_ex(&u, &md, b->kinetic,
			       precision == GDN_PRECISION_FP32 ? GDH_FP32 : GDH_FP64);
	double *dt = xcalloc(count, sizeof(double));
	for (int i = 0; i < count; i++)
		dt[i] = simus[i]->dt;
	gdh_batch_setup(b->ctx, count, b->e0, b->b0, par, dt, eps);
	if (b->kinetic) {
		const int nv = md.nb_v;
		uint8_t *cls = xcalloc((size_t)ne * 4 * nv, 1);
		for (int i = 0; i < count; i++) {
			uint8_t *ci = gdn_upwind_classes(simus[i]->tt, simus[i]->mdl);
			memcpy(cls + (size_t)b->e0[i] * 4 * nv, ci, (size_t)d[i].nelem * 4 * nv);
			free(ci);
		}
		int *items = xcalloc((size_t)nv * ne, sizeof(int));
		int *lptr = xcalloc((size_t)nv * ne + 1, sizeof(int));
		int nl = 0;
		gdn_sweep_schedule(ne, nv, nbr, cls, items, lptr, &nl);
		gdh_relax_setup(b->ctx, dt[0], cls, items, lptr, nl);
		free(items);
		free(lptr);
		free(cls);
	}
	gdh_batch_autotune(b->ctx, b->tiles);
	gdh_init_solution(b->ctx, NULL);
	gdh_synchronize(b->ctx);
	for (int i = 0; i < count; i++)
		gdn_mesh_desc_release(&d[i]);
	free(d);
	free(dt);
	free(geom);
	free(xnode);
	free(xghost);
	free(nbr);
	free(nbmap);
	b->m = b->kinetic ? md.nb_v * 6 : 6;
	return b;
}

void dg_batch_tiles(const gdn_batch *b, int out[2])
{
	out[0] = b->tiles[0];
	out[1] = b->tiles[1];
}

void dg_batch_advance(gdn_batch *b, int nsteps)
{
	gdh_batch_advance(b->ctx, nsteps);
	b->iter += nsteps;
}

void dg_batch_synchronize(gdn_batch *b)
{
	gdh_synchronize(b->ctx);
}

int dg_batch_iterations(const gdn_batch *b)
{
	return b->iter;
}

void dg_batch_errors(gdn_batch *b, double *err)
{
	gdh_batch_error_l2(b->ctx, err);
}

void dg_batch_times(gdn_batch *b, double *t)
{
	gdh_batch_times(b->ctx, t);
}

/* union state -> each simu->wn, in the reference layouts */
void dg_batch_finish(gdn_batch *b)
{
	gdh_ctx *c = b->ctx;
	gdh_synchronize(c); /* aborts if a sweep's bounded wait expired */
	if (b->kinetic)
		gdh_batch_ghosts(c); /* final ghost update at t (src/dg/relaxation.c:175) */
	const int ne = b->e0[b->count], nb = b->b0[b->count];
	const long neq = 10L * ne + 6L * nb;
	double *all = xcalloc((size_t)b->m * neq, sizeof(double));
	gdh_download_state(c, all);
	double *t = xcalloc(b->count, sizeof(double));
	gdh_batch_times(c, t);
	for (int i = 0; i < b->count; i++) {
		gdn_simulation *s = b->simus[i];
		const long nei = 10L * (b->e0[i + 1] - b->e0[i]);
		const long nbi = 6L * (b->b0[i + 1] - b->b0[i]);
		const long neqi = nei + nbi;
		const long ei = 10L * b->e0[i], gi = 10L * ne + 6L * b->b0[i];
		if (!b->kinetic) {
			memcpy(s->wn, all + 6 * ei, sizeof(double) * 6 * nei);
			memcpy(s->wn + 6 * nei, all + 6 * gi, sizeof(double) * 6 * nbi);
		} else {
			for (int p = 0; p < b->m; p++) {
				memcpy(s->wn + p * neqi, all + p * neq + ei, sizeof(double) * nei);
				memcpy(s->wn + p * neqi + nei, all + p * neq + gi,
				       sizeof(double) * nbi);
			}
		}
		s->t = t[i];
	}
	free(all);
	free(t);
}

void dg_batch_free(gdn_batch *b)
{
	if (!b)
		return;
	gdh_destroy(b->ctx);
	free(b->e0);
	free(b->b0);
	free(b->simus);
	free(b);
}

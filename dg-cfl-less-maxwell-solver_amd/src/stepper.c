/*
 * Stepper over the GPU layer: the time loops of src/dg/explicit.c:39-117
 * (RK3: per stage ghosts at t + C_i dt, fused stage) and
 * src/dg/relaxation.c:129-185 (per step ghosts at t + dt/2, transport,
 * relaxation; final ghosts at t), split into create / init / advance /
 * finish so that drivers can time or inspect single steps.
 */
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>

#include <dg/stepper.h>

#include "host_internal.h"

struct gdn_stepper {
	gdn_simulation *simu;
	gdh_ctx *ctx;
	gdh_model_desc md;
	double t;
	int iter;
	int levels;
	int lagged; /* upwind edges lagged to break cycles */
};

static void ghosts(gdn_stepper *st, double t)
{
	double *wg = NULL;
	if (st->md.imposed_kind == GDH_IMPOSED_HOST)
		wg = gdn_eval_imposed_ghosts(st->simu, t);
	gdh_update_boundary(st->ctx, t, wg);
	free(wg);
}

gdn_stepper *dg_stepper_create(gdn_simulation *simu)
{
	return dg_stepper_create_precision(simu, GDN_PRECISION_FP64);
}

gdn_stepper *dg_stepper_create_precision(gdn_simulation *simu, int precision)
{
	gdn_stepper *st = calloc(1, sizeof(*st));
	if (!st)
		gdn_fatal("out of memory");
	st->simu = simu;
	/* GDN_VERBOSE: the wall time of each phase */
	const int verbose = getenv("GDN_VERBOSE") != NULL;
	double t0 = omp_get_wtime();
#define SETUP_MARK(what)                                                          \
	do {                                                                      \
		if (verbose) {                                                    \
			const double t1 = omp_get_wtime();                        \
			fprintf(stderr, "gdon3d-hip: set-up, %s: %.3f s\n", what, \
				t1 - t0);                                         \
			t0 = t1;                                                  \
		}                                                                 \
	} while (0)
	st->ctx = gdn_ctx_open_ex(simu, &st->md,
				  precision == GDN_PRECISION_FP32 ? GDH_FP32 : GDH_FP64);
	SETUP_MARK("device context, mesh and geometry");
	if (simu->use_kinetic_scheme) {
		const gdn_mesh *msh = simu->tt;
		const int ne = msh->nbelems, nv = simu->mdl->nb_v;
		uint8_t *cls = gdn_upwind_classes(msh, simu->mdl);
		SETUP_MARK("upwind classes");
		gdh_mesh_desc d;
		gdn_mesh_describe(msh, &d);
		int *items = malloc(sizeof(int) * (size_t)nv * ne);
		int *lptr = malloc(sizeof(int) * ((size_t)nv * ne + 1));
		if (!items || !lptr)
			gdn_fatal("out of memory");
		int nl = 0;
		st->lagged = gdn_sweep_schedule(ne, nv, d.nbr, cls, items, lptr, &nl);
		gdn_mesh_desc_release(&d);
		st->levels = nl;
		SETUP_MARK("sweep schedule");
		gdh_relax_setup(st->ctx, simu->dt, cls, items, lptr, nl);
		free(items);
		free(lptr);
		free(cls);
	}
	gdh_synchronize(st->ctx);
	SETUP_MARK("relaxation set-up");
#undef SETUP_MARK
	return st;
}

void dg_stepper_set_permittivity(gdn_stepper *st, const gdn_real *eps)
{
	gdh_set_permittivity(st->ctx, eps);
}

void dg_stepper_init(gdn_stepper *st)
{
	double *wex = NULL;
	if (st->md.imposed_kind == GDH_IMPOSED_HOST)
		wex = gdn_eval_imposed_elements(st->simu, 0.0);
	gdh_init_solution(st->ctx, wex);
	free(wex);
	st->t = 0;
	st->iter = 0;
	st->simu->t = 0;
}

void dg_stepper_advance(gdn_stepper *st, int nsteps)
{
	static const double C[3] = { 0., 1. / 3., 3. / 4. };
	const double dt = st->simu->dt;
	for (int k = 0; k < nsteps; k++) {
		if (st->simu->use_kinetic_scheme) {
			ghosts(st, st->t + dt / 2);
			gdh_relax_step(st->ctx);
		} else {
			for (int i = 0; i < 3; i++) {
				ghosts(st, st->t + C[i] * dt);
				gdh_rk3_stage(st->ctx, i, dt);
			}
		}
		st->t += dt;
		st->iter++;
	}
	st->simu->t = st->t;
}

gdn_real dg_stepper_time(const gdn_stepper *st)
{
	return st->t;
}

int dg_stepper_iterations(const gdn_stepper *st)
{
	return st->iter;
}

int dg_stepper_levels(const gdn_stepper *st)
{
	return st->levels;
}

void dg_stepper_synchronize(gdn_stepper *st)
{
	gdh_synchronize(st->ctx);
}

gdn_real dg_stepper_error_l2(gdn_stepper *st)
{
	double *wex = NULL;
	if (st->md.imposed_kind == GDH_IMPOSED_HOST)
		wex = gdn_eval_imposed_elements(st->simu, st->t);
	const double e = gdh_error_l2(st->ctx, st->t, wex);
	free(wex);
	return e;
}

gdn_real dg_stepper_energy(gdn_stepper *st)
{
	return gdh_energy(st->ctx);
}

void dg_stepper_finish(gdn_stepper *st)
{
	if (st->simu->use_kinetic_scheme)
		ghosts(st, st->t);
	gdh_synchronize(st->ctx);
	gdh_download_state(st->ctx, st->simu->wn);
	st->simu->t = st->t;
}

void dg_stepper_profile(gdn_stepper *st, int enable, double out[4])
{
	gdh_profile(st->ctx, enable, out);
}

void dg_stepper_free(gdn_stepper *st)
{
	if (!st)
		return;
	gdh_destroy(st->ctx);
	free(st);
}

void dg_stepper_transport_stats(const gdn_stepper *st, int out[3])
{
	if (st->simu->use_kinetic_scheme) {
		gdh_transport_stats(st->ctx, out);
	} else {
		out[0] = out[1] = out[2] = 0;
	}
}

void *dg_stepper_device_context(gdn_stepper *st)
{
	return st->ctx;
}

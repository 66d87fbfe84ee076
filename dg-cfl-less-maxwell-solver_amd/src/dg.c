/*
 * dg_* entry points on the GPU (src/dg/dg.c:23-147). The standalone ones
 * (initial data, boundary update, relaxation) run on a device context kept
 * for the simulation between calls: each call uploads wn, runs the kernel
 * and copies wn back.
 */
#include <stdlib.h>

#include <dg/dg.h>
#include <dg/explicit.h>
#include <dg/relaxation.h>

#include "host_internal.h"

void dg_solve(gdn_simulation *simu)
{
	if (simu->use_kinetic_scheme)
		dg_solve_relaxation(simu);
	else
		dg_solve_rk3(simu);
}

void dg_init_sol_macro(gdn_simulation *simu)
{
	gdh_model_desc md;
	gdh_ctx *ctx = gdn_ctx_cached(simu, &md);
	double *wex = NULL;
	if (md.imposed_kind == GDH_IMPOSED_HOST)
		wex = gdn_eval_imposed_elements(simu, 0.0);
	gdh_init_solution(ctx, wex);
	free(wex);
	gdh_download_state(ctx, simu->wn);
}

void dg_update_boundary(gdn_simulation *simu, const gdn_real t)
{
	gdh_model_desc md;
	gdh_ctx *ctx = gdn_ctx_cached(simu, &md);
	double *wg = NULL;
	if (md.imposed_kind == GDH_IMPOSED_HOST)
		wg = gdn_eval_imposed_ghosts(simu, t);
	gdh_update_boundary(ctx, t, wg);
	free(wg);
	gdh_download_state(ctx, simu->wn);
}

void dg_apply_relaxation(gdn_simulation *simu)
{
	if (!simu->use_kinetic_scheme)
		gdn_fatal("dg_apply_relaxation needs the kinetic scheme");
	gdh_model_desc md;
	gdh_ctx *ctx = gdn_ctx_cached(simu, &md);
	gdh_apply_relaxation(ctx);
	gdh_download_state(ctx, simu->wn);
}

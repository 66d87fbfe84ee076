/*
 * Simulation state and the API around the solvers. Restates
 * src/simulation/simulation.c:22-214 (state, layouts, time parameters) and
 * :606-692 (L2 error, info table); the L2 reduction runs on the GPU.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <mesh/t10.h>
#include <simulation/simulation.h>

#include "host_internal.h"

/* kinetic layout: micro-field im = iw * nb_v + iv of DoF ipg at
 * wn[im * neq + ipg] (src/simulation/simulation.c:22-54) */
void simulation_get_f(gdn_simulation *simu, const int id_pg, gdn_real *f)
{
	for (int im = 0; im < simu->m; im++)
		f[im] = simu->wn[(size_t)im * simu->neq + id_pg];
}

void simulation_set_f(gdn_simulation *simu, const int id_pg, const gdn_real *f)
{
	for (int im = 0; im < simu->m; im++)
		simu->wn[(size_t)im * simu->neq + id_pg] = f[im];
}

void simulation_relax_f(gdn_simulation *simu, const int id_pg,
			const gdn_real *f_old, const gdn_real *f_new)
{
	const gdn_real om = simu->mdl->omega;
	for (int im = 0; im < simu->m; im++)
		simu->wn[(size_t)im * simu->neq + id_pg] =
			om * f_new[im] + (1. - om) * f_old[im];
}

int simulation_get_varindex(int m, int ipg, int iv)
{
	return iv + m * ipg;
}

/* src/simulation/simulation.c:144-164: dt from the CFL number unless
 * cfl == -1, then tmax truncated to a whole number of steps */
void simulation_set_time_parameters(gdn_simulation *simu, const gdn_real tmax,
				    const gdn_real dt, const gdn_real cfl)
{
	const gdn_model *mdl = simu->mdl;
	simu->tmax = tmax;
	if (cfl != -1) {
		simu->cfl = cfl;
		simu->dt = (simu->hmin * cfl) / mdl->lambda;
	} else {
		simu->dt = dt;
		simu->cfl = (mdl->lambda * simu->dt) / simu->hmin;
	}
	simu->itermax = (int)floor(tmax / simu->dt);
	simu->tmax = simu->itermax * simu->dt;
}

/* src/simulation/simulation.c:167-191 */
void simulation_init(gdn_mesh *msh, gdn_model *mdl, gdn_simulation *f)
{
	if (!msh || !mdl)
		gdn_fatal("simulation_init: mesh and model are required");
	f->tt = msh;
	f->mdl = mdl;
	f->use_kinetic_scheme = mdl->use_relax_scheme;
	f->m = f->use_kinetic_scheme ? mdl->nb_v * mdl->nb_w : mdl->nb_w;
	f->neq = msh->nbelems * NQN + msh->nboundaryfaces * NQV;
	f->wlen = f->m * f->neq;
	f->wn = calloc((size_t)f->wlen, sizeof(gdn_real));
	if (!f->wn)
		gdn_fatal("simulation_init: out of memory");
	f->t = 0;
	f->hmin = mesh_get_t4_hmin(msh);
	f->cpu_running_t = 0;
	f->cpu_assembly_t = 0;
}

/* src/simulation/simulation.c:193-209: frees wn and the mesh arrays */
void simulation_free(gdn_simulation *f)
{
	gdn_ctx_cache_drop(f); /* the standalone entry points' device context */
	f->m = 0;
	f->wlen = 0;
	f->neq = 0;
	f->t = 0;
	f->hmin = 0;
	if (f->tt)
		gdn_gdn_mesh_free(f->tt);
	f->mdl = NULL;
	free(f->wn);
	f->wn = NULL;
}

void simulation_display_info(gdn_simulation *simu)
{
	printf("[simulation] elements %d, DOF %d, itermax %d, tmax %.12f, "
	       "cfl %.12f, dt %.12f, hmin %.12f\n",
	       simu->tt->nbelems, simu->m * simu->tt->nbelems * NQN,
	       simu->itermax, simu->tmax, simu->cfl, simu->dt, simu->hmin);
}

/* src/simulation/simulation.c:606-666: mean over the nb_w components of
 * sqrt|sum_e sum_i (w_i - w_exact(x_i, t))^2 wpg_i det_e|, FP64 on the GPU */
gdn_real simulation_error_l2(gdn_simulation *simu)
{
	gdh_model_desc md;
	gdh_ctx *ctx = gdn_ctx_open(simu, &md);
	double *wex = NULL;
	if (md.imposed_kind == GDH_IMPOSED_HOST)
		wex = gdn_eval_imposed_elements(simu, simu->t);
	const double err = gdh_error_l2(ctx, simu->t, wex);
	free(wex);
	gdh_destroy(ctx);
	return err;
}

/* src/simulation/simulation.c:668-692 */
void simulation_dump_info(const int num_run, gdn_simulation *simu,
			  const gdn_real error, const gdn_real order)
{
	const int dof = simu->m * simu->tt->nbelems * NQN;
	if (num_run == 0)
		printf("\r%-30s %-12s %-12s %-12s %-14s  %-14s  %-14s  %-14s  "
		       "%-14s  %-14s  %-14s  %-14s\n",
		       "filename", "elements", "DOF", "iter", "tmax", "cfl", "dt",
		       "hmin", "cpu_run_t", "cpu_asm_t", "error", "order");
	printf("%-30s %-12d %-12d %-12d %-6.12f  %-6.12f  %-6.12f  %-6.12f  "
	       "%-6.12f  %-6.12f  %-6.12f  ",
	       simu->tt->file_name, simu->tt->nbelems, dof, simu->itermax,
	       simu->tmax, simu->cfl, simu->dt, simu->hmin, simu->cpu_running_t,
	       simu->cpu_assembly_t, error);
	if (num_run > 0)
		printf("%-6.12f\n", order);
	else
		printf("%-14s\n", "---");
	fflush(stdout);
}

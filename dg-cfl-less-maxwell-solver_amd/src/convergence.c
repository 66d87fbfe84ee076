/*
 * Time-refinement convergence harness (src/utils/convergence.c:26-98):
 * nb_run solves on one mesh with dt halved each run, itermax doubled, tmax
 * fixed at (floor(tmax/dt0) - 1) dt0, error and observed order printed.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <dg/dg.h>
#include <mesh/mesh.h>
#include <simulation/simulation.h>
#include <utils/convergence.h>

#include <io/io_hdf5.h>

#include "host_internal.h"

static gdn_real order_of(gdn_real e1, gdn_real e0, gdn_real d1, gdn_real d0)
{
	if (!(fabs(e0) > 1e-12) || !(fabs(d0) > 1e-12) || !(e1 / e0 > 0))
		gdn_fatal("convergence: non-positive error ratio");
	return log(e1 / e0) / log(d1 / d0);
}

/* the field dump of src/utils/convergence.c:88-95 (XDMF + raw binary) */
static void dump_fields(gdn_simulation *simu, int run)
{
	char name[1024];
	snprintf(name, sizeof(name), "wave_run_%03d", run);
	io_save_all(simu, name, simu->itermax);
}

void utils_l2_time_convergence(gdn_model *model, const int nb_run,
			       const char *msh_filename, const gdn_real init_dt,
			       const gdn_real tmax, const bool export_xdmf)
{
	gdn_real *err = calloc((size_t)nb_run, sizeof(gdn_real));
	gdn_real *dt = calloc((size_t)nb_run, sizeof(gdn_real));
	int *iter = calloc((size_t)nb_run, sizeof(int));
	gdn_simulation simu = { 0 };
	gdn_mesh mesh = { 0 };

	const int iter0 = (int)floor(tmax / init_dt) - 1;
	const gdn_real tmax0 = iter0 * init_dt;
	dt[0] = init_dt;
	iter[0] = iter0;
	for (int k = 1; k < nb_run; k++) {
		dt[k] = 0.5 * dt[k - 1];
		iter[k] = 2 * iter[k - 1];
	}

	mesh_read_from_msh22_file(&mesh, msh_filename);
	mesh_build_connectivity(&mesh);
	simulation_init(&mesh, model, &simu);
	for (int k = 0; k < nb_run; k++) {
		simulation_set_time_parameters(&simu, tmax0, dt[k], -1);
		if (simu.tmax != tmax0 || simu.itermax != iter[k])
			gdn_fatal("convergence: inconsistent step count");
		dg_solve(&simu);
		err[k] = simulation_error_l2(&simu);
		const gdn_real ord =
			k > 0 ? order_of(err[k], err[k - 1], dt[k], dt[k - 1]) : -1;
		simulation_dump_info(k, &simu, err[k], ord);
		if (export_xdmf)
			dump_fields(&simu, k + 1);
	}
	simulation_free(&simu);
	free(err);
	free(dt);
	free(iter);
}

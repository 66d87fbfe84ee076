/*
 * A C driver making the reference's own call sequence against include/ and
 * libgdon3d_hip.so (the order of src/utils/convergence.c:43-98):
 * mesh_read_from_msh22_file -> mesh_build_connectivity -> simulation_init
 * -> simulation_set_time_parameters -> dg_solve -> simulation_error_l2 ->
 * simulation_dump_info -> simulation_free.
 *
 * usage: solve_driver MESH SCHEME TMAX CFL OUT
 *   SCHEME: rk3 | d3q4 | d3q6; exact data: the plane wave
 *   k = 2 pi (1, 1, 0) / sqrt 2, E0 along z (maxwell_plane_wave_set);
 *   OUT: simu->wn after the solve, raw binary64, then a text line
 *   "itermax error" on stdout after the dump line.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <dg/dg.h>
#include <mesh/mesh.h>
#include <models/d3q4.h>
#include <models/d3q6.h>
#include <models/maxwell.h>
#include <models/model.h>
#include <simulation/simulation.h>

int main(int argc, char **argv)
{
	if (argc != 6) {
		fprintf(stderr, "usage: %s MESH rk3|d3q4|d3q6 TMAX CFL OUT\n", argv[0]);
		return 2;
	}
	const double pi = 3.14159265358979323846;
	const double k[3] = { 2 * pi / sqrt(2.0), 2 * pi / sqrt(2.0), 0.0 };
	const double e0[3] = { 0.0, 0.0, 1.0 };
	maxwell_plane_wave_set(k, e0, 0.0);

	gdn_mesh mesh = { 0 };
	gdn_model model = { 0 };
	gdn_simulation simu = { 0 };
	mesh_read_from_msh22_file(&mesh, argv[1]);
	mesh_build_connectivity(&mesh);

	model.nb_w = 6;
	model.get_num_flux = maxwell_num_flux_upwind;
	model.get_num_flux_boundary = maxwell_num_flux_upwind;
	model.get_imposed_data = maxwell_imposed_plane_wave;
	model.lambda = 1.0; /* wave speed of the RK3 model (the kinetic sets override it) */
	model.use_relax_scheme = strcmp(argv[2], "rk3") != 0;
	if (!strcmp(argv[2], "d3q4"))
		d3q4_set_model(&model);
	else if (!strcmp(argv[2], "d3q6"))
		d3q6_set_model(&model);
	else if (strcmp(argv[2], "rk3")) {
		fprintf(stderr, "unknown scheme %s\n", argv[2]);
		return 2;
	}

	simulation_init(&mesh, &model, &simu);
	simulation_set_time_parameters(&simu, atof(argv[3]), 0.0, atof(argv[4]));
	dg_solve(&simu);
	const double err = simulation_error_l2(&simu);
	simulation_dump_info(0, &simu, err, 0.0);

	FILE *f = fopen(argv[5], "wb");
	if (!f || fwrite(simu.wn, sizeof(double), (size_t)simu.wlen, f) != (size_t)simu.wlen) {
		fprintf(stderr, "cannot write %s\n", argv[5]);
		return 1;
	}
	fclose(f);
	printf("%d %.17g\n", simu.itermax, err);
	simulation_free(&simu);
	return 0;
}

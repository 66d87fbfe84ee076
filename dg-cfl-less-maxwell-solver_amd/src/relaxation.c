/*
 * CFL-less kinetic relaxation driver (src/dg/relaxation.c:129-185) over the
 * GPU stepper. cpu_assembly_t covers the device set-up: exact upwind
 * classification, level sets and the per-element block factors of
 * M1 = I + dt/2 M^-1 F_v (the reference's assemble_theta_scheme with its KLU
 * factorisation). cpu_running_t covers the time loop (ghosts at t + dt/2,
 * transport sweep, relaxation per step, final ghost update at t) up to a
 * device synchronisation, without the reference's per-step printing.
 */
#include <math.h>
#include <sys/time.h>

#include <dg/relaxation.h>
#include <dg/stepper.h>
#include <utils/timing.h>

#include "host_internal.h"

void dg_solve_relaxation(gdn_simulation *simu)
{
	struct timeval start;
	tic(&start);
	gdn_stepper *st = dg_stepper_create(simu);
	simu->cpu_assembly_t = toc(&start);

	dg_stepper_init(st);
	dg_stepper_synchronize(st);

	tic(&start);
	int n = 0;
	for (double t = 0; t < simu->tmax - 1e-12; t += simu->dt)
		n++;
	dg_stepper_advance(st, n);
	dg_stepper_finish(st); /* final ghost update at t, download */
	simu->cpu_running_t = toc(&start);

	if (dg_stepper_iterations(st) != simu->itermax ||
	    !(fabs(dg_stepper_time(st) - simu->tmax) < 1e-12))
		gdn_fatal("dg_solve_relaxation: step count / final time mismatch");
	dg_stepper_free(st);
}

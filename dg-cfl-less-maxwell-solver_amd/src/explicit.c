/*
 * Low-storage RK3 driver (src/dg/explicit.c:39-117) over the GPU stepper.
 * cpu_assembly_t covers the device set-up (geometry and state upload), the
 * reference's operator assembly; cpu_running_t covers the time loop (three
 * fused stage kernels and three ghost updates per step) up to a device
 * synchronisation, as the reference's tic/toc around its loop. The
 * reference's progress printing inside the loop is left out.
 */
#include <math.h>
#include <sys/time.h>

#include <dg/explicit.h>
#include <dg/stepper.h>
#include <utils/timing.h>

#include "host_internal.h"

void dg_solve_rk3(gdn_simulation *simu)
{
	struct timeval start;
	tic(&start);
	gdn_stepper *st = dg_stepper_create(simu);
	simu->cpu_assembly_t = toc(&start);

	dg_stepper_init(st);
	dg_stepper_synchronize(st);

	tic(&start);
	/* same step count as `while (t < tmax - 1e-12) t += dt` */
	int n = 0;
	for (double t = 0; t < simu->tmax - 1e-12; t += simu->dt)
		n++;
	dg_stepper_advance(st, n);
	dg_stepper_synchronize(st);
	simu->cpu_running_t = toc(&start);

	if (!(fabs(dg_stepper_time(st) - simu->tmax) < 1e-12) ||
	    dg_stepper_iterations(st) != simu->itermax)
		gdn_fatal("dg_solve_rk3: step count / final time mismatch");
	dg_stepper_finish(st);
	dg_stepper_free(st);
}

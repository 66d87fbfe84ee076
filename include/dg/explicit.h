/*
 * Low-storage RK3 solve. Replaces src/dg/explicit.h:14 (dg_solve_rk3,
 * src/dg/explicit.c:39-117): Williamson 2-register scheme
 * A = {0,-5/9,-153/128}, B = {1/3,15/16,8/15}, C = {0,1/3,3/4}; per stage the
 * ghosts are set at t + C_i dt, dw = M^-1 F w, k2 = A_i k2 - dt dw, w += B_i k2.
 * Initialises wn with dg_init_sol_macro, runs itermax steps, fills
 * cpu_assembly_t (GPU set-up) and cpu_running_t (time loop) in seconds.
 */
#ifndef DG_EXPLICIT_H
#define DG_EXPLICIT_H

#include <simulation/simulation.h>

#ifdef __cplusplus
extern "C" {
#endif

void dg_solve_rk3(gdn_simulation *simu);

#ifdef __cplusplus
}
#endif
#endif

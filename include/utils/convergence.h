/*
 * Time-refinement convergence harness. Replaces src/utils/convergence.h:12
 * (src/utils/convergence.c:43-98): nb_run solves with dt halved each run on
 * the same mesh, printing the simulation_dump_info table. With export_xdmf the
 * final field of each run is written as a raw binary dump (gdn_io_dump_raw)
 * instead of HDF5/XDMF.
 */
#ifndef UTILS_CONVERGENCE_H
#define UTILS_CONVERGENCE_H

#include <stdbool.h>
#include <models/model.h>

#ifdef __cplusplus
extern "C" {
#endif

void utils_l2_time_convergence(gdn_model *model, const int nb_run,
			       const char *msh_filename, const gdn_real init_dt,
			       const gdn_real tmax, const bool export_xdmf);

#ifdef __cplusplus
}
#endif
#endif

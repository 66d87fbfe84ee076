/*
 * Simulation state. Replaces src/simulation/simulation.h:17-88.
 *
 * gdn_simulation is field-for-field the reference struct
 * (src/simulation/simulation.h:17-38). Layouts of wn (reference semantics):
 *  - RK3 (macro) scheme: wn[ipg * m + iw], m = nb_w
 *    (src/simulation/simulation.c:211-214);
 *  - kinetic scheme: wn[(iw * nb_v + iv) * neq + ipg], m = nb_v * nb_w
 *    (src/simulation/simulation.c:22-54);
 * with neq = 10 nbelems + 6 nboundaryfaces, ghost DoF of boundary face bf and
 * face node a at ipg = 10 nbelems + 6 bf + a.
 *
 * Unlike the reference header this one does not pull in SuiteSparse: the
 * assembled-matrix prototypes taking gdn_sparse (src/simulation/simulation.h:
 * 60-81) have no counterpart, since the GPU operator is matrix-free.
 */
#ifndef SIMULATION_H
#define SIMULATION_H

#include <gdon3d.h>
#include <mesh/mesh.h>
#include <models/model.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gdn_simulation {
	int m;
	int neq;
	int wlen;
	int itermax;

	bool use_kinetic_scheme;
	gdn_real t;
	gdn_real tmax;
	gdn_real dt;
	gdn_real cfl;
	gdn_real hmin;

	gdn_real cpu_running_t;
	gdn_real cpu_assembly_t;

	gdn_real *wn;

	gdn_model *mdl;
	gdn_mesh *tt;

} gdn_simulation;

/* src/simulation/simulation.h:40-47: micro-field access at one DoF */
void simulation_get_f(gdn_simulation *simu, const int id_pg, gdn_real *f);
void simulation_set_f(gdn_simulation *simu, const int id_pg, const gdn_real *f);
void simulation_relax_f(gdn_simulation *simu, const int id_pg,
			const gdn_real *f_old, const gdn_real *f_new);
int simulation_get_varindex(int m, int ipg, int iv);

/* src/simulation/simulation.h:49-54 */
void simulation_init(gdn_mesh *mesh, gdn_model *model, gdn_simulation *simu);
void simulation_free(gdn_simulation *simu);
void simulation_set_time_parameters(gdn_simulation *simu, const gdn_real tmax,
				    const gdn_real dt, const gdn_real cfl);

/* Declared but never defined by the reference (src/simulation/simulation.h:56);
 * here it prints the time parameters. */
void simulation_display_info(gdn_simulation *simu);

/* src/simulation/simulation.h:84-87. simulation_error_l2 runs as an FP64
 * reduction on the GPU. */
gdn_real simulation_error_l2(gdn_simulation *simu);
void simulation_dump_info(const int num_run, gdn_simulation *simu,
			  const gdn_real error, const gdn_real order);

#ifdef __cplusplus
}
#endif
#endif

/* Internal declarations shared by the C99 host sources of libgdon3d_hip. */
#ifndef GDN_HOST_INTERNAL_H
#define GDN_HOST_INTERNAL_H

#include <gdn_hip.h>
#include <simulation/simulation.h>

/* Parameters of the parameterised exact solutions of models/maxwell.h. */
struct gdn_imposed_params {
	double pw_k[3], pw_e[3], pw_h[3], pw_omega, pw_phase;
	double cm_k[3], cm_a[3], cm_omega;
	double gp_c[3], gp_sigma, gp_amp;
};
const struct gdn_imposed_params *gdn_imposed_params(void);

/* Translate the gdn_model callbacks into a device model descriptor; aborts
 * on a callback the GPU layer has no twin for. */
void gdn_model_describe(const gdn_model *mdl, int kinetic, gdh_model_desc *out);

/* Build (and cache on the simulation's mesh) the device mesh descriptor. The
 * returned arrays stay valid until gdn_mesh_desc_release. */
void gdn_mesh_describe(const gdn_mesh *msh, gdh_mesh_desc *out);
void gdn_mesh_desc_release(gdh_mesh_desc *d);

/* Per (element, local face, velocity): 0 no flux, 1 outflow, 2 inflow, with
 * the sign of v . vn computed exactly as the reference assembly does
 * (src/simulation/simulation.c:127-136,511-537). Caller frees. */
uint8_t *gdn_upwind_classes(const gdn_mesh *msh, const gdn_model *mdl);

/* Level-ordered work list of the upwind transport sweep (longest-path
 * levels of each velocity's upwind graph, Kahn's algorithm). nbr is the
 * 4-per-element neighbour array of gdh_mesh_desc, cls the classes above.
 * Where a velocity's graph has cycles, the edges that close them (the back
 * edges of a depth-first search from the elements Kahn cannot place) are
 * lagged: the consumer's face class 2 becomes 3, the sweep takes that
 * inflow as a given value, and the device solves for those values exactly
 * (capacitance form, two sweeps per step). Fills items[nv * nelem] with v * nelem + e sorted by
 * (level, v, e) and level_ptr[nlevels + 1]; returns the number of lagged
 * edges (0 for acyclic graphs). */
int gdn_sweep_schedule(int nelem, int nv, const int *nbr, uint8_t *cls,
		       int *items, int *level_ptr, int *nlevels);

/* A device context for the simulation with its state uploaded (and the
 * imposed-data arrays of GDH_IMPOSED_HOST helpers). */
gdh_ctx *gdn_ctx_open(gdn_simulation *simu, gdh_model_desc *md);
/* the context of the standalone entry points, kept between calls for the
 * same simulation (state uploaded at every call); do not destroy it */
gdh_ctx *gdn_ctx_cached(gdn_simulation *simu, gdh_model_desc *md);
/* drop the kept context (of simu, or any when simu is NULL) */
void gdn_ctx_cache_drop(const gdn_simulation *simu);
/* the same with the device state precision (GDH_FP64 or GDH_FP32) */
gdh_ctx *gdn_ctx_open_ex(gdn_simulation *simu, gdh_model_desc *md, int precision);

/* Exact solution from a host callback at the 10 nelem element nodes
 * (nb_w each) or at the 6 nboundaryfaces ghost nodes. Caller frees. */
double *gdn_eval_imposed_elements(const gdn_simulation *simu, double t);
double *gdn_eval_imposed_ghosts(const gdn_simulation *simu, double t);

void gdn_fatal(const char *msg);

#endif

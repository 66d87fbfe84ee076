/*
 * Model plugin ABI. Replaces src/models/model.h:16-62; the struct is
 * field-for-field the reference's gdn_model.
 *
 * The GPU solver cannot call host function pointers from kernels. It
 * recognises the callbacks exported by this library by address and runs
 * their device twins:
 *  - get_num_flux / get_num_flux_boundary: maxwell_num_flux_upwind,
 *    maxwell_num_flux_boundary_metal;
 *  - get_w / get_feq / get_fR: d3q4_* and d3q6_* (f_R = f_eq(w_imposed));
 *  - get_imposed_data: the maxwell_imposed_* solutions of models/maxwell.h.
 * Any other get_imposed_data is evaluated on the host at the ghost nodes and
 * uploaded each stage (slower, same results). Any other flux or kinetic
 * callback is rejected with an error message and abort().
 */
#ifndef MODEL_H
#define MODEL_H

#include <stdbool.h>
#include <gdon3d.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void (*gdn_get_num_flux)(const gdn_real *wL, const gdn_real *wR,
				 const gdn_real *vnorm, gdn_real *flux);

typedef void (*gdn_get_num_flux_boundary)(const gdn_real *wL,
					  const gdn_real *wR,
					  const gdn_real *vnorm,
					  gdn_real *flux);

typedef void (*gdn_get_imposed_data)(const gdn_real *x, const gdn_real t,
				     gdn_real *w);

typedef void (*gdn_get_proj_pos)(const gdn_real *w, const gdn_real *n,
				 gdn_real *res);

typedef void (*gdn_get_proj_neg)(const gdn_real *w, const gdn_real *n,
				 gdn_real *res);

typedef void (*gdn_get_w)(const void *self, const gdn_real *f, gdn_real *w);

typedef void (*gdn_get_feq)(const void *self, const gdn_real *w, gdn_real *feq);

typedef void (*gdn_get_fR)(const void *self, const gdn_real *fL,
			   const gdn_real *vn, const gdn_real *wI,
			   gdn_real *fR);

typedef struct gdn_model {
	char *name;
	int nb_v;
	int nb_w;
	gdn_real *vi;
	gdn_real lambda;

	/* Macro scale functions */
	gdn_get_num_flux get_num_flux;
	gdn_get_num_flux_boundary get_num_flux_boundary;
	gdn_get_imposed_data get_imposed_data;
	gdn_get_proj_pos get_proj_pos;
	gdn_get_proj_neg get_proj_neg;

	/* Micro scale functions */
	bool use_relax_scheme;
	gdn_real omega;
	gdn_get_w get_w;
	gdn_get_feq get_feq;
	gdn_get_fR get_fR;

} gdn_model;

#ifdef __cplusplus
}
#endif
#endif

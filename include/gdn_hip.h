/*
 * gdn_hip.h -- C ABI of the GPU layer (HIP, gfx950).
 *
 * This is the thin boundary between the C99 host code that implements the
 * reference API (dg_solve_rk3, dg_solve_relaxation, simulation_error_l2, ...)
 * and the hand-written HIP kernels. Everything crossing it is a plain C type:
 * pointers to host arrays, sizes and small POD descriptors. No reference
 * function maps to it one-to-one; it is what src/dg/explicit.c:39-117,
 * src/dg/relaxation.c:93-185, src/dg/dg.c:32-147 and
 * src/simulation/simulation.c:606-666 are rebuilt on.
 *
 * Errors: every call checks the HIP status and aborts with a message naming
 * the failing call (the reference aborts through assert()).
 */
#ifndef GDN_HIP_H
#define GDN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ kinds */
enum gdh_flux_kind {
	GDH_FLUX_MAXWELL_UPWIND = 1, /* maxwell_num_flux_upwind */
	GDH_FLUX_MAXWELL_METAL = 2, /* maxwell_num_flux_boundary_metal */
};

enum gdh_imposed_kind {
	GDH_IMPOSED_HOST = 0, /* unknown callback: ghosts come from the host */
	GDH_IMPOSED_MACRO_COS = 1, /* maxwell_imposed_macro_cos */
	GDH_IMPOSED_CAVITY_TM = 2, /* maxwell_imposed_macro_cavity */
	GDH_IMPOSED_STATIONARY = 3, /* maxwell_imposed_stationary */
	GDH_IMPOSED_PLANE_WAVE = 4, /* maxwell_imposed_plane_wave */
	GDH_IMPOSED_CAVITY_MODE = 5, /* maxwell_imposed_cavity_mode */
	GDH_IMPOSED_GAUSSIAN_PULSE = 6, /* maxwell_imposed_gaussian_pulse */
};

enum gdh_kinetic_kind {
	GDH_KIN_NONE = 0,
	GDH_KIN_D3Q4 = 1, /* d3q4_get_w / d3q4_get_feq / d3q4_get_fR */
	GDH_KIN_D3Q6 = 2, /* d3q6_get_w / d3q6_get_feq / d3q6_get_fR */
};

/* Boundary state of the kinetic scheme (the model's get_fR) */
enum gdh_fr_kind {
	GDH_FR_FEQ = 0, /* d3q4_get_fR / d3q6_get_fR: f_eq of the imposed data */
	GDH_FR_PROJECTOR = 1, /* d3q4_get_fR_projector / d3q6_get_fR_projector */
};

#define GDH_MAX_V 8
#define GDH_IMPOSED_NPAR 16

/* Element geometry and neighbourhood, element-major, built by the host. */
typedef struct gdh_mesh_desc {
	int nelem; /* number of tetrahedra */
	int nbface; /* number of boundary faces (ghost blocks of 6 DoFs) */
	/* nelem * 10: codtau[3][3] row-major then det (src/mesh/mesh.c:81-102) */
	const double *geom;
	/* nelem * 4: neighbour element across local face f, or -1 - bf when
	 * local face f is boundary face bf (the bf-th entry of boundaryface) */
	const int *nbr;
	/* nelem * 4: for an interior face, 4 bits per own face node a (a = 0..5,
	 * own face order t10 face2node[f]) giving the neighbour's local T10 node
	 * that carries the same point */
	const int *nbmap;
	/* nelem * 10 * 3: coordinates of the T10 nodes, element by element */
	const double *xnode;
	/* nbface * 6 * 3: coordinates of the ghost DoFs (the L node of each face
	 * node, src/dg/dg.c:87-93) */
	const double *xghost;
} gdh_mesh_desc;

/* Physics recognised from the gdn_model callbacks. */
typedef struct gdh_model_desc {
	int flux_kind; /* interior faces */
	int bflux_kind; /* boundary faces */
	int imposed_kind;
	double imposed_par[GDH_IMPOSED_NPAR];
	int kinetic_kind;
	int nb_v;
	double vi[3 * GDH_MAX_V];
	double lambda;
	double omega;
	/* boundary state (GDH_FR_*) and, for GDH_FR_PROJECTOR, the model's
	 * get_proj_pos / get_proj_neg as P = c[0] I + c[1] A^2 + c[2] A with
	 * A = A(n / |n|) */
	int fr_kind;
	double proj_pos[3];
	double proj_neg[3];
} gdh_model_desc;

typedef struct gdh_ctx gdh_ctx;

/* Number of visible GPUs (0 when none); does not create a context. */
int gdh_device_count(void);

/* Upload the mesh and the model; scheme 0 = RK3 (macro), 1 = kinetic.
 * Replaces the set-up half of dg_solve_rk3 / dg_solve_relaxation
 * (src/dg/explicit.c:25-37 assemble_rk3, src/dg/relaxation.c:24-91
 * assemble_theta_scheme): the operators are applied matrix-free. */
gdh_ctx *gdh_create(const gdh_mesh_desc *mesh, const gdh_model_desc *model,
		    int scheme);

/* Precision of the device state (gdh_create_ex). GDH_FP64: the plain form,
 * matching the reference's binary64 path. GDH_FP32: fields stored in
 * binary32 and the RK3 flux / volume arithmetic done in binary32; the
 * geometry, the transport block factors, the transport and relaxation
 * arithmetic and the L2 reductions stay in binary64.
 * No counterpart in the reference (gdn_real is fixed at compile time,
 * src/gdon3d.h). */
#define GDH_FP64 0
#define GDH_FP32 1
gdh_ctx *gdh_create_ex(const gdh_mesh_desc *mesh, const gdh_model_desc *model,
		       int scheme, int precision);
void gdh_destroy(gdh_ctx *ctx);

/* State transfer in the reference layout of simu->wn (m * neq doubles). */
void gdh_upload_state(gdh_ctx *ctx, const double *wn);
void gdh_download_state(gdh_ctx *ctx, double *wn);

/* Initial condition at every element DoF (src/dg/dg.c:32-59). For
 * GDH_IMPOSED_HOST, wexact holds get_imposed_data(x, 0) for the 10 nelem
 * element nodes (nb_w values each); else it may be NULL. */
void gdh_init_solution(gdh_ctx *ctx, const double *wexact);

/* Ghost DoFs at time t (src/dg/dg.c:61-107). For GDH_IMPOSED_HOST, wghost
 * holds get_imposed_data(x, t) at the 6 nbface ghost nodes. */
void gdh_update_boundary(gdh_ctx *ctx, double t, const double *wghost);

/* One RK3 stage (0, 1, 2) of src/dg/explicit.c:76-96 on the device state. */
void gdh_rk3_stage(gdh_ctx *ctx, int stage, double dt);

/* Kinetic solver set-up for time step dt (src/dg/relaxation.c:24-91).
 * upwind_class: nelem * 4 * nb_v bytes, 0 no flux / 1 outflow / 2 inflow
 * per (element, local face, velocity); items (nb_v * nelem entries
 * v * nelem + e) and level_ptr (nlevels + 1 offsets): the sweep schedule in
 * upwind level order, every element after all its inflow neighbours. Builds
 * the per-item 10x10 block factors on the device. */
void gdh_relax_setup(gdh_ctx *ctx, double dt, const uint8_t *upwind_class,
		     const int *items, const int *level_ptr, int nlevels);

/* Transport (src/dg/relaxation.c:93-127) + swap + relaxation
 * (src/dg/dg.c:109-147) on the device state. */
void gdh_relax_step(gdh_ctx *ctx);

/* Node-local relaxation only (src/dg/dg.c:109-147). */
void gdh_apply_relaxation(gdh_ctx *ctx);

/* src/simulation/simulation.c:606-666 on the device state at time t. For
 * GDH_IMPOSED_HOST, wexact holds the exact solution at the element nodes. */
double gdh_error_l2(gdh_ctx *ctx, double t, const double *wexact);

/* Permittivity eps > 0 per mesh element (mu = 1), no counterpart in the
 * reference (which hard-wires eps = mu = 1 in its flux,
 * src/models/maxwell.c:152-192). RK3: eps dE/dt = curl H, the E rows of each
 * stage's M^-1 F w scaled by 1/eps of the element; the upwind flux stays the
 * reference's (consistent and dissipative for any eps). Kinetic scheme: the
 * state carries the conserved fields (D, B) = (eps E, H); the transport is
 * unchanged and the equilibrium f_eq(w) takes its physical flux on
 * (D / eps, B), so that sum_v f_eq = (D, B) and sum_v v f_eq = A(e_k)
 * (E, H); initial and ghost data are f_eq of (eps E, H) of the imposed
 * (E, H) (ghosts with the eps of their element), the L2 error compares
 * D / eps with E. Call before gdh_init_solution; not with the projector
 * boundary state. */
void gdh_set_permittivity(gdh_ctx *ctx, const double *eps);

/* Discrete energy 1/2 w^T M w of the device state (M the assembled P2 mass
 * matrix, src/simulation/simulation.c:218-261; w the macro fields, or
 * sum_v f_v for the kinetic scheme), element DoFs only, in binary64 with a
 * fixed reduction order; with a permittivity 1/2 (eps |E|^2 + |H|^2)
 * (RK3) or 1/2 (|D|^2 / eps + |B|^2) (kinetic). No counterpart in the
 * reference, which reports no energy; Maxwell conserves it and the upwind
 * scheme dissipates it. */
double gdh_energy(gdh_ctx *ctx);

/* The same energy of a state held in device memory by the caller, in the
 * reference layout of simu->wn (src/simulation/simulation.c:22-40,
 * 211-214): nv = 0 the macro layout wn[6 ipg + q], nv > 0 the kinetic layout
 * wn[(q nv + v) neq + ipg], neq = 10 ne + 6 nb; det[ne] the element
 * determinants (device); the result goes to out[0] (device) on the HIP
 * stream `stream` (NULL: the null stream), which is synchronised before the
 * call returns. The partial sums live in a device buffer allocated and freed
 * by the call. Returns 0, or -1 for invalid arguments. No counterpart in
 * the reference. */
int gdh_energy_dev(const double *wn, int nv, int ne, int nb, const double *det, double *out,
		   void *stream);

/* gdh_energy_dev without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc). No counterpart in the reference. Same arguments
 * and result (bit for bit: the same kernels and reduction order); the
 * partial sums go to a fixed device array of 1024 doubles (8 KiB per GPU,
 * held until the process ends) instead of a per-call buffer, so every input
 * size fits it. The call only enqueues the two kernels on `stream` and does
 * not synchronise: out[0] is valid once the stream has reached them. Calls
 * that overlap on one GPU share the array and race; serialise them (one
 * stream). Returns 0, or -1 for invalid arguments (nothing enqueued). */
int gdh_energy_dev_noalloc(const double *wn, int nv, int ne, int nb, const double *det,
			   double *out, void *stream);

/* Event timing of the main kernels: when enabled, each sweep (kinetic) or
 * stage kernel (RK3), relaxation kernel and ghost update is bracketed by HIP
 * events recorded on the context stream (no host wait in between). This call
 * waits for the stream, and out (may be NULL) receives the seconds of device
 * time accumulated since the previous call: [0] sweep / stage,
 * [1] relaxation, [2] ghost update, [3] sweep or stage launches; the
 * counters are then reset. */
void gdh_profile(gdh_ctx *ctx, int enable, double *out);

/* Form of the RK3 stage of this context (no counterpart in the reference):
 * out[0] = 0 one element per thread over the global state, 1 the tile's
 * state staged in LDS, 2 the same with two lanes per element; out[1] = 1
 * when k2 is loaded and stored with the non-temporal hint, 0 ordinary
 * accesses; out[2] = elements per block of the global-state stage. Fixed by
 * GDN_RK3 / GDN_RK3_K2 or, for batches, by gdh_batch_autotune. */
void gdh_rk3_info(gdh_ctx *ctx, int out[3]);

/* Wait for all queued work. */
void gdh_synchronize(gdh_ctx *ctx);

/* Device time between gdh_timer_start and gdh_timer_stop, seconds (HIP
 * events on the context stream). */
void gdh_timer_start(gdh_ctx *ctx);
double gdh_timer_stop(gdh_ctx *ctx);

/* Transport over upwind cycles: out[0] = inflow faces lagged to break the
 * cycles of the velocities' upwind graphs (0: every graph acyclic, one
 * sweep per step), out[1] = sweeps of the last step, out[2] = largest number
 * of sweeps in one step so far. Where the reference factors the whole M1
 * with KLU (src/maths/sparse.c:411-466), whose block triangular form keeps
 * each cycle as one block, the GPU sweep lags the edges that close cycles
 * and solves exactly in capacitance form: x = sweep(f, z) with the lagged
 * values z from (I - T S) z = T sweep(f, 0), S the sweep's response to unit
 * lagged inputs: up to 4096 lagged nodes per velocity (GDN_CAP_MAX_NODES)
 * the matrix is built and inverted on the device at set-up and each step
 * takes two sweeps; above, GMRES with sweeps as its products (2 + its
 * products per step). */
void gdh_transport_stats(gdh_ctx *ctx, int out[3]);

/* Workgroups of the one-launch transport sweep: the number measured to be
 * resident together at set-up (no counterpart in the reference). */
int gdh_sweep_grid(gdh_ctx *ctx);
/* 1 when the relaxation of each step is applied by the next step's sweep as
 * it reads its input, 0 when it runs as a pass of its own after the sweep:
 * chosen at gdh_relax_setup by timing both on the mesh (GDN_RELAX_FUSED=0|1
 * forces one); the results agree to rounding. No counterpart in the
 * reference. */
int gdh_relax_fused(gdh_ctx *ctx);
/* Form of the transport sweep (no counterpart in the reference; the
 * reference solves with KLU, src/maths/sparse.c:411-466): out[0] = 0 one
 * launch per upwind level, 1 one item per workgroup, 2 clusters (the
 * elements of a grid cell of one velocity solved by one wavefront in level
 * order, values passed inside the cluster through LDS); out[1] = clusters,
 * out[2] = levels of the cluster graphs, out[3] = items in clusters of more
 * than one element, out[4] = resident workgroups of the chosen sweep.
 * Chosen at gdh_relax_setup by timing (GDN_SWEEP=level|item|cluster forces
 * one); all forms give bit-identical results. */
void gdh_sweep_info(gdh_ctx *ctx, int out[5]);

/* The cluster plan of the sweep on the host, without a device (no
 * counterpart in the reference): ne elements, nv velocities, vtx the 4
 * vertices of each element (12 floats), nbr the neighbours (nelem * 4, as
 * gdh_mesh_desc.nbr), ranges[0..nranges] the element ranges of batch
 * instances (nranges = 0: one range), cls / items_ve / level_ptr / nlevels
 * the schedule of gdn_sweep_schedule. Fills head (4 ints per cluster: first
 * item, n | nx << 8, inflow faces from outside as bytes 4 j + f), perm
 * (cluster-order item -> level-order item, nv * ne) and code (per item, 4
 * bits per face: external face x < 8 or 8 + in-cluster producer; bit 16:
 * a downstream neighbour lies in another cluster);
 * stats = { levels of the cluster graphs, items in clusters of more than one
 * element }. Returns the number of clusters (at most nv * ne), or -1. */
int gdh_cluster_plan(int ne, int nv, const float *vtx, const int *nbr, const int *ranges,
		     int nranges, const uint8_t *cls, const int *items_ve, const int *level_ptr,
		     int nlevels, int *head, int *perm, int *code, int stats[2]);

/* ---- batches (no counterpart in the reference: one drop-in simulation
 * per call there). The context's mesh is the disjoint union of ninst
 * instance meshes: instance i owns elements [elem_first[i], elem_first[i+1])
 * and boundary faces [bface_first[i], bface_first[i+1]), with its own
 * closed-form exact data (the model's imposed_kind with the parameter block
 * par[GDH_IMPOSED_NPAR * i ...]), time step dt[i] and clock (starting at 0).
 * eps: permittivity per element (as gdh_set_permittivity), or NULL.
 * Call after gdh_create_ex and before gdh_relax_setup (kinetic) and
 * gdh_init_solution. */
void gdh_batch_setup(gdh_ctx *ctx, int ninst, const int *elem_first,
		     const int *bface_first, const double *par, const double *dt,
		     const double *eps);
/* Tile sizes by timing candidates on the batch: out[0] elements per block of
 * the RK3 stage (32: the tile's state staged in LDS, two lanes per element;
 * 64: the same with one lane per element; 128/256: one element per thread
 * over the global state; GDN_RK3=pair|tile|soa keeps that kernel's
 * candidates only), out[1] nodes per block of the relaxation (128/256/512).
 * Call before the run starts (it overwrites the second state buffer). */
void gdh_batch_autotune(gdh_ctx *ctx, int out[2]);
/* nsteps time steps of every instance (RK3: ghosts at t + c_s dt and the
 * stage, three times; kinetic: ghosts at t + dt/2, transport, relaxation),
 * replayed from a captured HIP graph of two steps. */
void gdh_batch_advance(gdh_ctx *ctx, int nsteps);
void gdh_batch_times(gdh_ctx *ctx, double *t);
void gdh_batch_ghosts(gdh_ctx *ctx);
void gdh_batch_error_l2(gdh_ctx *ctx, double *err);

/* ---- learned block preconditioner of the transport solve (no counterpart
 * in the reference, whose KLU factors each block exactly). A network
 * 7 -> 64 -> 64 -> 100 (tanh; bf16 MFMA products, fp32 accumulation,
 * fp32 master weights) maps an item's 7 block parameters (theta dt/det times
 * v . gamma_k, k < 3, and times the outflow coefficients s_f, f < 4) to
 * P - I with P ~ D^-1. Parameters are a flat fp32 vector of
 * gdh_mlp_nparams() values: W1[64][16] (7 inputs used), b1[64], W2[64][64],
 * b2[64], W3[112][64] (100 rows used), b3[112]. */
typedef struct gdh_mlp gdh_mlp;
int gdh_mlp_nparams(void);
/* Xavier-uniform weights from SplitMix64(seed), zero biases */
gdh_mlp *gdh_mlp_create(uint64_t seed);
void gdh_mlp_destroy(gdh_mlp *mlp);
void gdh_mlp_get_params(gdh_mlp *mlp, float *out);
void gdh_mlp_set_params(gdh_mlp *mlp, const float *in);
/* mu_x[7], sd_x[7], mu_y[100], sd_y[100]: inputs and targets are used as
 * (x - mu) / sd */
void gdh_mlp_set_normalization(gdh_mlp *mlp, const float *norm214);
/* gradient (grad[nparams]) of the mean squared normalised error over the
 * n items x[n][7], y[n][100]; returns that loss */
double gdh_mlp_grad(gdh_mlp *mlp, const float *x, const float *y, int n, float *grad);
/* Adam (lr; beta 0.9, 0.999) on `steps` minibatches of `batch` items drawn
 * with SplitMix64(seed); loss_hist[steps] or NULL; returns the last loss */
double gdh_mlp_train(gdh_mlp *mlp, const float *x, const float *y, int n, int steps,
		     int batch, float lr, uint64_t seed, double *loss_hist);
/* y[n][100] = prediction of D^-1 - I (denormalised) */
void gdh_mlp_predict(gdh_mlp *mlp, const float *x, int n, float *y);
/* training data of a kinetic context after gdh_relax_setup: x[n][7] and
 * y[n][100] = (stored exact D^-1) - I of the sweep items idx[n] */
void gdh_learned_data(gdh_ctx *ctx, int n, const int *idx, float *x, float *y);
/* block solves D u = r (6 random sides per item) by Richardson with P from
 * mlp (identity if NULL) to |r - Du| <= tol |r|, exact D^-1 r after maxit
 * iterations; out[5] = mean iterations (converged), largest, fallback rate,
 * largest relative L2 to the exact solution, number of solves; iters[6 n]
 * (or NULL) per solve, -1 = fallback */
void gdh_learned_solve_stats(gdh_ctx *ctx, gdh_mlp *mlp, int n, const int *idx,
			     double tol, int maxit, double *out, int *iters);
/* the same by Richardson (method 0) or right-preconditioned GMRES (method
 * 1, restarted every 10 products, from u = 0); iterations count products
 * with D */
void gdh_learned_solve_stats_ex(gdh_ctx *ctx, gdh_mlp *mlp, int n, const int *idx,
				double tol, int maxit, int method, double *out, int *iters);

/* Learned block inverses in the transport sweep: the network mlp predicts
 * D^-1 - I of every sweep item from its seven numbers; the predictions are
 * stored in bfloat16 (200 B per item, where the exact binary64 inverse takes
 * 800 B), and the item and level sweeps rebuild D from the numbers and
 * refine the prediction by Newton-Schulz steps (R = I - D X,
 * X <- X + X R) to the binary64 inverse. An item whose residual before the
 * last step exceeds tol keeps the stored exact inverse (the fallback, and
 * the judge: out[1] counts them). A context on the cluster sweep moves to
 * the item sweep. out[0] items on the network's inverse, out[1] on the
 * stored one, out[2] largest max|X - D^-1| / max|D^-1| of the refined
 * inverses against the stored ones, out[3] median max|I - D X0| of the
 * predictions. mlp NULL: every item back on the stored inverse. No
 * counterpart in the reference (KLU factors every block exactly,
 * src/maths/sparse.c:411-466). */
void gdh_learned_enable(gdh_ctx *ctx, gdh_mlp *mlp, double tol, double out[4]);

/* Copies `bytes` (read) into a second buffer (written) `reps` times with a
 * plain streaming kernel (k_copy_calib) after one warm-up launch, at 8, 16,
 * 32 and 64 blocks per CU; returns the fastest mean milliseconds per launch. No counterpart in the reference: the
 * benchmark uses it for a measured copy bandwidth and to calibrate the
 * profiler's HBM byte counters against a kernel of known traffic. */
double gdh_copy_calibration(size_t bytes, int reps);

#ifdef __cplusplus
}
#endif
#endif

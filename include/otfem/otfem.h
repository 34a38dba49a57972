/*
 * otfem — MI355X-native corotated tetrahedral FEM step (C ABI).
 *
 * This is the drop-in boundary for the OpenTissue::fem path.  The reference
 * is a header-only C++ template API with no FFI of its own; the entry points
 * below are what a binding of that path needs, and each names the reference
 * function it replaces (paths relative to the OpenTissue tree).  The C++
 * drop-in header <otfem/fem.hpp> keeps the reference's fem::init /
 * fem::simulate / fem::update_original_coord / fem::export_K signatures on top
 * of these calls.
 *
 * Conventions
 *   - Plain pointers and sizes only; every function returns an otfem_status.
 *   - Host arrays are in the caller's node / tetrahedron order, xyz
 *     interleaved (node i at [3i, 3i+1, 3i+2]); the library keeps its own
 *     device layout and permutes at this boundary.
 *   - One HIP stream per handle.  Calls on one handle must be serialised by
 *     the caller; distinct handles are independent.
 *   - No exception crosses the ABI; otfem_last_error() gives a message.
 */
#ifndef OTFEM_OTFEM_H
#define OTFEM_OTFEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  OTFEM_OK = 0,
  OTFEM_INVALID_ARG = 1,   /* bad pointer, size, index or parameter */
  OTFEM_INVERTED_TET = 2,  /* a rest volume V <= 0 (reference asserts V>0) */
  OTFEM_NONFINITE = 3,     /* a CG scalar became NaN/Inf in some body */
  OTFEM_NOT_CONVERGED = 4, /* informational: a body hit the iteration cap */
  OTFEM_HIP_ERROR = 5,     /* HIP runtime error or device barrier time-out */
  OTFEM_NO_DEVICE = 6,     /* no usable gfx950 device */
  OTFEM_STATE = 7,         /* call out of order (e.g. step before init) */
  OTFEM_OUT_OF_MEMORY = 8  /* a host allocation failed */
} otfem_status;

/* Rotation extraction.  GRAM_SCHMIDT is the reference's
 * (fem_update_orientation.h:208 -> math_matrix3x3.h:535-557); POLAR is the
 * eigen-based polar decomposition the reference leaves commented out
 * (fem_update_orientation.h:210-213, math_polar_decomposition.h:49-91). */
enum { OTFEM_ROT_GRAM_SCHMIDT = 0, OTFEM_ROT_POLAR = 1 };

/* CG stopping rule.  ABS_R2 is the reference's
 * `iteration >= min && r.r < tol` (fem_conjugate_gradients.h:118); REL_B stops
 * when sqrt(r.r) <= tol * sqrt(b.b) over free degrees of freedom. */
enum { OTFEM_TOL_ABS_R2 = 0, OTFEM_TOL_REL_B = 1 };

/* Warm start of CG.  PREVIOUS_VELOCITY is the reference's
 * (fem_conjugate_gradients.h:48-67: CG starts from m_velocity). */
enum { OTFEM_WARM_PREVIOUS_VELOCITY = 0, OTFEM_WARM_LEARNED = 1 };

typedef struct {
  double alpha;          /* mass damping: C = alpha*M + beta*K (ref: 2, fem_simulate.h:104) */
  double beta;           /* stiffness damping (ref: 0, fem_dynamics_assembly.h:39-41) */
  int cg_min_iterations; /* ref: 20 (fem_simulate.h:108) */
  int cg_max_iterations; /* ref: 20 (fem_simulate.h:109) */
  int tol_mode;          /* OTFEM_TOL_* (ref: ABS_R2) */
  double tol;            /* ref: 1e-3 (fem_conjugate_gradients.h:44) */
  double tiny;           /* |p.Ap|, |r.r| clamp (ref: 1e-10, :43); 0 disables */
  int rotation_mode;     /* OTFEM_ROT_* (ref: GRAM_SCHMIDT) */
  int warm_start;        /* OTFEM_WARM_* (ref: PREVIOUS_VELOCITY) */
  int store_rotations;   /* 1: keep per-tet Re for otfem_get_rotations */
  int subspace;          /* 0 (ref): CG starts from the warm start above.  1..10:
                            subspace start -- CG starts from v^n + D c, where D holds
                            v^n, the differences of the last `subspace` velocities
                            and, with OTFEM_WARM_LEARNED, the network's start minus
                            v^n, and c minimises the A-norm error of the start over
                            their span (per body); no reference counterpart */
} otfem_params;

typedef struct {
  int iterations;        /* max CG iterations over bodies this step */
  int bodies_converged;  /* bodies whose stopping rule fired */
  int bodies_capped;     /* bodies that reached cg_max_iterations unconverged */
  int flags;             /* 1: a CG scalar became non-finite; 2: the learned start
                            was not finite at some node (that node used v^n) */
  double rr0;            /* body 0: r.r after the warm start */
  double rr;             /* body 0: final recursive r.r */
  double bb;             /* body 0: b.b over free dofs */
  double elapsed_ms;     /* device time of the step (HIP events) */
} otfem_stats;

typedef struct {
  int num_nodes;
  int num_tets;
  const double *coord;         /* [3N] world coords x           (m_coord) */
  const double *model_coord;   /* [3N] rest coords x0, or NULL = coord (m_model_coord) */
  const double *velocity;      /* [3N] or NULL = 0               (m_velocity) */
  const unsigned char *fixed;  /* [N]  or NULL = none            (m_fixed) */
  const int32_t *tets;         /* [4T] node indices i, j, k, m   (t4mesh_t4tetrahedron.h:39) */
  int num_bodies;              /* 0 or 1: one body; >1: CSR of bodies below */
  const int64_t *body_node_offsets; /* [B+1]: body b owns nodes [o[b], o[b+1]) */
  const int64_t *body_tet_offsets;  /* [B+1]: body b owns tets  [o[b], o[b+1]) */
  int target_tile_nodes;       /* owned nodes per device tile, 0 = default: ~512, or for
                                  one body a multiple of 256 tiles, or up to 248 smaller
                                  tiles when the body is too small for that */
} otfem_mesh_desc;

typedef struct otfem_solver *otfem_handle;

/* Library / device. */
const char *otfem_version(void);
const char *otfem_status_string(int status);
/* Reference settings (mode 0) or the relative-residual settings (mode 1:
 * alpha 0.1, beta 0.001, tol 1e-6, cap 200, tiny off). */
otfem_status otfem_default_params(otfem_params *params, int mode);
otfem_status otfem_device_count(int *count);

/* Create a solver for one mesh (or a CSR-of-bodies batch).  Replaces building
 * a fem::Mesh (fem_mesh.h:21-34, t4mesh_t4mesh.h:178-252) and is the owner of
 * all solver state the reference keeps in node/tet traits
 * (fem_node_traits.h:37-52, fem_tetrahedron_traits.h:29-48). */
otfem_status otfem_create(const otfem_mesh_desc *desc, otfem_handle *out);
/* Why the calling thread's last otfem_create failed ("" after a success); the
 * reference has no counterpart (its mesh building asserts or throws). */
const char *otfem_create_last_error(void);
otfem_status otfem_destroy(otfem_handle h);
const char *otfem_last_error(otfem_handle h);

/* fem::init (fem_init.h:45-65): uniform materials, element stiffness,
 * lumped mass (fixed nodes -> DBL_MAX), plastic parameters. */
otfem_status otfem_init(otfem_handle h, double young, double poisson, double density,
                        double c_yield, double c_creep, double c_max);
/* The user-written init that fem_init.h:32-35 invites: per-tet E, nu, rho. */
otfem_status otfem_init_materials(otfem_handle h, const double *young, const double *poisson,
                                  const double *density, double c_yield, double c_creep,
                                  double c_max);

/* Node data the reference's caller writes into traits before simulate. */
otfem_status otfem_set_external_forces(otfem_handle h, const double *f);   /* m_f_external */
otfem_status otfem_get_external_forces(otfem_handle h, double *f);
otfem_status otfem_set_state(otfem_handle h, const double *x, const double *v); /* NULL skips */
otfem_status otfem_get_state(otfem_handle h, double *x, double *v);             /* NULL skips */
otfem_status otfem_set_fixed(otfem_handle h, const unsigned char *fixed);       /* m_fixed */
otfem_status otfem_get_mass(otfem_handle h, double *mass);                      /* m_mass */
otfem_status otfem_get_volumes(otfem_handle h, double *V);                      /* m_V */
otfem_status otfem_get_rotations(otfem_handle h, double *Re);  /* [9T] row-major m_Re */
otfem_status otfem_get_plastic(otfem_handle h, double *plastic); /* [6T] m_plastic */

/* fem::simulate (fem_simulate.h:49-113): one implicit step.  params NULL means
 * the reference's settings.  stats may be NULL. */
otfem_status otfem_step(otfem_handle h, double dt, int use_stiffness_warping,
                        const otfem_params *params, otfem_stats *stats);
/* With stats == NULL otfem_step only enqueues the step on the handle's stream
 * (no host synchronisation, as a device-resident caller wants); the results
 * and any error of the enqueued steps are reported by otfem_synchronize.  No
 * reference counterpart: fem::simulate is synchronous. */
otfem_status otfem_synchronize(otfem_handle h, otfem_stats *stats);
/* otfem_step without any allocation, for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid it).
 * No reference counterpart.  It enqueues the same kernels as otfem_step with
 * stats == NULL (no captured graphs, no timing events; results and errors
 * through otfem_synchronize), and keeps its per-step scratch -- the tile
 * partial sums of the element, CG and subspace-start kernels, the
 * stopped-body and active-tile counters and the active-tile list -- in one
 * static device array of 25.1 MiB that every GPU the library runs on holds
 * until the process ends.  Limits and preconditions (OTFEM_INVALID_ARG /
 * OTFEM_STATE otherwise): at most 32,768 tiles (otfem_layout_info); no
 * recording; a subspace start or FP64-stored SpMV needs one otfem_step with
 * the same settings first (it allocates their history and record buffers);
 * the solver's device is current.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (steps of different solvers in flight on
 * different streams) race. */
otfem_status otfem_step_noalloc(otfem_handle h, double dt, int use_stiffness_warping,
                                const otfem_params *params);
/* Per-kernel HIP-event timing of every step (element kernel, CG kernels,
 * and the learned-start network when it runs), accumulated until the next
 * otfem_set_timing call.  No reference counterpart (fem has no timing hooks). */
otfem_status otfem_set_timing(otfem_handle h, int enable);
otfem_status otfem_get_timing(otfem_handle h, double *element_ms, double *cg_ms,
                              double *network_ms, long long *steps);
/* The same accumulation with the subspace start's kernels (Gram partials,
 * start and its residual) apart: ms[0] network, ms[1] element kernel,
 * ms[2] subspace start, ms[3] CG.  No reference counterpart. */
otfem_status otfem_get_timing_detail(otfem_handle h, double *ms, long long *steps);
/* CG iterations per body summed over all steps since the last reset
 * (reset != 0 clears after reading; totals may be NULL).  No reference
 * counterpart; used to count algorithmic bytes of asynchronous steps. */
otfem_status otfem_iteration_totals(otfem_handle h, int64_t *totals, int reset);
/* Per body, the number of steps since the last reset whose CG stopped at the
 * iteration cap without meeting its stopping rule.  No reference counterpart
 * (the reference reports no convergence at all). */
otfem_status otfem_unconverged_totals(otfem_handle h, int64_t *totals, int reset);
/* Kernel configuration: threads per tile in the CG kernels (256/512/1024),
 * SpMV precision inside CG (OTFEM_SPMV_*: FP32 arithmetic with FP32 or FP64
 * accumulation, FP64 arithmetic on FP32-stored element data (default), or FP64
 * arithmetic on FP64-stored element data), and whether steps are replayed as
 * captured hipGraphs (1) or launched kernel by kernel (0).  In the default
 * mode with tol_mode 1 the CG direction p is stored in FP32 (rounded before
 * every use, so r stays the residual of v), and so are r and A p once a
 * body's relative residual is below 1e4 tol (each rounded before any use;
 * the drift this leaves between r and b - A v stays ~1% of tol); v and every
 * sum stay FP64.  No reference counterpart (tuning of this implementation). */
enum { OTFEM_SPMV_F32 = 0, OTFEM_SPMV_F32_ACC64 = 1, OTFEM_SPMV_F64 = 2, OTFEM_SPMV_F64_STORE64 = 3 };
otfem_status otfem_set_kernel_config(otfem_handle h, int cg_block, int spmv_mode, int use_graphs);
/* Deterministic mode (on = 1): every per-node sum of the step has a result
 * that does not depend on the order of the atomics -- K' p in every CG
 * iteration is summed as 64-bit fixed point (integer atomics, scale chosen
 * per tile from a bound on the sums), the once-per-step sums (elastic
 * forces, K' v_start, plastic forces, the subspace start's SpMV) add the
 * node's slot terms in slot order -- so a step from the same state, with
 * the same kernel configuration, gives the same bits on every run (the dot
 * products are fixed-order trees in both modes).  CG arithmetic is FP64 in
 * this mode.  It costs a buffer of 288 B per slot and about 1.1x the step
 * time (64^3 cube: 8.29 ms against 7.42 ms).  No reference counterpart (the
 * reference is single-threaded, hence deterministic). */
otfem_status otfem_set_deterministic(otfem_handle h, int on);
/* Autotuner: times `steps` steps with each CG block size (256/512/1024
 * threads per tile) on this mesh and these settings, restores the state
 * (x, v, f_ext, plastic strain) after every trial, and keeps the fastest.
 * No reference counterpart (tuning of this implementation). */
otfem_status otfem_autotune(otfem_handle h, double dt, int use_stiffness_warping,
                            const otfem_params *params, int steps, int *best_block);
/* Layout autotuner: for each candidate target_tile_nodes it builds a
 * throw-away solver of `desc` (materials E, nu, rho per tet, loads f_ext or
 * NULL), runs the otfem_autotune sweep on it, and returns the tile target and
 * CG block size of the fastest (element + CG device time of `steps` steps;
 * *best_ms may be NULL).  candidates NULL / num_candidates 0: for one body
 * the default tiling, 512 owned nodes per tile, and 120, 256, 768 and 1,024
 * tiles (where that leaves 16 / 64 nodes per tile on average); for several
 * bodies 256, 512 and 1,024 owned nodes per tile.  Candidates that cannot be
 * built (a tile beyond the LDS budget) are skipped.  Create the real solver with
 * desc.target_tile_nodes = *best_tile_nodes and apply the block size with
 * otfem_set_kernel_config.  No reference counterpart (tuning of this
 * implementation). */
otfem_status otfem_autotune_layout(const otfem_mesh_desc *desc, const double *young,
                                   const double *poisson, const double *density,
                                   const double *f_ext, double dt, int use_stiffness_warping,
                                   const otfem_params *params, int steps,
                                   const int32_t *candidates, int num_candidates,
                                   int32_t *best_tile_nodes, int *best_block, double *best_ms);
/* Message of the calling thread's last failed otfem_autotune_layout. */
const char *otfem_autotune_layout_last_error(void);
/* Per-body CG iterations and final relative residual of the last step. */
otfem_status otfem_body_stats(otfem_handle h, int *iterations, double *rel_residual);
/* Per-body relative residual ||b - A v_start|| / ||b|| of the last step's CG
 * start (after the warm start and any subspace start).  No reference
 * counterpart (the reference reports no residuals). */
otfem_status otfem_body_start_residuals(otfem_handle h, double *rel_residual);
/* Coefficients c of the last subspace start, [B][12] (body-major; c[0] is
 * v^n's, then the differences from the newest, then the network's; unused
 * entries 0).  No reference counterpart (diagnostics of this start). */
otfem_status otfem_subspace_coefficients(otfem_handle h, double *coef);

/* fem::update_original_coord (fem_update_original_coord.h:28-33): x0 := x.
 * As in the reference, otfem_init must be called again afterwards. */
otfem_status otfem_update_original_coord(otfem_handle h);

/* fem::export_K (fem_export_k.h:23-68) as block-CSR over nodes, from the last
 * step's rotations: fixed rows hold only an identity diagonal block, fixed
 * columns are dropped.  Call with blocks == NULL to get *nnz_blocks first.
 * row_ptr [N+1], cols [nnz], blocks [9*nnz] row-major. */
otfem_status otfem_export_K(otfem_handle h, int64_t *nnz_blocks, int64_t *row_ptr,
                            int32_t *cols, double *blocks);

/* Layout facts (no device work): tiles, slots (tet copies) and halo nodes. */
otfem_status otfem_layout_info(otfem_handle h, int64_t *num_tiles, int64_t *num_slots,
                               int64_t *num_halo, int64_t *max_tile_nodes);

/* ------------------------------------------------------------------------
 * Learned CG warm start.  No reference counterpart: the reference's CG starts
 * from m_velocity (fem_conjugate_gradients.h:48-67).  A per-node network
 * 12 -> 128 -> 128 -> 128 -> 3 (SiLU), BF16 weights, FP32 accumulation,
 * predicts the CG start from the node's last three velocities (the
 * displacement increments divided by dt) and its rest stiffness diagonal:
 *   s = |v^n| + |v^n-1| + |v^n-2|,  features = (v^n, v^n-1, v^n-2) / s (9),
 *   log(1 + cK K_ii,dd / (cM m_i)) / 4 (3),  cM = 1 + alpha dt, cK = dt^2 + beta dt;
 *   start = v^n + y0 (v^n - v^n-1) + y1 (v^n-1 - v^n-2) + y2 v^n
 * (y = network output; y = 0 is the plain start, a new model's y = (2, -1, 0)
 * is quadratic extrapolation).  CG then runs to the same stopping rule as
 * with the plain start.
 * Parameters are FP32 masters, in the order W1[128][12], b1[128],
 * W2[128][128], b2, W3[128][128], b3, W4[3][128], b4[3] (row-major
 * [out][in]); kernels use BF16 copies refreshed after every change.
 * Samples are (features[12], target[4] = u*(3), loss weight); a solver with
 * recording on appends one per free node and step: u* = (v^n+1 - v^n) / s,
 * weight s^2.  With u = (start - v^n) / s = y0 D1 + y1 D2 + y2 f0..2
 * (D1 = f0..2 - f3..5, D2 = f3..5 - f6..8), the per-sample loss is
 *   L = sum_c (w_c / mean w) |u_c - u*_c|^2 / (2 n).
 * A model and the solvers it is attached to must be used from one thread. */
typedef struct otfem_model *otfem_model_handle;
enum { OTFEM_MLP_INPUTS = 12, OTFEM_MLP_OUTPUTS = 3, OTFEM_MLP_PARAMS = 35075 };

/* New model on the current device: hidden layers uniform(+-1/sqrt(fan_in)),
 * output weights zero, output bias (2, -1, 0). */
otfem_status otfem_model_create(uint64_t seed, otfem_model_handle *out);
otfem_status otfem_model_destroy(otfem_model_handle m);
const char *otfem_model_last_error(otfem_model_handle m);
otfem_status otfem_model_get_params(otfem_model_handle m, float *theta);
/* Replaces the parameters and clears the Adam state.  Non-finite values are
 * accepted: a network that produces non-finite starts falls back to v^n. */
otfem_status otfem_model_set_params(otfem_model_handle m, const float *theta);
/* Network outputs for n feature rows ([n][12] -> [n][3]) on the device. */
otfem_status otfem_model_forward(otfem_model_handle m, const float *features, int64_t n,
                                 float *out);
/* Loss and its gradient ([OTFEM_MLP_PARAMS]) over n samples, computed by the
 * device's training kernels (grad may be NULL). */
otfem_status otfem_model_gradient(otfem_model_handle m, const float *features,
                                  const float *targets, int64_t n, double *loss, float *grad);
/* Device sample store: capacity, count, clear, read back. */
otfem_status otfem_model_reserve(otfem_model_handle m, int64_t capacity);
otfem_status otfem_model_samples(otfem_model_handle m, int64_t *count, int64_t *capacity);
otfem_status otfem_model_clear_samples(otfem_model_handle m);
otfem_status otfem_model_get_samples(otfem_model_handle m, int64_t offset, int64_t n,
                                     float *features, float *targets);
/* Append n caller-made samples (features [n][12], targets [n][4]). */
otfem_status otfem_model_add_samples(otfem_model_handle m, const float *features,
                                     const float *targets, int64_t n);
/* Adam (beta 0.9, 0.999, eps 1e-8) on the per-sample loss for `steps`
 * minibatches of `batch` samples (a multiple of 128) drawn uniformly from the
 * store; loss[steps] (may be NULL) receives each minibatch's loss before its
 * update. */
otfem_status otfem_model_train(otfem_model_handle m, int steps, int batch, double lr,
                               uint64_t seed, float *loss);

/* Adam on the residual loss, for samples recorded by solver h (the store
 * must hold whole steps of h's free nodes, nothing else).  Each minibatch is
 * `steps_per_batch` recorded steps; with e = v_start - v^{n+1} and
 * e0 = v^n - v^{n+1} over the free nodes of a step,
 *   L = sum over the minibatch's steps of |A e|^2 / the same sum of |A e0|^2,
 * A = cM M + cK K_rest (lumped mass, rest stiffness; cM, cK from dt and
 * params, NULL = the relative-residual settings).  The ratio is the squared
 * residual of the learned start relative to the plain start's, with the rest
 * operator in place of the step's warped one. */
otfem_status otfem_model_train_residual(otfem_model_handle m, otfem_handle h, double dt,
                                        const otfem_params *params, int steps,
                                        int steps_per_batch, double lr, uint64_t seed,
                                        float *loss);
/* That loss and its gradient over the given recorded steps (no update). */
otfem_status otfem_model_residual_gradient(otfem_model_handle m, otfem_handle h, double dt,
                                           const otfem_params *params, const int32_t *steps,
                                           int num_steps, double *loss, float *grad);

/* Attach a model to a solver (NULL detaches).  While attached, every step
 * keeps the velocity history the features need (cleared by otfem_init and by
 * otfem_set_state with velocities).  Steps with params->warm_start == OTFEM_WARM_LEARNED
 * start CG from the network's prediction; a node whose prediction is not
 * finite starts from v^n and sets bit 2 of otfem_stats.flags. */
otfem_status otfem_set_warmstart_model(otfem_handle h, otfem_model_handle m);
/* While on, each step appends one sample per free node (in the caller's node
 * order) to the attached model's store (otfem_model_reserve first; a full
 * store fails the step). */
otfem_status otfem_set_recording(otfem_handle h, int on);

/* ---- Checkpoint / resume (no reference counterpart: the reference saves
 * geometry only) ---- */

/* Writes the solver's simulation state to `path` (binary): rest and current
 * positions, velocities, external forces, fixed flags, per-tet materials and
 * plastic parameters, plastic strain, the learned warm start's velocity
 * history (when a model is attached) and the per-body iteration counters.
 * Needs otfem_init first.  Pending steps are completed first. */
otfem_status otfem_save_state(otfem_handle h, const char *path);
/* Restores such a file into a solver created for the same mesh (node, tet and
 * body counts, connectivity and body offsets are checked; a mismatch is
 * OTFEM_INVALID_ARG): it re-initialises the element data from the saved rest
 * shape and materials and restores every array bit for bit.  The velocity
 * history is restored when a model is attached to this solver. */
otfem_status otfem_load_state(otfem_handle h, const char *path);
/* The network's FP32 parameters, Adam moments and step count. */
otfem_status otfem_model_save(otfem_model_handle m, const char *path);
otfem_status otfem_model_load(otfem_model_handle m, const char *path);

/* ---- Surface coupling (skinning a render surface onto the tetrahedra) ---- */

/* t4mesh::mesh_coupling::bind_surface (t4mesh_mesh_coupling.h:100-200): for
 * each of `num_vertices` surface vertices vertices[3n] (caller frame), a
 * tetrahedron containing it at the solver's current node positions -- all
 * barycentric coordinates in (-1e-4, 1 + 1e-4) -- as tet_index[n] (caller
 * tetrahedron order; the lowest such index; -1 if none) and the coordinates
 * of its corners j, k, m as barycentric[3n].  Host computation (uniform grid
 * over the tetrahedra's bounding boxes). */
otfem_status otfem_surface_bind(otfem_handle h, int64_t num_vertices, const double *vertices,
                                double *barycentric, int32_t *tet_index);

typedef struct otfem_surface *otfem_surface_handle;
/* A bound surface on the device (as otfem_surface_bind returns it, or any
 * caller-made binding); vertices[3n] are kept for unbound vertices
 * (tet_index -1), which otfem_surface_update leaves where they are.  The
 * surface refers to the solver: destroy it first.  No reference counterpart
 * (the reference updates surfaces on the host). */
otfem_status otfem_surface_create(otfem_handle h, int64_t num_vertices, const double *vertices,
                                  const double *barycentric, const int32_t *tet_index,
                                  otfem_surface_handle *out);
/* t4mesh::mesh_coupling::update_surface (t4mesh_mesh_coupling.h:202-238):
 * every bound vertex at w0 p_i + w1 p_j + w2 p_k + w3 p_m (w0 = 1 - w1 - w2 -
 * w3) of the solver's current node positions, computed on the device after
 * the solver's queued steps, into vertices_out[3n] (caller frame).
 * kernel_ms (may be NULL) receives the update kernel's time. */
otfem_status otfem_surface_update(otfem_surface_handle s, double *vertices_out, double *kernel_ms);
/* Vertex normals carried by the element rotations, n = Re n0, as the coupling
 * notes of t4mesh_mesh_coupling.h:171-176 describe (the reference leaves it to
 * the caller): set the rest normals n0 (normals[3n], caller frame) once, then
 * otfem_surface_update_normals writes Re n0 for every bound vertex, Re being
 * its tetrahedron's rotation from the last warped step taken with
 * store_rotations = 1 (Re = I before one; unbound vertices keep n0), into
 * normals_out[3n].  No reference counterpart as a function. */
otfem_status otfem_surface_set_normals(otfem_surface_handle s, const double *normals);
otfem_status otfem_surface_update_normals(otfem_surface_handle s, double *normals_out);
const char *otfem_surface_last_error(otfem_surface_handle s);
void otfem_surface_destroy(otfem_surface_handle s);

/* ---- Lattice bodies built on the device (the step before otfem_create) ---- */

/* Random source of the jitter: fills out[3n] with the displacements
 * (x, y, z) of the n listed nodes, in list order, and returns 0 (anything
 * else aborts the build).  Called once per round with the nodes to move. */
typedef int (*otfem_jitter_fn)(void *user, int64_t n, double *out);

typedef struct {
  int64_t num_bodies;          /* B >= 1 */
  const int32_t *dims;         /* [3B]: I, J, K cubes of body b (each >= 1) */
  double width, height, depth; /* cube edges; node (i, j, k) of a body sits at
                                  (width i, height k, depth j) */
  /* Jitter (off when jitter == 0): every interior node (not on its body's
   * bounding box) moves, then the nodes of any tetrahedron whose signed
   * volume is below min_volume are moved again from their lattice
   * position, round after round (at most max_rounds). */
  double jitter;               /* built-in source: U(-jitter e, jitter e) per axis, e that axis's edge */
  double min_volume;
  int32_t max_rounds;
  uint64_t seed;               /* built-in source */
  otfem_jitter_fn draw;        /* NULL: the built-in source */
  void *draw_user;
  /* Boundary conditions on bounding faces of every body: world axis 0 (x),
   * 1 (y) or 2 (z) and side 0 (minimum) or 1 (maximum); axis -1 for none. */
  int32_t fixed_axis, fixed_side;
  int32_t load_axis, load_side;
  const double *traction;      /* [3B] Pa per body, or [3] for all when traction_per_body == 0 */
  int32_t traction_per_body;
} otfem_blocks_desc;

/* Node and tetrahedron counts of the bodies (sum of (I+1)(J+1)(K+1) and
 * 5 I J K). */
otfem_status otfem_blocks_count(const otfem_blocks_desc *d, int64_t *num_nodes, int64_t *num_tets);

/* t4mesh::generate_blocks (t4mesh_block_generator.h:32-110) for B bodies at
 * once, packed as CSR of bodies: node order x, then J, then K; five
 * tetrahedra per cube with the parity switch of :91-106; tets[4T] hold
 * global node ids and body b owns nodes [body_node_offsets[b], [b+1]) and
 * tetrahedra [body_tet_offsets[b], [b+1]).  With a load face, f_ext[3N]
 * receives the traction as nodal forces t A / 3 per boundary triangle of
 * that face (the faces of T4BoundaryFaces, t4mesh_t4boundary_faces.h:87-150,
 * that lie on it), zero at fixed nodes; fixed[N] marks the fixed face.
 * fixed, f_ext and rounds (jitter rounds taken) may be NULL.  Computed on
 * the device; the arrays are host memory of the sizes otfem_blocks_count
 * gives. */
otfem_status otfem_blocks_build(const otfem_blocks_desc *d, double *coords, int32_t *tets,
                                uint8_t *fixed, double *f_ext, int64_t *body_node_offsets,
                                int64_t *body_tet_offsets, int32_t *rounds);
/* Message of the calling thread's last failed otfem_blocks_* call. */
const char *otfem_blocks_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* OTFEM_OTFEM_H */

/*
 * cism_gpu.h -- C ABI of libcism_gpu.so, the MI355X (gfx950) implementation of
 * CISM Glissade's higher-order (Blatter-Pattyn) velocity solve.
 *
 * Every entry point is callable from Fortran through ISO_C_BINDING (see
 * cism_amd/fortran/) and from C, C++ or Python (ctypes).  Arrays are plain
 * pointers to HOST buffers in the reference's Fortran (column-major) layouts:
 *
 *   matrix blocks  (27, nz, nx-1, ny-1)   Auu, Auv, Avu, Avv     double
 *   node vectors   (nz, nx-1, ny-1)       u, v, b, residuals      double
 *   node masks     (nz, nx-1, ny-1)       Dirichlet masks         int32
 *   vertex fields  (nx-1, ny-1)           coordinates, beta, ...  double
 *   vertex masks   (nx-1, ny-1)           active_vertex           int32 (0/1)
 *   cell masks     (nx, ny)               active_cell             int32 (0/1)
 *   layer fields   (nz-1, nx, ny)         flwafact, efvs          double
 *   indxA          (-1:1,-1:1,-1:1)       27 int32, must be the standard map
 *
 * nx, ny are the scalar-grid sizes including two halo cells on each side
 * (nx = ewn + 4), as in the reference.  Fortran LOGICAL arrays are passed as
 * int32 by the Fortran shim (LOGICAL's kind and .true. encoding are compiler
 * specific).  Units are the solver's internal ones: m, m/yr, Pa, yr.
 *
 * The library owns all device memory, inside an opaque context.  Every call is
 * synchronous (results are valid on return), runs on the context's single HIP
 * stream, and is not re-entrant for one context.  Non-zero return values are
 * the status codes below; cism_last_error() gives the message.
 */
#ifndef CISM_GPU_H
#define CISM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CISM_OK               0
#define CISM_ERR_HIP          1  /* HIP runtime error                                        */
#define CISM_ERR_SINGULAR     2  /* "Jacobian matrix is singular" (velo_higher:6458-6473)    */
#define CISM_ERR_ASYMMETRIC   3  /* element matrix asymmetric (velo_higher:9497-9527)        */
#define CISM_ERR_NAN          4  /* "PCG solver has failed, alpha = NaN" (pcg:460-463)        */
#define CISM_ERR_ARG          5  /* invalid argument (sizes, indxA, options)                  */
#define CISM_ERR_NOMEM        6  /* device allocation failed                                  */
#define CISM_ERR_UNSUPPORTED  7  /* option outside the implemented set                        */

typedef struct cism_ctx cism_ctx;

/* Context: the analogue of glissade_velo_higher_init plus the grid of
 * parallel_slap's distributed_grid (parallel_slap.F90:563-686).  periodic = 1
 * for periodic global boundaries (the serial default), 0 for outflow.  nhalo
 * must be 2 (the higher-order dycores require it, parallel_slap.F90:33-38).
 * device = HIP device ordinal. */
int  cism_ctx_create(int nx, int ny, int nz, int nhalo, int periodic, int device, cism_ctx **out);
int  cism_ctx_destroy(cism_ctx *ctx);
const char *cism_last_error(void);

/* Linear-solver constants maxiters and tolerance (pcg:69-75); defaults 200, 1e-8. */
int  cism_ctx_set_linear_solver(cism_ctx *ctx, int maxiters, double tolerance);

/* matvec_multiply_structured_3d (glissade_velo_higher_pcg.F90:2533-2647) */
int  cism_matvec_multiply_structured_3d(cism_ctx *ctx, const int32_t *indxA,
                                        const int32_t *active_vertex,
                                        const double *Auu, const double *Auv,
                                        const double *Avu, const double *Avv,
                                        const double *xu, const double *xv,
                                        double *yu, double *yv);

/* pcg_solver_standard_3d (glissade_velo_higher_pcg.F90:84-549).
 * xu, xv: initial guess in, solution out.  precond: 0 none, 1 diagonal,
 * 2 column-tridiagonal (SIA).  err, niters out. */
int  cism_pcg_solver_standard_3d(cism_ctx *ctx, const int32_t *indxA,
                                 const int32_t *active_vertex,
                                 const double *Auu, const double *Auv,
                                 const double *Avu, const double *Avv,
                                 const double *bu, const double *bv,
                                 double *xu, double *xv, int precond,
                                 double *err, int *niters);

/* pcg_solver_chrongear_3d (glissade_velo_higher_pcg.F90:965-1610) */
int  cism_pcg_solver_chrongear_3d(cism_ctx *ctx, const int32_t *indxA,
                                  const int32_t *active_vertex,
                                  const double *Auu, const double *Auv,
                                  const double *Avu, const double *Avv,
                                  const double *bu, const double *bv,
                                  double *xu, double *xv, int precond,
                                  double *err, int *niters);

/* assemble_stiffness_matrix_3d (glissade_velo_higher.F90:4954-5204), with
 * get_basis_function_derivatives_3d (:6360-6579), compute_effective_viscosity
 * (:7199-7378), compute_element_matrix (:7868-7986), the element symmetry
 * check (:9468-9529) and element_to_global_matrix_3d (:7990-8054).
 * Outputs efvs and the four blocks (all entries written). */
int  cism_assemble_stiffness_matrix_3d(cism_ctx *ctx, const double *sigma,
                                       const int32_t *active_cell,
                                       const double *xVertex, const double *yVertex,
                                       const double *uvel, const double *vvel,
                                       const double *stagusrf, const double *stagthck,
                                       const double *flwafact, int whichapprox,
                                       double *efvs, int whichefvs, double efvs_constant,
                                       double *Auu, double *Auv, double *Avu, double *Avv);

/* basal_sliding_bc (glissade_velo_higher.F90:8125-8381) on the basal layer
 * k = nz of the full 3D blocks; whichassemble_beta must be 1 (local). */
int  cism_basal_sliding_bc_3d(cism_ctx *ctx, double dx, double dy,
                              const int32_t *active_vertex, const double *beta,
                              int whichassemble_beta, double *Auu, double *Avv);

/* basal_sliding_bc (glissade_velo_higher.F90:8125-8381) with the reference's arguments:
 * whichassemble_beta 1 local (as cism_basal_sliding_bc_3d), 0 standard finite-element
 * assembly (:8236-8363): beta phi_r phi_c integrated by 2x2 Gauss points over each active
 * cell into the horizontal slots of the basal layer of the full 3D blocks. */
int  cism_basal_sliding_bc(cism_ctx *ctx, double dx, double dy, const int32_t *active_cell,
                           const int32_t *active_vertex, const double *beta,
                           const double *xVertex, const double *yVertex,
                           int whichassemble_beta, double *Auu, double *Avv);

/* dirichlet_boundary_conditions_3d (glissade_velo_higher.F90:8385-8578) */
int  cism_dirichlet_boundary_conditions_3d(cism_ctx *ctx, const int32_t *active_vertex,
                                           const int32_t *umask_dirichlet,
                                           const int32_t *vmask_dirichlet,
                                           const double *uvel, const double *vvel,
                                           double *Auu, double *Auv, double *Avu, double *Avv,
                                           double *bu, double *bv);

/* check_symmetry_assembled_matrix_3d (glissade_velo_higher.F90:9533-9741).
 * nwarn (optional) = number of pairs too asymmetric to repair. */
int  cism_check_symmetry_assembled_matrix_3d(cism_ctx *ctx, const int32_t *active_vertex,
                                             double *Auu, double *Auv, double *Avu, double *Avv,
                                             int64_t *nwarn);

/* count_nonzeros_3d (glissade_velo_higher.F90:9355-9410) */
int  cism_count_nonzeros_3d(cism_ctx *ctx, const double *Auu, const double *Auv,
                            const double *Avu, const double *Avv,
                            const int32_t *active_vertex, int64_t *nNonzeros);

/* compute_residual_vector_3d (glissade_velo_higher.F90:8768-8927) */
int  cism_compute_residual_vector_3d(cism_ctx *ctx, const int32_t *active_vertex,
                                     const double *Auu, const double *Auv,
                                     const double *Avu, const double *Avv,
                                     const double *bu, const double *bv,
                                     const double *uvel, const double *vvel,
                                     double *resid_u, double *resid_v,
                                     double *L2_norm, double *L2_norm_relative);

/* No counterpart in the reference: compute_residual_vector_3d (the entry point above) on
 * caller-owned DEVICE arrays, allocating nothing -- for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that must not call hipMalloc).  The
 * arrays are device pointers in the layouts above; norms (device, 2 doubles) receives
 * L2_norm and L2_norm_relative.  Enqueued on `stream` (a hipStream_t; NULL = the null
 * stream) and not synchronized.  The only scratch is a fixed per-GPU array of 2 x 1024
 * partial sums (16 KiB) inside the library; grids above 2^30 nodes are rejected
 * (CISM_ERR_ARG).  Calls that overlap on one GPU (other streams or threads) race on that
 * scratch.  periodic as in cism_ctx_create; device = HIP device ordinal. */
int  cism_compute_residual_vector_3d_noalloc(int nx, int ny, int nz, int periodic, int device,
                                             void *stream, const int32_t *active_vertex,
                                             const double *Auu, const double *Auv,
                                             const double *Avu, const double *Avv,
                                             const double *bu, const double *bv,
                                             const double *uvel, const double *vvel,
                                             double *resid_u, double *resid_v, double *norms);

/* compute_residual_velocity_3d (glissade_velo_higher.F90:9111-9234) */
int  cism_compute_residual_velocity_3d(cism_ctx *ctx, int whichresid,
                                       const double *uvel, const double *vvel,
                                       const double *usav, const double *vsav,
                                       double *resid_velo);

/* load_vector_gravity (glissade_velo_higher.F90:4317-4532); adds into loadu/v.
 * whichassemble_taud: 0 standard FE quadrature, 1 local. */
int  cism_load_vector_gravity(cism_ctx *ctx, const double *sigma, const double *stagwbndsigma,
                              double dx, double dy, const int32_t *active_cell,
                              const int32_t *active_vertex,
                              const double *xVertex, const double *yVertex,
                              const double *stagusrf, const double *stagthck,
                              const double *dusrf_dx, const double *dusrf_dy,
                              int whichassemble_taud, double *loadu, double *loadv);

/* load_vector_lateral_bc (glissade_velo_higher.F90:4536-4950, with
 * lateral_shelf_bc): adds the net ice-ocean pressure on faces of active cells
 * that border an ocean cell (or an inactive calving-front cell) to loadu/v.
 * land_mask is accepted for the reference's argument list and not used. */
int  cism_load_vector_lateral_bc(cism_ctx *ctx, const double *sigma, const int32_t *land_mask,
                                 const int32_t *ocean_mask, const int32_t *calving_front_mask,
                                 const int32_t *active_cell,
                                 const double *xVertex, const double *yVertex,
                                 const double *stagusrf, const double *stagthck,
                                 double *loadu, double *loadv);

/* staggered_parallel_halo_real8_3d/_4d (parallel_slap.F90:2807-2925) for an
 * array of shape (lead, nx-1, ny-1): lead = nz (vectors) or 27*nz (blocks). */
int  cism_staggered_parallel_halo(cism_ctx *ctx, int lead, double *a);

/* Device-resident Picard loop of glissade_velo_higher_solve, 3D BP branch
 * (glissade_velo_higher.F90:2003-2016, 2206-3727): the inputs are the arrays the
 * reference prepares before its loop; uvel/vvel are the initial guess in and
 * the solution out.  Host memory is touched only at entry and exit. */
typedef struct {
    int    whichbabc;          /* 4 beta large, 5 external, 6 no-slip, others: beta as given */
    int    whichefvs;          /* 0 constant, 1 flow factor, 2 nonlinear                    */
    int    whichresid;         /* 0..4, default 3 (L2 norm)                                  */
    int    whichsparse;        /* 2 standard PCG, 3 Chron-Gear                               */
    int    whichprecond;       /* 0 none, 1 diagonal, 2 SIA column tridiagonal               */
    int    whichapprox;        /* 0 SIA, 1 SSA, 2 BP (element terms of the 3D assembly)     */
    int    maxiter_nonlinear;  /* glissade_maxiter, default 100                              */
    int    precision;          /* 0 FP64 (reference), 1 FP32 matrix + FP64 refinement        */
    double efvs_constant;      /* Pa yr                                                      */
    double dx, dy;             /* m                                                          */
    double target_override;    /* > 0: replace the Picard target (1e-4 for L2)               */
    double inner_rtol;         /* precision 1: relative tolerance of each FP32 inner solve   */
    int    adaptive_linear;    /* 1: linear tolerance max(tol, gamma * |Ax-b|/|b|) per Picard
                                  iteration (inexact Picard); 0: the reference's fixed tol */
    double adaptive_gamma;     /* forcing factor of adaptive_linear (default 0.05)           */
    double stop_rel_err;       /* > 0: also stop once |x - x_ref| / |x_ref| <= stop_rel_err,
                                  x_ref set by cism_set_reference_velocity                   */
    int    anderson_m;         /* precision 1: Anderson acceleration of the Picard map with
                                  this many history columns (0 off, at most 4)              */
    int    initial_guess;      /* cism_velo_higher_solve_3d: 0 start from uvel, vvel; 2 from
                                  the context's velocity (the last solve's, or the guess of
                                  cism_learned_guess), uvel and vvel then only receive the
                                  solution                                                  */
    int    whichassemble_beta; /* which_ho_assemble_beta: 0 standard finite-element basal
                                  friction, 1 local (the default)                            */
    int    coarse;             /* precision 1 with whichprecond 2: 1 adds an aggregation coarse
                                  correction (depth-uniform aggregates of 4 x 4 columns, one
                                  l1-Jacobi V-cycle) to the SIA preconditioner of the FP32
                                  inner solves; 0 (default) the SIA preconditioner alone.  The
                                  FP64 solver keeps the reference's preconditioners.        */
    int    skip_final_residual; /* precision 1: 1 ends a linear solve after its first
                                  refinement step, without the FP64 residual that would
                                  confirm it, when the FP32 inner solve's own relative
                                  residual e (>= 1e-4, above FP32's rounding floor) times
                                  the step's FP64 residual is below half the tolerance; the
                                  next Picard iteration's residual checks the result.  Where
                                  the gather forms the FP32 matrix itself (no Dirichlet row,
                                  local basal friction) the FP64 matrix is then not stored:
                                  the few FP64 residuals left come from the element
                                  matrices, and cism_column_scale_target and the FP64
                                  matvec timing need a new assembly.  2: no FP64 matrix
                                  either, but every step confirmed, its FP64 residual from
                                  the element matrices (the same values as from the stored
                                  matrix).  0 (default): every step is confirmed in FP64
                                  with the stored matrix.                                  */
} cism_solve_opts;

typedef struct {
    int    picard_iters;
    int    linear_iters;       /* total PCG iterations (FP64 + FP32)                          */
    int    fp64_matvecs;
    int    fp32_matvecs;
    double final_L2_norm;      /* residual |Ax-b| of the last Picard iteration (old x)      */
    double final_resid_velo;
    double seconds_assembly;
    double seconds_linear;
    double seconds_total;
    int    status;
    int    refinement_steps;   /* FP64 residual evaluations of the mixed-precision solves    */
    double final_rel_err;      /* |x - x_ref| / |x_ref| at exit (-1 without a reference)     */
    double seconds_device;     /* the Picard loop alone: inputs resident, outputs not copied  */
    int    converged;          /* converged_soln of the reference (velo_higher:3731-3745):
                                  the loop ended before maxiter_nonlinear (or on stop_rel_err) */
    int    fallbacks;          /* precision 1: linear solves redone by the FP64 solver after the
                                  mixed-precision refinement stalled (the rest of the solve then
                                  stays on FP64)                                            */
    int    guess_rejected;     /* precision 1: 1 if the initial guess was no closer than zero
                                  to the solution of the first Picard system in its energy
                                  norm (x.A x / 2 - b.x > 0) and was replaced by zero         */
    int    unconverged_linear; /* linear solves that ended on their iteration budget above
                                  their tolerance (the reference's niters = maxiters case)   */
} cism_solve_stats;

void cism_solve_opts_default(cism_solve_opts *o);

int  cism_velo_higher_solve_3d(cism_ctx *ctx, const cism_solve_opts *opts,
                               const double *sigma, const int32_t *active_cell,
                               const int32_t *active_vertex,
                               const double *xVertex, const double *yVertex,
                               const double *stagusrf, const double *stagthck,
                               const double *flwafact, const double *beta_internal,
                               const int32_t *umask_dirichlet, const int32_t *vmask_dirichlet,
                               const double *loadu, const double *loadv,
                               double *uvel, double *vvel, double *efvs,
                               double *bu, double *bv, double *resid_u, double *resid_v,
                               cism_solve_stats *stats);

/* FP64 reference velocity for cism_solve_opts.stop_rel_err and the error report
 * (host arrays (nz, nx-1, ny-1)); NULL pointers clear it.  No reference
 * counterpart: the error of every fast solve is judged against an FP64 solve. */
int  cism_set_reference_velocity(cism_ctx *ctx, const double *uref, const double *vref);

/* Kernel timing on the context's stream with HIP events, for roofline reports
 * (no reference counterpart).  Each times `reps` launches on the matrix and
 * vectors left by the last solve and returns the mean duration in ms.
 *   which = 0 FP64 matvec as the FP64 solver runs it (with the fused Chron-Gear dot
 *             products): on the symmetric copy that a symmetry pass of the current
 *             matrix leaves when no pair stays unrepaired (the pass runs first, once,
 *             if the copy is not current), else, or after
 *             cism_set_fp64_symmetric(ctx, 0), on the full storage.  Not after a
 *             mixed solve with skip_final_residual, which keeps no FP64 matrix:
 *             assemble with which = 2 first
 *           1 FP32 matvec (same kernel, FP32 matrix; needs a mixed solve first)
 *           2 assembly (element kernel + node gather) of the current velocity
 *           3 plain copy of 1 GiB (k_copy_calib: 1 GiB read + 1 GiB written), the
 *             calibration of hardware byte counters
 *           4 FP32 fused Chron-Gear update with the column preconditioner (needs a
 *             mixed solve first; it changes that solve's leftover vectors)
 *           5 FP64 matvec on the full storage (all 27 slots)
 *           6 FP64 matvec on the symmetric copy in the other launch order (plain
 *             instead of XCD-aware); an error when there is no exactly symmetric copy */
int  cism_time_kernel(cism_ctx *ctx, int which, int reps, double *ms_per_launch);

/* Diagnostics of the aggregation coarse correction of the mixed-precision inner solves
 * (cism_solve_opts.coarse; no reference counterpart, for tests and tools).
 * cism_coarse_build: uploads the active-vertex and Dirichlet masks and a matrix in the
 * reference's layout (as cism_check_symmetry_assembled_matrix_3d takes it), runs the
 * symmetry pass that writes the FP32 matrix and builds the hierarchy from it; *nlevels
 * receives its level count.  A later solve replaces both.
 * cism_coarse_level: the aggregate grid of level (1-based) and, when blocks is not null,
 * its FP64 Galerkin stencil, 36 values per aggregate a = ai + ax*aj: for direction
 * d = (di+1) + 3*(dj+1) the blocks uu, uv, vu, vv of the coupling to aggregate
 * (ai+di, aj+dj), wrapped on periodic grids.
 * cism_coarse_correction: z = P V(P^T r), the term the coarse correction adds to the SIA
 * preconditioner's output, for r in the reference's vector layout (FP32 on the device). */
int  cism_coarse_build(cism_ctx *ctx, const int32_t *active_vertex, const int32_t *umask_dirichlet,
                       const int32_t *vmask_dirichlet, const double *Auu, const double *Auv,
                       const double *Avu, const double *Avv, int *nlevels);
int  cism_coarse_level(cism_ctx *ctx, int level, int *ax, int *ay, double *blocks);
int  cism_coarse_correction(cism_ctx *ctx, const double *ru, const double *rv, double *zu, double *zv);

/* The SSA branch of glissade_velo_higher_solve (solve_2d with which_ho_approx = 1,
 * glissade_velo_higher.F90:2206-3767, assemble_stiffness_matrix_2d :5208-5595, and
 * pcg_solver_*_2d, glissade_velo_higher_pcg.F90:553-961, 1614-2146) on a one-level
 * context (cism_ctx_create with nz = 1).  The arrays are the 3D solve's, with nz3
 * levels: the 2D velocity starts from the bed level of uvel/vvel and is written to
 * every level; the 2D load is the column sum of loadu/loadv; the Dirichlet masks are
 * the bed level's; flwafact (nz3-1, nx, ny) is averaged over the column; efvs
 * (nz3-1, nx, ny) out.  bu, bv, resid_u, resid_v are one-level (nx-1, ny-1). */
int  cism_velo_higher_solve_2d(cism_ctx *ctx, const cism_solve_opts *opts, int nz3,
                               const double *sigma, const int32_t *active_cell,
                               const int32_t *active_vertex, const double *xVertex,
                               const double *yVertex, const double *stagusrf,
                               const double *stagthck, const double *flwafact,
                               const double *beta_internal, const int32_t *umask_dirichlet,
                               const int32_t *vmask_dirichlet, const double *loadu,
                               const double *loadv, double *uvel, double *vvel, double *efvs,
                               double *bu, double *bv, double *resid_u, double *resid_v,
                               cism_solve_stats *stats);

/* The L1L2 and DIVA branches of glissade_velo_higher_solve (solve_2d with which_ho_approx
 * = 3 or 4; glissade_velo_higher.F90:1117-1173, 2206-3830, assemble_stiffness_matrix_2d
 * :5208-5595, compute_integrals_diva :5601-5726, compute_3d_velocity_diva :5730-5830,
 * compute_3d_velocity_L1L2 :5834-6356, compute_effective_viscosity_L1L2/_diva
 * :7382-7864) on a one-level context (cism_ctx_create with nz = 1), the arrays of the 3D
 * solve with nz3 levels:
 *   ice_mask, ice_plus_land_mask (nx, ny): L1L2's gradient and staggering mask; DIVA's
 *     staggering mask of omega (active ice or land);
 *   thck (nx, ny), dusrf_dx/dusrf_dy (nx-1, ny-1), flwa and flwafact (nz3-1, nx, ny);
 *   uvel/vvel (nz3, nx-1, ny-1) in/out: L1L2 starts from their bed level; the 3D velocity
 *     comes out (L1L2 integrated from the bed after the loop, DIVA after every solve);
 *   uvel_2d/vvel_2d (nx-1, ny-1): DIVA's depth-mean velocity in; the 2D solution out;
 *   btractx/btracty (nx-1, ny-1): DIVA's basal traction in/out (null: zero in, not out);
 *   efvs (nz3-1, nx, ny): DIVA's layer viscosities in (their quadrature-point start);
 *     every layer's viscosity of the last assembly out;
 *   bu, bv, resid_u, resid_v one-level (nx-1, ny-1).
 * The 2D load is the column sum of loadu/loadv and the Dirichlet masks are the bed
 * level's, as in cism_velo_higher_solve_2d.  DIVA requires opts->anderson_m = 0. */
int  cism_velo_higher_solve_depth(cism_ctx *ctx, const cism_solve_opts *opts, int nz3,
                                  const double *sigma, const int32_t *ice_mask,
                                  const int32_t *ice_plus_land_mask, const int32_t *active_cell,
                                  const int32_t *active_vertex, const double *xVertex,
                                  const double *yVertex, const double *thck, const double *stagthck,
                                  const double *dusrf_dx, const double *dusrf_dy, const double *flwa,
                                  const double *flwafact, const double *beta_internal,
                                  const int32_t *umask_dirichlet, const int32_t *vmask_dirichlet,
                                  const double *loadu, const double *loadv, double *uvel, double *vvel,
                                  double *uvel_2d, double *vvel_2d, double *btractx, double *btracty,
                                  double *efvs, double *bu, double *bv, double *resid_u, double *resid_v,
                                  cism_solve_stats *stats);

/* The basal_physics fields calcbeta reads (glide_types.F90:1308-1370; reference
 * defaults ho_beta_const 1000, ho_beta_large 1e10, powerlaw_c 1e4, powerlaw_m 3,
 * coulomb_c 0.42, beta_grounded_min 1, beta_powerlaw_umax 0). whichbabc is
 * which_ho_babc: 0 constant, 4 large, 5 external, 9 power law, 11 Coulomb-power
 * law (Schoof), 12 Coulomb-power law (Tsai), 14 simple. */
typedef struct {
    int whichbabc;
    int reserved;
    double ho_beta_const, ho_beta_large, powerlaw_c, powerlaw_m, coulomb_c;
    double beta_grounded_min, beta_powerlaw_umax;
} cism_basal_physics;

/* glissade_velo_higher_solve (glissade_velo_higher.F90:654-3986, the 3D branch) from the
 * ice geometry, every step on the device: the inputs as cism_glissade_velo_inputs computes
 * them (masks, staggered thickness and surface, surface gradient, vertex geometry,
 * flwafact), the grounded fraction and calcbeta in every Picard iteration (bp; NULL only
 * with opts->whichbabc = 6, no-slip), the gravity load and the lateral ice-ocean load,
 * the Dirichlet masks (the bed level under no-slip; kinbcmask and no-penetration
 * boundaries are not supported), the initial guess and the Picard loop.  Host transfers:
 * thck, usrf, topg (nx, ny) and flwa (nz-1, nx, ny) in, uvel, vvel out, and efvs when
 * non-NULL.  The call leaves bp set as the context's basal physics (as
 * cism_set_basal_physics does). */
typedef struct {
    double dx, dy;             /* m                                                          */
    double thklim;             /* ice thickness above which a cell is active (m)             */
    double eus;                /* sea level (m)                                              */
    double max_slope;          /* surface-slope limit (the reference's default is 1)         */
    int    whichground;        /* which_ho_ground: 0 no GLP, 2 all grounded                  */
    int    whichassemble_taud; /* 0 standard finite-element quadrature, 1 local             */
    int    lateral_bc;         /* 1: lateral ice-ocean pressure at marine margins            */
    int    initial_guess;      /* 0 uvel, vvel as given; 1 the learned guess (needs
                                  cism_learned_guess_load); 2 the context's last velocity    */
    int    reserved[4];
} cism_velo_setup;

int  cism_glissade_velo_higher_solve(cism_ctx *ctx, const cism_solve_opts *opts,
                                     const cism_velo_setup *setup, const double *sigma,
                                     const double *thck, const double *usrf, const double *topg,
                                     const double *flwa, const cism_basal_physics *bp,
                                     const double *beta_external, const double *effecpress_stag,
                                     double *uvel, double *vvel, double *efvs,
                                     cism_solve_stats *stats);

/* ---------------------------------------------------------------------------
 * The solve's inputs computed on the device from the ice geometry (the set-up of
 * glissade_velo_higher_solve, glissade_velo_higher.F90:1449-1939):
 *   masks            glissade_get_masks without a subgrid calving front
 *                    (glissade_masks.F90:130-210)
 *   stagthck/usrf    glissade_stagger, stagger_margin_in = 1, over ice-or-land cells
 *                    (glissade_grid_operators.F90:118-146)
 *   dusrf_dx/dy      glissade_surface_elevation_gradient, centered, hybrid margin,
 *                    |slope| <= max_slope, staggered halo (glissade_grid_operators.F90:830-990)
 *   vertex geometry  get_vertex_geometry (glissade_velo_higher.F90:4163-4313)
 *   flwafact         0.5 A^(-1/3) (glissade_velo_higher.F90:1930-1939)
 * Inputs: cell fields thck, usrf, topg (nx, ny), flwa (nz-1, nx, ny).  The results stay
 * in the context; every output pointer may be NULL (not copied). */
int  cism_glissade_velo_inputs(cism_ctx *ctx, double dx, double dy, double thklim, double eus,
                               double max_slope, const double *thck, const double *usrf,
                               const double *topg, const double *flwa, int32_t *ice_mask,
                               int32_t *floating_mask, int32_t *ocean_mask, int32_t *land_mask,
                               int32_t *active_cell, int32_t *active_vertex, double *xVertex,
                               double *yVertex, double *stagusrf, double *stagthck,
                               double *dusrf_dx, double *dusrf_dy, double *flwafact);

/* glissade_grounded_fraction (glissade_grounding_line.F90:55-178): f_ground at
 * vertices (nx-1, ny-1) from the cell masks (nx, ny).  whichground 0 (HO_GROUND_NO_GLP)
 * or 2 (HO_GROUND_ALL); 1 (GLP) is CISM_ERR_UNSUPPORTED (the reference stops on it). */
int  cism_glissade_grounded_fraction(cism_ctx *ctx, const int32_t *ice_mask,
                                     const int32_t *floating_mask, const int32_t *land_mask,
                                     int whichground, double *f_ground);

/* calcbeta (glissade_basal_traction.F90:67-598): beta at vertices from the basal
 * velocity ubas, vbas (nx-1, ny-1), weighted by f_ground, floored at
 * beta_grounded_min where f_ground > 0, staggered halo applied.  beta_external is read
 * for whichbabc 5, effecpress_stag for 11 and 12 (else may be NULL).  A negative or NaN
 * beta is CISM_ERR_NAN (the reference's GM_FATAL check). */
int  cism_calcbeta(cism_ctx *ctx, const cism_basal_physics *bp, const double *ubas,
                   const double *vbas, const double *beta_external,
                   const double *effecpress_stag, const double *f_ground, double *beta);

/* Basal physics of the device Picard loop: with bp non-NULL, every Picard iteration of
 * cism_velo_higher_solve_3d/_2d first recomputes beta by calcbeta from the current
 * basal velocity, as glissade_velo_higher_solve does (velo_higher:2502-2517), and the
 * beta_internal argument of the solve is not read.  bp = NULL switches it off. */
int  cism_set_basal_physics(cism_ctx *ctx, const cism_basal_physics *bp,
                            const double *beta_external, const double *effecpress_stag,
                            const double *f_ground);

/* ---------------------------------------------------------------------------
 * Post-solve diagnostics.  Outputs are Fortran-layout host arrays; each entry
 * point is synchronous.
 * ------------------------------------------------------------------------- */
/* compute_internal_stress (glissade_velo_higher.F90:6968-7195): element means of
 * the stress components (Pa) at cell layers (nz-1, nx, ny).  whichapprox 0..2;
 * efvs is accepted for the reference's dummy list (its L1L2 branch) and unused. */
int  cism_compute_internal_stress(cism_ctx *ctx, const double *sigma, const int32_t *active_cell,
                                  const double *xVertex, const double *yVertex,
                                  const double *stagusrf, const double *stagthck,
                                  const double *flwafact, const double *efvs, int whichefvs,
                                  double efvs_constant, int whichapprox, const double *uvel,
                                  const double *vvel, double *tau_xz, double *tau_yz,
                                  double *tau_xx, double *tau_yy, double *tau_xy, double *tau_eff);

/* compute_basal_friction_heatflx (glissade_velo_higher.F90:6755-6964): W/m^2 at
 * cells (nx, ny), halo filled (parallel_halo).  ubas, vbas, beta: vertex fields.
 * whichassemble_bfric 0 finite-element quadrature, 1 vertex values unstaggered. */
int  cism_compute_basal_friction_heatflx(cism_ctx *ctx, const int32_t *active_cell,
                                         const int32_t *active_vertex, const double *xVertex,
                                         const double *yVertex, const double *ubas,
                                         const double *vbas, const double *beta,
                                         int whichassemble_bfric, double *bfricflx);

/* basal traction (glissade_velo_higher.F90:3893-3894): beta * velocity at the bed */
int  cism_compute_basal_traction(cism_ctx *ctx, const double *beta, const double *uvel,
                                 const double *vvel, double *btractx, double *btracty);

/* All three from the context's last solve (its geometry, beta and velocity stay
 * on the device); any output pointer may be NULL. */
int  cism_velo_diagnostics(cism_ctx *ctx, int whichapprox, int whichefvs, double efvs_constant,
                           int whichassemble_bfric, double *tau_xz, double *tau_yz,
                           double *tau_xx, double *tau_yy, double *tau_xy, double *tau_eff,
                           double *bfricflx, double *btractx, double *btracty);

/* Storage of the FP64 matrix in the FP64 solver (no counterpart in the reference).
 * enable = 1 (the default): each Picard iteration's symmetry pass also writes the
 * matrix in symmetric storage (14 of 27 slots, a coupling before the diagonal read as
 * the transpose of its neighbour's), and the FP64 PCG and its residuals read that copy
 * whenever the pass left no coefficient pair unrepaired; the products, and so every
 * iterate, are the same bit for bit, from about half the matrix bytes.  enable = 0:
 * always the full storage. */
int  cism_set_fp64_symmetric(cism_ctx *ctx, int enable);

/* Per-column preconditioner scale of the mixed-precision solves (no counterpart in
 * the reference): the SIA preconditioner becomes z = s M^-1 r with one positive
 * scale per column and component, given as vertex fields scale_u, scale_v
 * (nx-1, ny-1).  NULL clears it.  The plain FP64 path (precision 0) never uses it. */
int  cism_set_precond_scale(cism_ctx *ctx, const double *scale_u, const double *scale_v);

/* Training target of that scale from the context's last solve (its FP64 matrix,
 * rhs and velocity): s = (x_c . M_c x_c) / (x_c . (A x)_c) per owned active column
 * and component (M_c: the column's vertical block), clamped to [1/16, 16]; 1
 * elsewhere.  Out: vertex fields (nx-1, ny-1). */
int  cism_column_scale_target(cism_ctx *ctx, double *scale_u, double *scale_v);

/* ---------------------------------------------------------------------------
 * Learned initial guess (no counterpart in the reference).  A fully
 * convolutional 4-level encoder-decoder (about 1.0M parameters, FP32 master
 * weights, BF16 MFMA convolutions with FP32 accumulation, periodic padding)
 * on batches of H x W images (H, W multiples of 8):
 *   input  (B, H, W, 5)  float32, NHWC: thickness, surface slopes, log beta,
 *                        log column-mean flow factor at owned vertices
 *   output (B, H, W, 24) float32: u at levels 0..10, v at levels 0..10, and
 *                        two per-column log preconditioner scales
 * Parameters are one float32 vector: per layer W[co][tap*Ci + ci] then b[co],
 * layers in forward order.  Loss: sum_p mask_p sum_{c<nch} w_c (out-tgt)^2 /
 * sum_p mask_p (w_c = 1 unless set).  One stream per network; calls are
 * synchronous and not re-entrant per handle.
 * ------------------------------------------------------------------------- */
typedef struct cism_surrogate cism_surrogate;

int  cism_surrogate_create(int H, int W, int batch, int device, cism_surrogate **out);
int  cism_surrogate_destroy(cism_surrogate *s);
long long cism_surrogate_num_params(const cism_surrogate *s);
int  cism_surrogate_get_params(cism_surrogate *s, float *params);
/* also resets the Adam moments and step count */
int  cism_surrogate_set_params(cism_surrogate *s, const float *params);
int  cism_surrogate_set_channel_weights(cism_surrogate *s, const float *w24);
int  cism_surrogate_forward(cism_surrogate *s, const float *in, float *out);
/* device time (ms) of the last forward pass: its kernels, without the host transfers */
double cism_surrogate_last_forward_ms(const cism_surrogate *s);
/* loss and its gradient w.r.t. every parameter (grad may be NULL), no update */
int  cism_surrogate_loss_grad(cism_surrogate *s, const float *in, const float *target,
                              const float *mask, int nch, double *loss, float *grad);
/* loss_grad followed by one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) */
int  cism_surrogate_train_step(cism_surrogate *s, const float *in, const float *target,
                               const float *mask, int nch, float lr, double *loss);

/* The learned initial guess of a solver context, on the device (no counterpart in the
 * reference).  cism_learned_guess_load builds the network for the context's grid and
 * sets its parameters; nz must equal the context's (the output channels hold nz levels
 * of u from channel 0 and of v from channel 11); encoding 0 decodes v = uscale o, 1
 * decodes v = sign(o) uscale (exp|o| - 1), outputs clipped to |o| <= 1000 (0) or 10 (1).
 * A periodic grid needs ewn and nsn divisible by 8; an outflow grid is embedded in a
 * zero-feature canvas.  cism_learned_guess_load_file reads a weight file: the 8 bytes
 * "CISMSG01", int32 nz, int32 encoding, float64 uscale, int64 nparams, then nparams
 * float32 (little endian).  cism_learned_guess computes the features (stagthck / 1000,
 * 100 ds/dx, 100 ds/dy, (ln max(beta, 1e-3) - 7) / 2, log10 of the column-mean flow
 * factor + 16, zero at inactive vertices), runs the network and writes the decoded
 * velocity, halos filled, into the context's velocity, where a solve with initial_guess
 * 2 starts from it.  Its inputs are host arrays (vertex fields (nx-1, ny-1), flwafact
 * (nz-1, nx, ny)); a NULL input keeps the context's current field (the one the last
 * solve or cism_glissade_velo_inputs left).  uvel, vvel (optional) receive the guess. */
int  cism_learned_guess_load(cism_ctx *ctx, const float *params, long long nparams, int nz,
                             int encoding, double uscale);
int  cism_learned_guess_load_file(cism_ctx *ctx, const char *path);
int  cism_learned_guess(cism_ctx *ctx, const double *stagthck, const double *dusrf_dx,
                        const double *dusrf_dy, const double *beta_internal,
                        const double *flwafact, const int32_t *active_vertex,
                        double *uvel, double *vvel);
/* device time (ms) of the network's forward pass in the last guess */
double cism_learned_guess_last_ms(const cism_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif

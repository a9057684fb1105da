/*
 * cism_oracle.h -- CPU oracle for the Glissade Blatter-Pattyn velocity solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm, used by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg as the checker and the timed CPU baseline.  The product
 * (libcism_gpu.so) never links, loads or calls it.
 *
 * All arrays use the reference's Fortran (column-major) layouts with 1-based
 * logical indices:
 *   matrix blocks  A(27, nz, nx-1, ny-1)     -> Auu, Auv, Avu, Avv
 *   node vectors   v(nz, nx-1, ny-1)
 *   vertex fields  s(nx-1, ny-1)
 *   cell fields    c(nx, ny)
 *   layer fields   f(nz-1, nx, ny)           (flwafact, efvs)
 * Fortran LOGICAL arrays are passed as int32 (0 / 1).
 *
 * Parity pins (which check fixes which function) are listed in the header of
 * cism_oracle.c.
 */
#ifndef CISM_ORACLE_H
#define CISM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nx, ny, nz, nhalo;   /* scalar grid (incl. halo), velocity levels, halo width */
    int ilo, ihi, jlo, jhi;  /* owned staggered vertices, 1-based inclusive          */
    int periodic;            /* 1: periodic global BC, 0: outflow                      */
} orc_grid;

/* status codes shared with the GPU library's C-ABI */
#define ORC_OK              0
#define ORC_ERR_SINGULAR    2   /* Jacobian determinant is zero                 */
#define ORC_ERR_ASYMMETRIC  3   /* element matrix asymmetric beyond 1e-10       */
#define ORC_ERR_NAN         4   /* PCG alpha is NaN                             */
#define ORC_ERR_ARG         5   /* invalid argument                             */

/* parallel_slap.F90:563-686 bounds for nhalo=2, staggered halo (2,1) */
void orc_grid_init(orc_grid *g, int nx, int ny, int nz, int periodic);

/* glissade_velo_higher_init tables (velo_higher:270-650) */
void orc_tables(double *phi, double *dxr, double *dyr, double *dzr,
                int *ishift, int *jshift, int *kshift, int *indxA);

int  orc_basis_derivatives(const double *x, const double *y, const double *z, int p,
                           double *dx, double *dy, double *dz, double *detJ);

double orc_efvs_qp(int whichefvs, int whichapprox, double efvs_constant,
                   const double *dphx, const double *dphy, const double *dphz,
                   const double *u, const double *v, double flwafact);

int  orc_element_matrix(const orc_grid *g, int i, int j, int k,
                        const double *sigma, const double *xV, const double *yV,
                        const double *uvel, const double *vvel,
                        const double *stagusrf, const double *stagthck,
                        const double *flwafact, int whichapprox, int whichefvs,
                        double efvs_constant,
                        double *Kuu, double *Kuv, double *Kvu, double *Kvv,
                        double *efvs_elem);

int  orc_assemble_3d(const orc_grid *g, const double *sigma, const int32_t *active_cell,
                     const double *xV, const double *yV,
                     const double *uvel, const double *vvel,
                     const double *stagusrf, const double *stagthck,
                     const double *flwafact, int whichapprox,
                     double *efvs, int whichefvs, double efvs_constant,
                     double *Auu, double *Auv, double *Avu, double *Avv);

/* basal_sliding_bc, HO_ASSEMBLE_BETA_STANDARD (velo_higher:8236-8363): the 2x2-point
 * finite-element integral of beta phi_r phi_c over each active cell, into the kA = 0 slots
 * of the basal layer k = nz */
void orc_basal_sliding_standard(const orc_grid *g, const int32_t *active_cell,
                                const double *xV, const double *yV, const double *beta,
                                double *Auu, double *Avv);
void orc_basal_sliding_local(const orc_grid *g, double dx, double dy,
                             const int32_t *active_vertex, const double *beta,
                             double *Auu, double *Avv);

void orc_dirichlet_3d(const orc_grid *g, const int32_t *active_vertex,
                      const int32_t *umask, const int32_t *vmask,
                      const double *uvel, const double *vvel,
                      double *Auu, double *Auv, double *Avu, double *Avv,
                      double *bu, double *bv);

/* staggered_parallel_halo for arrays of shape (lead, nx-1, ny-1) */
void orc_halo(const orc_grid *g, int lead, double *a);

int  orc_check_symmetry_3d(const orc_grid *g, const int32_t *active_vertex,
                           double *Auu, double *Auv, double *Avu, double *Avv,
                           double *maxdiff);

long orc_count_nonzeros_3d(const orc_grid *g, const int32_t *active_vertex,
                           const double *Auu, const double *Auv,
                           const double *Avu, const double *Avv);

void orc_residual_vector_3d(const orc_grid *g, const int32_t *active_vertex,
                            const double *Auu, const double *Auv,
                            const double *Avu, const double *Avv,
                            const double *bu, const double *bv,
                            const double *uvel, const double *vvel,
                            double *resid_u, double *resid_v,
                            double *L2_norm, double *L2_norm_relative);

double orc_residual_velocity_3d(const orc_grid *g, int whichresid,
                                const double *uvel, const double *vvel,
                                const double *usav, const double *vsav);

void orc_load_gravity_local(const orc_grid *g, const double *stagwbndsigma,
                            double dx, double dy, const int32_t *active_vertex,
                            const double *stagthck, const double *dusrf_dx,
                            const double *dusrf_dy, double *loadu, double *loadv);

void orc_load_gravity_standard(const orc_grid *g, const double *sigma, const int32_t *active_cell,
                               const double *xV, const double *yV,
                               const double *stagusrf, const double *stagthck,
                               const double *dusrf_dx, const double *dusrf_dy,
                               double *loadu, double *loadv);

void orc_load_lateral_bc(const orc_grid *g, const double *sigma, const int32_t *ocean_mask,
                         const int32_t *calving_front_mask, const int32_t *active_cell,
                         const double *xV, const double *yV, const double *stagusrf,
                         const double *stagthck, double *loadu, double *loadv);

void orc_matvec_3d(const orc_grid *g, const int32_t *active_vertex,
                   const double *Auu, const double *Auv,
                   const double *Avu, const double *Avv,
                   const double *xu, const double *xv, double *yu, double *yv);

void orc_tridiag(int n, const double *sb, const double *d, const double *sp,
                 const double *rhs, double *soln);

int  orc_pcg_standard_3d(const orc_grid *g, const int32_t *active_vertex,
                         const double *Auu, const double *Auv,
                         const double *Avu, const double *Avv,
                         const double *bu, const double *bv,
                         double *xu, double *xv, int precond,
                         int maxiters, double tolerance,
                         double *err, int *niters);

int  orc_pcg_chrongear_3d(const orc_grid *g, const int32_t *active_vertex,
                          const double *Auu, const double *Auv,
                          const double *Avu, const double *Avv,
                          const double *bu, const double *bv,
                          double *xu, double *xv, int precond,
                          int maxiters, double tolerance,
                          double *err, int *niters);

/* post-solve diagnostics (velo_higher:3842-3894, 6755-7195) */
void orc_cell_halo(const orc_grid *g, double *a);
int  orc_internal_stress(const orc_grid *g, const double *sigma, const int32_t *active_cell,
                         const double *xV, const double *yV, const double *stagusrf,
                         const double *stagthck, const double *flwafact, const double *efvs,
                         int whichefvs, double efvs_constant, int whichapprox,
                         const double *uvel, const double *vvel,
                         double *tau_xz, double *tau_yz, double *tau_xx, double *tau_yy,
                         double *tau_xy, double *tau_eff);
void orc_basal_friction_heatflx(const orc_grid *g, const int32_t *active_cell,
                                const int32_t *active_vertex, const double *ubas,
                                const double *vbas, const double *beta, int whichassemble_bfric,
                                double *bfricflx);
void orc_basal_traction(const orc_grid *g, const double *beta, const double *uvel,
                        const double *vvel, double *btractx, double *btracty);

/* the basal_physics fields calcbeta reads (glide_types.F90:1308-1370) */
typedef struct {
    int whichbabc, reserved;
    double ho_beta_const, ho_beta_large, powerlaw_c, powerlaw_m, coulomb_c;
    double beta_grounded_min, beta_powerlaw_umax;
} orc_basal_physics;

int orc_grounded_fraction(int nx, int ny, const int32_t *ice, const int32_t *floating,
                          const int32_t *land, int whichground, double *f_ground);
int orc_calcbeta(const orc_grid *g, const orc_basal_physics *bp, const double *ubas,
                 const double *vbas, const double *beta_external, const double *effecpress_stag,
                 const double *f_ground, double *beta);

/* Picard loop of glissade_velo_higher_solve, 3D BP branch (velo_higher:2206-3727).
 * Inputs are the arrays the reference computes before the loop. */
typedef struct {
    int whichbabc, whichefvs, whichresid, whichsparse, whichprecond;
    int whichapprox, maxiter_nonlinear, linear_maxiters;
    double efvs_constant, linear_tolerance, dx, dy;
    /* outputs */
    int picard_iters, total_linear_iters, status;
    double final_L2_norm, final_resid_velo;
    /* optional: with bp set, beta is recomputed by calcbeta from the current basal
     * velocity at the start of every Picard iteration (velo_higher:2502-2517), and
     * beta_internal is not read */
    const orc_basal_physics *bp;
    const double *beta_external, *effecpress_stag, *f_ground;
    /* optional: > 0 replaces the Picard target of whichresid (1e-4 for the L2 norm) */
    double target_override;
    /* optional output: the residual norm L2_norm measured before the linear solve of
     * each Picard iteration (velo_higher:3310-3320), for the first L2_history_cap ones */
    double *L2_history;
    int L2_history_cap;
    /* output: converged_soln of the reference, counter < maxiter_nonlinear
     * (velo_higher:3731-3745) */
    int converged;
    /* 1: basal friction by the standard finite-element assembly (which_ho_assemble_beta =
     * HO_ASSEMBLE_BETA_STANDARD); 0: the local assembly (the default, HO_ASSEMBLE_BETA_LOCAL) */
    int assemble_beta_standard;
} orc_picard_opts;

/* z = M^-1 r for precond 0 (identity), 1 (diagonal) or 2 (column tridiagonal) with the
 * preconditioner of setup_preconditioner_3d (pcg:2150-2265) built from Auu, Avv and applied
 * as the PCG loops do (pcg:300-365, easy_sia_solver pcg:2759-2848); z is left unchanged
 * in inactive columns under precond 2, as easy_sia_solver leaves it. */
void orc_precond_apply(const orc_grid *g, const int32_t *active_vertex, int precond,
                       const double *Auu, const double *Avv, const double *ru, const double *rv,
                       double *zu, double *zv);

int orc_picard_solve(const orc_grid *g, orc_picard_opts *o,
                     const double *sigma, const int32_t *active_cell,
                     const int32_t *active_vertex,
                     const double *xV, const double *yV,
                     const double *stagusrf, const double *stagthck,
                     const double *flwafact, const double *beta_internal,
                     const int32_t *umask_dirichlet, const int32_t *vmask_dirichlet,
                     const double *loadu, const double *loadv,
                     double *uvel, double *vvel, double *efvs,
                     double *bu, double *bv, double *resid_u, double *resid_v);

/* 2D (SSA) solve on a one-level grid: the nine 2D couplings sit in the 27-slot
 * layout's kA = 0 slots, so the 3D routines above serve the 2D system unchanged */
void orc_vertical_average(int nx, int ny, int nz, const double *sigma, const double *var,
                          double *var_2d);
int  orc_assemble_2d_ssa(const orc_grid *g, int nz3, const double *sigma,
                         const int32_t *active_cell, const double *xV, const double *yV,
                         const double *uvel, const double *vvel, const double *stagthck,
                         const double *flwafact, int whichefvs, double efvs_constant,
                         double *efvs, double *Auu, double *Auv, double *Avu, double *Avv);
int  orc_picard_solve_2d(const orc_grid *g, orc_picard_opts *o,
                         const double *sigma, const int32_t *active_cell,
                         const int32_t *active_vertex,
                         const double *xV, const double *yV,
                         const double *stagusrf, const double *stagthck,
                         const double *flwafact, const double *beta_internal,
                         const int32_t *umask, const int32_t *vmask,
                         const double *loadu, const double *loadv,
                         double *uvel, double *vvel, double *efvs,
                         double *bu, double *bv, double *resid_u, double *resid_v);

/* Depth-integrated 2D solvers, L1L2 (which_ho_approx = 3) and DIVA (= 4)
 * (velo_higher:5208-6356, 7382-7864; glissade_grid_operators.F90:60-410) */
void orc_stagger(int nx, int ny, const double *var, double *stagvar, const int32_t *mask, int margin);
void orc_gradient(int nx, int ny, double dx, double dy, const double *f, double *dfdx, double *dfdy,
                  const int32_t *mask, int margin);
void orc_efvs_l1l2(int whichefvs, double efvs_constant, int nz3, const double *sigma,
                   const double *phi, const double *dphi_dx, const double *dphi_dy,
                   const double *u, const double *v, const double *h, const double *dsdx,
                   const double *dsdy, const double *flwa, const double *flwafact, double *efvs);
void orc_efvs_diva(int whichefvs, double efvs_constant, int nz3, const double *stagsigma,
                   const double *phi, const double *dphi_dx, const double *dphi_dy,
                   const double *u, const double *v, const double *bx, const double *by,
                   const double *flwafact, double *efvs);
void orc_integrals_diva(int nz3, const double *sigma, double thck, const double *efvs_qp,
                        double *omega_k, double *omega);
int  orc_assemble_2d_depth(const orc_grid *g, int whichapprox, int nz3, const double *sigma,
                           const int32_t *active_cell, const double *xV, const double *yV,
                           const double *uvel, const double *vvel, const double *stagthck,
                           const double *dusrf_dx, const double *dusrf_dy, const double *thck,
                           const double *flwa, const double *flwafact, const double *btractx,
                           const double *btracty, int whichefvs, double efvs_constant, double *efvs,
                           double *efvs_qp3, double *omega_k, double *omega,
                           double *Auu, double *Auv, double *Avu, double *Avv);
int  orc_velocity_3d_l1l2(const orc_grid *g, const double *sigma, double dx, double dy,
                          const int32_t *ice_mask, const int32_t *active_cell,
                          const int32_t *active_vertex, const int32_t *umask2, const int32_t *vmask2,
                          const double *xV, const double *yV, const double *thck, const double *stagthck,
                          const double *dusrf_dx, const double *dusrf_dy, const double *flwa,
                          const double *efvs, double *uvel, double *vvel);
void orc_velocity_3d_diva(const orc_grid *g, const int32_t *active_vertex, const double *stag_omega_k,
                          const double *stag_omega, const double *btx, const double *bty,
                          const double *u2, const double *v2, double *uvel, double *vvel);

/* the fields the L1L2 and DIVA loops read beyond the SSA one's */
typedef struct {
    const double *thck;                 /* (nx, ny) m                                  */
    const double *dusrf_dx, *dusrf_dy;  /* (nx-1, ny-1) L1L2                           */
    const double *flwa;                 /* (nz-1, nx, ny) Pa^-3 yr^-1, L1L2             */
    const int32_t *ice_mask;            /* (nx, ny) L1L2 gradients and staggering       */
    const int32_t *ice_plus_land_mask;  /* (nx, ny) DIVA staggering of omega            */
    double *uvel_2d, *vvel_2d;          /* (nx-1, ny-1): DIVA in (mean velocity); out   */
    double *btractx, *btracty;          /* (nx-1, ny-1): DIVA in/out basal traction     */
} orc_depth_fields;

int  orc_picard_solve_depth(const orc_grid *g, orc_picard_opts *o, const orc_depth_fields *df,
                            const double *sigma, const int32_t *active_cell,
                            const int32_t *active_vertex, const double *xV, const double *yV,
                            const double *stagthck, const double *flwafact, const double *beta_internal,
                            const int32_t *umask, const int32_t *vmask,
                            const double *loadu, const double *loadv,
                            double *uvel, double *vvel, double *efvs,
                            double *bu, double *bv, double *resid_u, double *resid_v);

#ifdef __cplusplus
}
#endif
#endif

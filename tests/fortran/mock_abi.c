/* Test double of libcism_gpu.so for the CPU test of the Fortran module's argument
 * marshalling (tests/test_fortran_shim.py).  Each entry point checks the indxA map
 * and writes a simple, documented function of its inputs into its outputs, so the
 * test can see which Fortran array reached which C argument, in which layout. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/cism_gpu.h"

struct cism_ctx {
    int nx, ny, nz;
};

static char err_buf[128] = "";

static long nodes(const cism_ctx *c) { return (long)c->nz * (c->nx - 1) * (c->ny - 1); }

static int bad_indxA(const int32_t *ind) {
    for (int n = 0; n < 27; ++n)
        if (ind[n] != n + 1) {
            snprintf(err_buf, sizeof err_buf, "indxA[%d] = %d", n, ind[n]);
            return 1;
        }
    return 0;
}

int cism_ctx_create(int nx, int ny, int nz, int nhalo, int periodic, int device, cism_ctx **out) {
    (void)nhalo, (void)periodic, (void)device;
    cism_ctx *c = (cism_ctx *)malloc(sizeof *c);
    c->nx = nx, c->ny = ny, c->nz = nz;
    *out = c;
    return CISM_OK;
}

int cism_ctx_destroy(cism_ctx *c) {
    free(c);
    return CISM_OK;
}

const char *cism_last_error(void) { return err_buf; }

int cism_ctx_set_linear_solver(cism_ctx *c, int maxiters, double tolerance) {
    (void)c, (void)maxiters, (void)tolerance;
    return CISM_OK;
}

/* y = x + diag(A) + (active ? 1000 : 0) for u; v uses Avv and 2000 */
int cism_matvec_multiply_structured_3d(cism_ctx *c, const int32_t *indxA, const int32_t *av,
                                       const double *Auu, const double *Auv, const double *Avu,
                                       const double *Avv, const double *xu, const double *xv,
                                       double *yu, double *yv) {
    (void)Auv, (void)Avu;
    if (bad_indxA(indxA)) return CISM_ERR_ARG;
    for (long q = 0; q < nodes(c); ++q) {
        const int a = av[q / c->nz] != 0;
        yu[q] = xu[q] + Auu[27 * q + 13] + 1000.0 * a;
        yv[q] = xv[q] + Avv[27 * q + 13] + 2000.0 * a;
    }
    return CISM_OK;
}

/* x += b + Auv slot 0 (u) / Avu slot 26 (v); err = precond + 0.25; niters = 11 or 22 */
static int pcg(cism_ctx *c, int kind, const int32_t *indxA, const double *Auv, const double *Avu,
               const double *bu, const double *bv, double *xu, double *xv, int precond, double *err,
               int *niters) {
    if (bad_indxA(indxA)) return CISM_ERR_ARG;
    for (long q = 0; q < nodes(c); ++q) {
        xu[q] += bu[q] + Auv[27 * q];
        xv[q] += bv[q] + Avu[27 * q + 26];
    }
    *err = precond + 0.25;
    *niters = kind ? 22 : 11;
    return CISM_OK;
}

int cism_pcg_solver_standard_3d(cism_ctx *c, const int32_t *indxA, const int32_t *av, const double *Auu,
                                const double *Auv, const double *Avu, const double *Avv,
                                const double *bu, const double *bv, double *xu, double *xv,
                                int precond, double *err, int *niters) {
    (void)av, (void)Auu, (void)Avv;
    return pcg(c, 0, indxA, Auv, Avu, bu, bv, xu, xv, precond, err, niters);
}

int cism_pcg_solver_chrongear_3d(cism_ctx *c, const int32_t *indxA, const int32_t *av, const double *Auu,
                                 const double *Auv, const double *Avu, const double *Avv,
                                 const double *bu, const double *bv, double *xu, double *xv,
                                 int precond, double *err, int *niters) {
    (void)av, (void)Auu, (void)Avv;
    return pcg(c, 1, indxA, Auv, Avu, bu, bv, xu, xv, precond, err, niters);
}

/* efvs = 2 flwafact + active_cell + sigma(nz) ; A_b(m, q) = uvel(q) + 10 b + m/100
 * + vvel(q) * whichapprox + stagthck(col) + whichefvs * efvs_constant */
int cism_assemble_stiffness_matrix_3d(cism_ctx *c, const double *sigma, const int32_t *ac,
                                      const double *xV, const double *yV, const double *uvel,
                                      const double *vvel, const double *susrf, const double *sthck,
                                      const double *flwafact, int whichapprox, double *efvs,
                                      int whichefvs, double efvs_constant, double *Auu, double *Auv,
                                      double *Avu, double *Avv) {
    (void)xV, (void)yV, (void)susrf;
    const long ne = (long)(c->nz - 1) * c->nx * c->ny;
    for (long e = 0; e < ne; ++e) efvs[e] = 2.0 * flwafact[e] + ac[e / (c->nz - 1)] + sigma[c->nz - 1];
    double *A[4] = {Auu, Auv, Avu, Avv};
    for (long q = 0; q < nodes(c); ++q)
        for (int b = 0; b < 4; ++b)
            for (int m = 0; m < 27; ++m)
                A[b][27 * q + m] = uvel[q] + 10.0 * b + m / 100.0 + vvel[q] * whichapprox +
                                   sthck[q / c->nz] + whichefvs * efvs_constant;
    return CISM_OK;
}

void cism_solve_opts_default(cism_solve_opts *o) {
    memset(o, 0, sizeof *o);
    o->whichresid = 3;
    o->maxiter_nonlinear = 100;
    o->inner_rtol = 1e-3;
    o->adaptive_gamma = 0.05;
}

/* uvel = loadu + umask + beta(col) + whichprecond; vvel = loadv + vmask + 10 active_vertex(col)
 * + dx / 1000; efvs = flwafact + ac; bu = sigma(1) + stagusrf(col); stats: picard =
 * maxiter_nonlinear + whichbabc, linear = 100 whichsparse + 10 whichefvs + whichapprox */
int cism_velo_higher_solve_3d(cism_ctx *c, const cism_solve_opts *o, const double *sigma,
                              const int32_t *ac, const int32_t *av, const double *xV, const double *yV,
                              const double *susrf, const double *sthck, const double *flwafact,
                              const double *beta, const int32_t *umask, const int32_t *vmask,
                              const double *loadu, const double *loadv, double *uvel, double *vvel,
                              double *efvs, double *bu, double *bv, double *ru, double *rv,
                              cism_solve_stats *st) {
    (void)xV, (void)yV, (void)sthck;
    const long ne = (long)(c->nz - 1) * c->nx * c->ny;
    for (long q = 0; q < nodes(c); ++q) {
        const long col = q / c->nz;
        uvel[q] = loadu[q] + umask[q] + beta[col] + o->whichprecond;
        vvel[q] = loadv[q] + vmask[q] + 10.0 * av[col] + o->dx / 1000.0;
        bu[q] = sigma[0] + susrf[col];
        bv[q] = o->efvs_constant;
        ru[q] = rv[q] = 0.0;
    }
    for (long e = 0; e < ne; ++e) efvs[e] = flwafact[e] + ac[e / (c->nz - 1)];
    memset(st, 0, sizeof *st);
    st->picard_iters = o->maxiter_nonlinear + o->whichbabc;
    st->linear_iters = 100 * o->whichsparse + 10 * o->whichefvs + o->whichapprox;
    return CISM_OK;
}

static int guess_loaded = 0;

/* reads the 8-byte magic of the weight file; records that a guess was loaded */
int cism_learned_guess_load_file(cism_ctx *c, const char *path) {
    (void)c;
    FILE *f = fopen(path, "rb");
    char magic[9] = "";
    if (!f) return CISM_ERR_ARG;
    size_t n = fread(magic, 1, 8, f);
    fclose(f);
    if (n != 8 || memcmp(magic, "CISMSG01", 8) != 0) return CISM_ERR_ARG;
    guess_loaded = 1;
    return CISM_OK;
}

/* uvel = thck(i, j) + flwa(1, i, j) + precision + 10 initial_guess; vvel = usrf(i, j) + topg(i, j)
 * + beta_external(i, j) (0 without bp) + anderson_m; efvs = sigma(nz) + inner_rtol;
 * stats: picard = whichbabc + 100 adaptive_linear, linear = lateral_bc + 10 whichassemble_taud,
 * converged = 1, fallbacks = 100 thklim (vertex (i, j) of a node is its cell (i, j)) */
int cism_glissade_velo_higher_solve(cism_ctx *c, const cism_solve_opts *o, const cism_velo_setup *su,
                                    const double *sigma, const double *thck, const double *usrf,
                                    const double *topg, const double *flwa, const cism_basal_physics *bp,
                                    const double *bext, const double *N, double *uvel, double *vvel,
                                    double *efvs, cism_solve_stats *st) {
    (void)N;
    if (su->initial_guess == 1 && !guess_loaded) return CISM_ERR_ARG;
    const long ne = (long)(c->nz - 1) * c->nx * c->ny;
    for (long q = 0; q < nodes(c); ++q) {
        const long col = q / c->nz;
        const long i = col % (c->nx - 1), j = col / (c->nx - 1);
        const long cell = i + (long)c->nx * j;
        uvel[q] = thck[cell] + flwa[(long)(c->nz - 1) * cell] + o->precision + 10.0 * su->initial_guess;
        vvel[q] = usrf[cell] + topg[cell] + (bp ? bext[col] : 0.0) + o->anderson_m;
    }
    for (long e = 0; e < ne; ++e) efvs[e] = sigma[c->nz - 1] + o->inner_rtol;
    memset(st, 0, sizeof *st);
    st->picard_iters = o->whichbabc + 100 * o->adaptive_linear;
    st->linear_iters = su->lateral_bc + 10 * su->whichassemble_taud;
    st->converged = 1;
    st->fallbacks = (int)(100.0 * su->thklim) + 1000 * o->coarse + 2000 * o->skip_final_residual;
    return CISM_OK;
}

/* 2D solvers on a one-level context (c->nz == 1), arrays of nz3 levels.
 * SSA: uvel = loadu + umask + beta(col) + nz3; vvel = loadv + vmask + 10 active_vertex(col)
 * + whichapprox; efvs = flwafact + ac; stats: picard = maxiter_nonlinear + whichbabc,
 * linear = 100 whichsparse + 10 whichefvs + whichprecond, converged = 1 */
int cism_velo_higher_solve_2d(cism_ctx *c, const cism_solve_opts *o, int nz3, const double *sigma,
                              const int32_t *ac, const int32_t *av, const double *xV, const double *yV,
                              const double *susrf, const double *sthck, const double *flwafact,
                              const double *beta, const int32_t *umask, const int32_t *vmask,
                              const double *loadu, const double *loadv, double *uvel, double *vvel,
                              double *efvs, double *bu, double *bv, double *ru, double *rv,
                              cism_solve_stats *st) {
    (void)sigma, (void)xV, (void)yV, (void)susrf, (void)sthck;
    if (c->nz != 1) return CISM_ERR_ARG;
    const long ncol = (long)(c->nx - 1) * (c->ny - 1), ne = (long)(nz3 - 1) * c->nx * c->ny;
    for (long q = 0; q < nz3 * ncol; ++q) {
        const long col = q / nz3;
        uvel[q] = loadu[q] + umask[q] + beta[col] + nz3;
        vvel[q] = loadv[q] + vmask[q] + 10.0 * av[col] + o->whichapprox;
    }
    for (long col = 0; col < ncol; ++col) bu[col] = bv[col] = ru[col] = rv[col] = 0.0;
    for (long e = 0; e < ne; ++e) efvs[e] = flwafact[e] + ac[e / (nz3 - 1)];
    memset(st, 0, sizeof *st);
    st->picard_iters = o->maxiter_nonlinear + o->whichbabc;
    st->linear_iters = 100 * o->whichsparse + 10 * o->whichefvs + o->whichprecond;
    st->converged = 1;
    return CISM_OK;
}

/* L1L2 / DIVA: (cell of vertex column col = (i, j) is cell (i, j))
 * uvel = loadu + umask + thck(cell) + flwa(1, cell) + ice(cell) + 10 ipl(cell);
 * vvel = loadv + vmask + dusrf_dx(col) + 2 dusrf_dy(col) + beta(col) + 100 active_vertex(col);
 * uvel_2d += btractx + xVertex + stagthck; vvel_2d += btracty + yVertex + sigma(nz3);
 * btractx *= 2; btracty *= 3; efvs += flwafact + ac; stats: picard = maxiter_nonlinear +
 * whichbabc + 1000 whichapprox, linear = 100 whichsparse + 10 whichefvs + whichprecond */
int cism_velo_higher_solve_depth(cism_ctx *c, const cism_solve_opts *o, int nz3, const double *sigma,
                                 const int32_t *ice, const int32_t *ipl, const int32_t *ac, const int32_t *av,
                                 const double *xV, const double *yV, const double *thck, const double *sthck,
                                 const double *dsdx, const double *dsdy, const double *flwa,
                                 const double *flwafact, const double *beta, const int32_t *umask,
                                 const int32_t *vmask, const double *loadu, const double *loadv, double *uvel,
                                 double *vvel, double *u2, double *v2, double *btx, double *bty, double *efvs,
                                 double *bu, double *bv, double *ru, double *rv, cism_solve_stats *st) {
    if (c->nz != 1) return CISM_ERR_ARG;
    const long sx = c->nx - 1, ncol = sx * (c->ny - 1), L = nz3 - 1, ne = L * c->nx * c->ny;
    for (long q = 0; q < nz3 * ncol; ++q) {
        const long col = q / nz3, i = col % sx, j = col / sx, cell = i + (long)c->nx * j;
        uvel[q] = loadu[q] + umask[q] + thck[cell] + flwa[L * cell] + ice[cell] + 10.0 * ipl[cell];
        vvel[q] = loadv[q] + vmask[q] + dsdx[col] + 2.0 * dsdy[col] + beta[col] + 100.0 * av[col];
    }
    for (long col = 0; col < ncol; ++col) {
        u2[col] = u2[col] + btx[col] + xV[col] + sthck[col];
        v2[col] = v2[col] + bty[col] + yV[col] + sigma[nz3 - 1];
        btx[col] = 2.0 * btx[col];
        bty[col] = 3.0 * bty[col];
        bu[col] = bv[col] = ru[col] = rv[col] = 0.0;
    }
    for (long e = 0; e < ne; ++e) efvs[e] = efvs[e] + flwafact[e] + ac[e / L];
    memset(st, 0, sizeof *st);
    st->picard_iters = o->maxiter_nonlinear + o->whichbabc + 1000 * o->whichapprox;
    st->linear_iters = 100 * o->whichsparse + 10 * o->whichefvs + o->whichprecond;
    st->converged = 1;
    return CISM_OK;
}

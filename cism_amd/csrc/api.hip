// api.hip -- context management, host<->device layout conversion and the
// C-ABI entry points of libcism_gpu.so (declared in include/cism_gpu.h).
#include "internal.h"

#include <chrono>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

static thread_local char g_err[512] = "";

bool roctx_sync_enabled() {
    static const bool on = [] {
        const char *e = getenv("CISM_ROCTX_SYNC");
        return e && e[0] == '1';
    }();
    return on;
}

void cism_set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

extern "C" const char *cism_last_error(void) { return g_err; }

// ---------------------------------------------------------------------------
// layout conversion: Fortran (27, nn) block <-> tiled matrix; (nn) u, (nn) v
// <-> interleaved (u, v) pairs.  One workgroup per 64-node tile.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fortran_to_tiled(long long nn, int b,
                                                          const double *__restrict__ src,
                                                          double *__restrict__ A) {
    const long long q0 = (long long)blockIdx.x * 64;
    const int cnt = (int)((q0 + 64 <= nn ? 64 : nn - q0) * 27);
    for (int t = threadIdx.x; t < cnt; t += blockDim.x)
        A[aidx(b, t % 27, q0 + t / 27)] = src[q0 * 27 + t];
}

__global__ __launch_bounds__(256) void k_tiled_to_fortran(long long nn, int b,
                                                          const double *__restrict__ A,
                                                          double *__restrict__ dst) {
    const long long q0 = (long long)blockIdx.x * 64;
    const int cnt = (int)((q0 + 64 <= nn ? 64 : nn - q0) * 27);
    for (int t = threadIdx.x; t < cnt; t += blockDim.x)
        dst[q0 * 27 + t] = A[aidx(b, t % 27, q0 + t / 27)];
}

__global__ __launch_bounds__(256) void k_interleave(long long nn, const double *__restrict__ s,
                                                    double *__restrict__ x) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 2 * nn) x[t] = s[(t & 1) * nn + (t >> 1)];
}

__global__ __launch_bounds__(256) void k_deinterleave(long long nn, const double *__restrict__ x,
                                                      double *__restrict__ s) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 2 * nn) s[(t & 1) * nn + (t >> 1)] = x[t];
}

int ensure_stage(cism_ctx *c) {
    if (!c->stage) HIP_TRY(hipMalloc(&c->stage, sizeof(double) * 27 * (size_t)c->g.nn));
    return CISM_OK;
}

int ensure_ke(cism_ctx *c) {
    if (!c->Ke) {
        // the element tiles, then one tile of zeros that the gather reads in place of a
        // missing element (no element kernel writes it)
        const size_t tile = (size_t)CISM_KE_ENTRIES * 64;
        hipError_t e = hipMalloc(&c->Ke, sizeof(double) * (CISM_KE_ENTRIES * (size_t)c->g.nepad + tile));
        if (e != hipSuccess) {
            c->Ke = nullptr;
            cism_set_error("element scratch allocation failed: %s", hipGetErrorString(e));
            return CISM_ERR_NOMEM;
        }
        HIP_TRY(hipMemsetAsync(c->Ke + CISM_KE_ENTRIES * (size_t)c->g.nepad, 0, sizeof(double) * tile, c->stream));
    }
    return CISM_OK;
}

static int upload_block(cism_ctx *c, int b, const double *h) {
    c->a_zero_ok = 0;   // the caller's matrix may hold anything in inactive rows
    c->a64_valid = 1;
    c->a64s_valid = 0;
    HIP_TRY(hipMemcpyAsync(c->stage, h, sizeof(double) * 27 * c->g.nn, hipMemcpyHostToDevice,
                           c->stream));
    hipLaunchKernelGGL(k_fortran_to_tiled, dim3((unsigned)((c->g.nn + 63) / 64)), dim3(256), 0,
                       c->stream, c->g.nn, b, c->stage, c->A);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

static int download_block(cism_ctx *c, int b, double *h) {
    hipLaunchKernelGGL(k_tiled_to_fortran, dim3((unsigned)((c->g.nn + 63) / 64)), dim3(256), 0,
                       c->stream, c->g.nn, b, c->A, c->stage);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(h, c->stage, sizeof(double) * 27 * c->g.nn, hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));   // the staging buffer is reused by the next block
    return CISM_OK;
}

int upload_matrix(cism_ctx *c, const double *Auu, const double *Auv, const double *Avu,
                  const double *Avv) {
    CISM_TRY(ensure_stage(c));
    const double *h[4] = {Auu, Auv, Avu, Avv};
    for (int b = 0; b < 4; ++b) {
        CISM_TRY(upload_block(c, b, h[b]));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return CISM_OK;
}

int download_matrix(cism_ctx *c, double *Auu, double *Auv, double *Avu, double *Avv) {
    CISM_TRY(ensure_stage(c));
    double *h[4] = {Auu, Auv, Avu, Avv};
    for (int b = 0; b < 4; ++b) CISM_TRY(download_block(c, b, h[b]));
    return CISM_OK;
}

int upload_vec(cism_ctx *c, double *dst, const double *hu, const double *hv) {
    const long long nn = c->g.nn;
    HIP_TRY(hipMemcpyAsync(c->vstage, hu, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->vstage + nn, hv, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_interleave, dim3(nblocks_for(2 * nn)), dim3(256), 0, c->stream, nn, c->vstage,
                       dst);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));   // vstage is reused by the next transfer
    return CISM_OK;
}

int download_vec(cism_ctx *c, const double *src, double *hu, double *hv) {
    const long long nn = c->g.nn;
    hipLaunchKernelGGL(k_deinterleave, dim3(nblocks_for(2 * nn)), dim3(256), 0, c->stream, nn, src,
                       c->vstage);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(hu, c->vstage, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(hv, c->vstage + nn, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CISM_OK;
}

template <typename T>
static int up(cism_ctx *c, T *dst, const T *src, size_t n) {
    HIP_TRY(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyHostToDevice, c->stream));
    return CISM_OK;
}

// uploads the Dirichlet masks (n entries each) and records whether any node is
// prescribed (mask value 1); with none, the Dirichlet update changes nothing and is skipped
static int upload_dirichlet_masks(cism_ctx *c, const int32_t *um, const int32_t *vm, size_t n) {
    CISM_TRY(up(c, c->umask, um, n));
    CISM_TRY(up(c, c->vmask, vm, n));
    int any = 0;
    for (size_t t = 0; t < n && !any; ++t) any = um[t] == 1 || vm[t] == 1;
    c->dirichlet_any = any;
    return CISM_OK;
}
template <typename T>
static int down(cism_ctx *c, T *dst, const T *src, size_t n) {
    HIP_TRY(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyDeviceToHost, c->stream));
    return CISM_OK;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
static int check_indxA(const int32_t *indxA) {
    if (!indxA) return CISM_OK;
    for (int kA = -1; kA <= 1; ++kA)
        for (int jA = -1; jA <= 1; ++jA)
            for (int iA = -1; iA <= 1; ++iA)
                if (indxA[(iA + 1) + 3 * (jA + 1) + 9 * (kA + 1)] != slot_of(iA, jA, kA) + 1) {
                    cism_set_error("indxA is not the standard 27-point map");
                    return CISM_ERR_ARG;
                }
    return CISM_OK;
}

template <typename T>
static int dalloc(T **p, size_t n) {
    hipError_t e = hipMalloc(p, sizeof(T) * (n ? n : 1));
    if (e != hipSuccess) {
        cism_set_error("device allocation of %zu bytes failed: %s", sizeof(T) * n, hipGetErrorString(e));
        return CISM_ERR_NOMEM;
    }
    // hipMemset runs on the null stream, which the context's non-blocking stream does
    // not wait for: finish it here, or a later upload on the context stream can land
    // first and be zeroed (the first call on a fresh context then sees wiped inputs)
    HIP_TRY(hipMemset(*p, 0, sizeof(T) * (n ? n : 1)));
    HIP_TRY(hipDeviceSynchronize());
    return CISM_OK;
}

extern "C" int cism_ctx_create(int nx, int ny, int nz, int nhalo, int periodic, int device,
                               cism_ctx **out) {
    if (!out) return CISM_ERR_ARG;
    *out = nullptr;
    if (nhalo != CISM_NHALO || nx < 7 || ny < 7 || nz < 1) {
        cism_set_error("invalid grid nx=%d ny=%d nz=%d nhalo=%d (need nhalo=2, nx,ny>=7, nz>=1)", nx,
                       ny, nz, nhalo);
        return CISM_ERR_ARG;
    }
    HIP_TRY(hipSetDevice(device));
    cism_ctx *c = (cism_ctx *)calloc(1, sizeof(cism_ctx));
    if (!c) return CISM_ERR_NOMEM;
    GridDev &g = c->g;
    g.nx = nx; g.ny = ny; g.nz = nz; g.sx = nx - 1; g.sy = ny - 1;
    g.periodic = periodic ? 1 : 0;
    // parallel_slap.F90:661-673, 0-based
    g.ilo = (periodic ? 3 : 2) - 1;
    g.jlo = (periodic ? 3 : 2) - 1;
    g.ihi = (nx - 2) - 1;
    g.jhi = (ny - 2) - 1;
    g.nn = (long long)nz * g.sx * g.sy;
    g.npad = (g.nn + 255) / 256 * 256;
    g.ne = (long long)(nz - 1) * nx * ny;
    g.nepad = (g.ne + 255) / 256 * 256;
    if (g.npad > (1LL << 30) || g.nepad > (1LL << 30)) {
        // the assembly kernels index nodes and elements with 32-bit integers
        free(c);
        cism_set_error("grid too large: %lld nodes, %lld elements (limit 2^30 each)", g.nn, g.ne);
        return CISM_ERR_ARG;
    }
    c->device = device;
    c->maxiters = 200;
    c->tolerance = 1.0e-8;
    int st = CISM_OK;
    auto fail = [&](int s) { cism_ctx_destroy(c); return s; };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(CISM_ERR_HIP);
    const size_t v2 = 2 * (size_t)g.npad, ncol = (size_t)g.sx * g.sy;
    if ((st = dalloc(&c->A, 108 * (size_t)g.npad))) return fail(st);
    double **vecs[] = {&c->b, &c->x, &c->r, &c->d, &c->z, &c->q, &c->Az, &c->t, &c->load, &c->usav,
                       &c->vstage};
    for (double **v : vecs)
        if ((st = dalloc(v, v2))) return fail(st);
    if ((st = dalloc(&c->F64, 6 * (size_t)g.npad))) return fail(st);
    if ((st = dalloc(&c->av, ncol)) || (st = dalloc(&c->ac, (size_t)nx * ny)) ||
        (st = dalloc(&c->umask, (size_t)g.nn)) || (st = dalloc(&c->vmask, (size_t)g.nn)))
        return fail(st);
    double **geo[] = {&c->xV, &c->yV, &c->susrf, &c->sthck, &c->beta, &c->dsdx, &c->dsdy};
    for (double **v : geo)
        if ((st = dalloc(v, ncol))) return fail(st);
    if ((st = dalloc(&c->sigma, (size_t)nz)) || (st = dalloc(&c->wsig, (size_t)nz + 1)) ||
        (st = dalloc(&c->flwafact, (size_t)g.nepad)) || (st = dalloc(&c->efvs, (size_t)g.nepad)))
        return fail(st);
    const size_t nparts = (size_t)(xcd_blocks_for(g.nn) > CISM_RED_BLOCKS ? xcd_blocks_for(g.nn) : CISM_RED_BLOCKS);
    // per-block partial sums: up to CISM_MAX_NV values per block (the widest reduction)
    if ((st = dalloc(&c->partials, nparts * CISM_MAX_NV))) return fail(st);
    if ((st = dalloc(&c->red_groups, (size_t)64 * CISM_MAX_NV))) return fail(st);
    if ((st = dalloc(&c->av_zero, ncol)) || (st = dalloc(&c->ac_zero, (size_t)nx * ny)) ||
        (st = dalloc(&c->as_av_prev, ncol)) || (st = dalloc(&c->as_ac_prev, (size_t)nx * ny)))
        return fail(st);
    if ((st = dalloc(&c->sc, 1))) return fail(st);
    if (hipHostMalloc(&c->sc_host, sizeof(DevScal), hipHostMallocDefault) != hipSuccess)
        return fail(CISM_ERR_HIP);
    c->a64_valid = 1;
    c->sym64_on = 1;
    *out = c;
    return CISM_OK;
}

extern "C" int cism_ctx_destroy(cism_ctx *c) {
    if (!c) return CISM_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    learned_guess_free(c);
    coarse_free(c);
    void *ptrs[] = {c->A, c->A32, c->A64s, c->b, c->x, c->r, c->d, c->z, c->q, c->Az, c->t, c->load, c->usav, c->vstage,
                    c->x32, c->b32, c->r32, c->d32, c->z32, c->q32, c->Az32, c->F64, c->F32, c->av, c->ac, c->umask, c->vmask,
                    c->xV, c->yV, c->susrf, c->sthck, c->beta, c->sigma, c->wsig, c->flwafact, c->efvs,
                    c->dsdx, c->dsdy, c->Ke, c->av_zero, c->ac_zero, c->as_av_prev, c->as_ac_prev, c->xref, c->partials, c->red_groups, c->sc, c->stage,
                    c->cscale, c->aa, c->aa_sums, c->ssa_Ke, c->ssa_ff2, c->ssa_efvs2, c->ssa_sigma,
                    c->ssa_flw3, c->bp_bext, c->bp_N, c->bp_fg, c->planes, c->dep_flwa, c->dep_thck,
                    c->dep_efvs3, c->dep_eqp, c->dep_omk, c->dep_om, c->dep_beff, c->dep_som, c->dep_btr,
                    c->dep_u3, c->dep_v3, c->dep_cs, c->dep_ub, c->dep_ice, c->dep_ipl};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (c->sc_host) (void)hipHostFree(c->sc_host);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    free(c);
    return CISM_OK;
}

extern "C" int cism_ctx_set_linear_solver(cism_ctx *c, int maxiters, double tolerance) {
    if (!c || maxiters < 1 || !(tolerance > 0)) return CISM_ERR_ARG;
    c->maxiters = maxiters;
    c->tolerance = tolerance;
    return CISM_OK;
}

#define CTX_BEGIN(c)                                \
    do {                                            \
        if (!(c)) {                                 \
            cism_set_error("null context");         \
            return CISM_ERR_ARG;                    \
        }                                           \
        HIP_TRY(hipSetDevice((c)->device));         \
    } while (0)

// ---------------------------------------------------------------------------
// stateless entry points (host arrays in and out)
// ---------------------------------------------------------------------------
extern "C" int cism_matvec_multiply_structured_3d(cism_ctx *c, const int32_t *indxA,
                                                  const int32_t *active_vertex,
                                                  const double *Auu, const double *Auv,
                                                  const double *Avu, const double *Avv,
                                                  const double *xu, const double *xv,
                                                  double *yu, double *yv) {
    CTX_BEGIN(c);
    CISM_TRY(check_indxA(indxA));
    CISM_TRY(up(c, c->av, active_vertex, (size_t)c->g.sx * c->g.sy));
    CISM_TRY(upload_matrix(c, Auu, Auv, Avu, Avv));
    CISM_TRY(upload_vec(c, c->x, xu, xv));
    CISM_TRY(launch_matvec64(c, c->A, c->x, c->q));
    return download_vec(c, c->q, yu, yv);
}

static int pcg_entry(cism_ctx *c, int chrongear, const int32_t *indxA, const int32_t *active_vertex,
                     const double *Auu, const double *Auv, const double *Avu, const double *Avv,
                     const double *bu, const double *bv, double *xu, double *xv, int precond,
                     double *err, int *niters) {
    CTX_BEGIN(c);
    CISM_TRY(check_indxA(indxA));
    if (precond < 0 || precond > 2) {
        cism_set_error("precond must be 0, 1 or 2 (got %d)", precond);
        return CISM_ERR_ARG;
    }
    CISM_TRY(up(c, c->av, active_vertex, (size_t)c->g.sx * c->g.sy));
    CISM_TRY(upload_matrix(c, Auu, Auv, Avu, Avv));
    CISM_TRY(upload_vec(c, c->b, bu, bv));
    CISM_TRY(upload_vec(c, c->x, xu, xv));
    double e = 0.0;
    int n = 0;
    int st = solver_pcg64(c, chrongear, precond, c->maxiters, c->tolerance, &e, &n);
    if (st) return st;
    if (err) *err = e;
    if (niters) *niters = n;
    return download_vec(c, c->x, xu, xv);
}

extern "C" int cism_pcg_solver_standard_3d(cism_ctx *c, const int32_t *indxA,
                                           const int32_t *active_vertex, const double *Auu,
                                           const double *Auv, const double *Avu, const double *Avv,
                                           const double *bu, const double *bv, double *xu, double *xv,
                                           int precond, double *err, int *niters) {
    return pcg_entry(c, 0, indxA, active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond, err,
                     niters);
}

extern "C" int cism_pcg_solver_chrongear_3d(cism_ctx *c, const int32_t *indxA,
                                            const int32_t *active_vertex, const double *Auu,
                                            const double *Auv, const double *Avu, const double *Avv,
                                            const double *bu, const double *bv, double *xu,
                                            double *xv, int precond, double *err, int *niters) {
    return pcg_entry(c, 1, indxA, active_vertex, Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond, err,
                     niters);
}

static int check_options(int whichapprox, int whichefvs) {
    if (whichapprox < 0 || whichapprox > 2 || whichefvs < 0 || whichefvs > 2) {
        cism_set_error("unsupported whichapprox=%d / whichefvs=%d for the 3D assembly", whichapprox,
                       whichefvs);
        return CISM_ERR_UNSUPPORTED;
    }
    return CISM_OK;
}

static int upload_geometry(cism_ctx *c, const double *sigma, const int32_t *active_cell,
                           const double *xVertex, const double *yVertex, const double *stagusrf,
                           const double *stagthck, const double *flwafact) {
    const GridDev &g = c->g;
    const size_t ncol = (size_t)g.sx * g.sy;
    CISM_TRY(up(c, c->sigma, sigma, (size_t)g.nz));
    CISM_TRY(up(c, c->ac, active_cell, (size_t)g.nx * g.ny));
    CISM_TRY(up(c, c->xV, xVertex, ncol));
    CISM_TRY(up(c, c->yV, yVertex, ncol));
    CISM_TRY(up(c, c->susrf, stagusrf, ncol));
    CISM_TRY(up(c, c->sthck, stagthck, ncol));
    CISM_TRY(up(c, c->flwafact, flwafact, (size_t)g.ne));
    return CISM_OK;
}

static int check_flags(cism_ctx *c) {
    CISM_TRY(sync_scalars(c));
    if (c->sc_host->singular_flag) {
        cism_set_error("Jacobian matrix is singular");
        return CISM_ERR_SINGULAR;
    }
    return CISM_OK;
}

extern "C" int cism_assemble_stiffness_matrix_3d(cism_ctx *c, const double *sigma,
                                                 const int32_t *active_cell, const double *xVertex,
                                                 const double *yVertex, const double *uvel,
                                                 const double *vvel, const double *stagusrf,
                                                 const double *stagthck, const double *flwafact,
                                                 int whichapprox, double *efvs, int whichefvs,
                                                 double efvs_constant, double *Auu, double *Auv,
                                                 double *Avu, double *Avv) {
    CTX_BEGIN(c);
    if (c->g.nz < 2) {
        cism_set_error("assemble_stiffness_matrix_3d needs a 3D context (nz >= 2)");
        return CISM_ERR_ARG;
    }
    CISM_TRY(check_options(whichapprox, whichefvs));
    CISM_TRY(upload_geometry(c, sigma, active_cell, xVertex, yVertex, stagusrf, stagthck, flwafact));
    CISM_TRY(upload_vec(c, c->x, uvel, vvel));
    HIP_TRY(hipMemsetAsync(c->sc, 0, sizeof(DevScal), c->stream));
    CISM_TRY(assemble_device(c, whichapprox, whichefvs, efvs_constant, 0, 0.0, 0.0));
    CISM_TRY(check_flags(c));
    CISM_TRY(down(c, efvs, c->efvs, (size_t)c->g.ne));
    return download_matrix(c, Auu, Auv, Avu, Avv);
}

__global__ void k_basal_local(GridDev g, const int32_t *__restrict__ av, const double *__restrict__ beta,
                              double scale, double *__restrict__ A) {
    const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (long long)g.sx * g.sy || !av[col]) return;
    const long long q = (g.nz - 1) + (long long)g.nz * col;
    A[aidx(0, CISM_DIAG, q)] = A[aidx(0, CISM_DIAG, q)] + scale * beta[col];
    A[aidx(3, CISM_DIAG, q)] = A[aidx(3, CISM_DIAG, q)] + scale * beta[col];
}

extern "C" int cism_basal_sliding_bc_3d(cism_ctx *c, double dx, double dy,
                                        const int32_t *active_vertex, const double *beta,
                                        int whichassemble_beta, double *Auu, double *Avv) {
    CTX_BEGIN(c);
    if (whichassemble_beta != 1) {
        cism_set_error("cism_basal_sliding_bc_3d is the local assembly (which_ho_assemble_beta = 1); "
                       "cism_basal_sliding_bc takes the standard one");
        return CISM_ERR_UNSUPPORTED;
    }
    const size_t ncol = (size_t)c->g.sx * c->g.sy;
    CISM_TRY(up(c, c->av, active_vertex, ncol));
    CISM_TRY(up(c, c->beta, beta, ncol));
    CISM_TRY(ensure_stage(c));
    CISM_TRY(upload_block(c, 0, Auu));
    HIP_TRY(hipStreamSynchronize(c->stream));
    CISM_TRY(upload_block(c, 3, Avv));
    c->a64s_valid = 0;
    hipLaunchKernelGGL(k_basal_local, dim3(nblocks_for((long long)ncol)), dim3(256), 0, c->stream,
                       c->g, c->av, c->beta, dx * dy / CISM_VOL0, c->A);
    HIP_TRY(hipGetLastError());
    CISM_TRY(download_block(c, 0, Auu));
    CISM_TRY(download_block(c, 3, Avv));
    return CISM_OK;
}

// basal_sliding_bc (velo_higher:8125-8381) with the reference's arguments, both assemblies,
// on the basal layer of the full 3D blocks (standard: velo_higher:8236-8363)
extern "C" int cism_basal_sliding_bc(cism_ctx *c, double dx, double dy, const int32_t *active_cell,
                                     const int32_t *active_vertex, const double *beta,
                                     const double *xVertex, const double *yVertex,
                                     int whichassemble_beta, double *Auu, double *Avv) {
    CTX_BEGIN(c);
    if (whichassemble_beta == 1) return cism_basal_sliding_bc_3d(c, dx, dy, active_vertex, beta, 1, Auu, Avv);
    if (whichassemble_beta != 0) {
        cism_set_error("whichassemble_beta must be 0 (standard) or 1 (local)");
        return CISM_ERR_ARG;
    }
    const GridDev &g = c->g;
    const size_t ncol = (size_t)g.sx * g.sy;
    CISM_TRY(up(c, c->ac, active_cell, (size_t)g.nx * g.ny));
    CISM_TRY(up(c, c->beta, beta, ncol));
    CISM_TRY(up(c, c->xV, xVertex, ncol));
    CISM_TRY(up(c, c->yV, yVertex, ncol));
    CISM_TRY(ensure_stage(c));
    CISM_TRY(upload_block(c, 0, Auu));
    HIP_TRY(hipStreamSynchronize(c->stream));
    CISM_TRY(upload_block(c, 3, Avv));
    CISM_TRY(basal_standard_device(c));
    CISM_TRY(download_block(c, 0, Auu));
    CISM_TRY(download_block(c, 3, Avv));
    return CISM_OK;
}

extern "C" int cism_dirichlet_boundary_conditions_3d(cism_ctx *c, const int32_t *active_vertex,
                                                     const int32_t *umask_dirichlet,
                                                     const int32_t *vmask_dirichlet,
                                                     const double *uvel, const double *vvel,
                                                     double *Auu, double *Auv, double *Avu,
                                                     double *Avv, double *bu, double *bv) {
    CTX_BEGIN(c);
    CISM_TRY(up(c, c->av, active_vertex, (size_t)c->g.sx * c->g.sy));
    CISM_TRY(upload_dirichlet_masks(c, umask_dirichlet, vmask_dirichlet, (size_t)c->g.nn));
    CISM_TRY(upload_vec(c, c->x, uvel, vvel));
    CISM_TRY(upload_vec(c, c->b, bu, bv));
    CISM_TRY(upload_matrix(c, Auu, Auv, Avu, Avv));
    CISM_TRY(dirichlet_device(c));
    CISM_TRY(download_matrix(c, Auu, Auv, Avu, Avv));
    return download_vec(c, c->b, bu, bv);
}

extern "C" int cism_check_symmetry_assembled_matrix_3d(cism_ctx *c, const int32_t *active_vertex,
                                                       double *Auu, double *Auv, double *Avu,
                                                       double *Avv, int64_t *nwarn) {
    CTX_BEGIN(c);
    CISM_TRY(up(c, c->av, active_vertex, (size_t)c->g.sx * c->g.sy));
    CISM_TRY(upload_matrix(c, Auu, Auv, Avu, Avv));
    CISM_TRY(symmetry_device(c, nullptr));
    CISM_TRY(sync_scalars(c));
    if (nwarn) *nwarn = (int64_t)c->sc_host->sums[0];
    return download_matrix(c, Auu, Auv, Avu, Avv);
}

extern "C" int cism_count_nonzeros_3d(cism_ctx *c, const double *Auu, const double *Auv,
                                      const double *Avu, const double *Avv,
                                      const int32_t *active_vertex, int64_t *nNonzeros) {
    CTX_BEGIN(c);
    CISM_TRY(up(c, c->av, active_vertex, (size_t)c->g.sx * c->g.sy));
    CISM_TRY(upload_matrix(c, Auu, Auv, Avu, Avv));
    CISM_TRY(count_nonzeros_device(c));
    CISM_TRY(sync_scalars(c));
    if (nNonzeros) *nNonzeros = (int64_t)llround(c->sc_host->sums[0]);
    return CISM_OK;
}

extern "C" int cism_compute_residual_vector_3d(cism_ctx *c, const int32_t *active_vertex,
                                               const double *Auu, const double *Auv,
                                               const double *Avu, const double *Avv,
                                               const double *bu, const double *bv,
                                               const double *uvel, const double *vvel,
                                               double *resid_u, double *resid_v, double *L2_norm,
                                               double *L2_norm_relative) {
    CTX_BEGIN(c);
    CISM_TRY(up(c, c->av, active_vertex, (size_t)c->g.sx * c->g.sy));
    CISM_TRY(upload_matrix(c, Auu, Auv, Avu, Avv));
    CISM_TRY(upload_vec(c, c->b, bu, bv));
    CISM_TRY(upload_vec(c, c->x, uvel, vvel));
    CISM_TRY(residual_vector_device(c, c->t, nullptr));
    CISM_TRY(sync_scalars(c));
    const double L2 = sqrt(c->sc_host->sums[0]), rhs = sqrt(c->sc_host->sums[1]);
    if (L2_norm) *L2_norm = L2;
    if (L2_norm_relative) *L2_norm_relative = rhs > 0.0 ? L2 / rhs : 0.0;
    return download_vec(c, c->t, resid_u, resid_v);
}

extern "C" int cism_compute_residual_velocity_3d(cism_ctx *c, int whichresid, const double *uvel,
                                                 const double *vvel, const double *usav,
                                                 const double *vsav, double *resid_velo) {
    CTX_BEGIN(c);
    CISM_TRY(upload_vec(c, c->x, uvel, vvel));
    CISM_TRY(upload_vec(c, c->usav, usav, vsav));
    CISM_TRY(residual_velocity_device(c, whichresid));
    CISM_TRY(sync_scalars(c));
    double r = c->sc_host->sums[0];
    if (whichresid == 2) r = c->sc_host->sums[1] > 0 ? r / c->sc_host->sums[1] : 0.0;
    if (resid_velo) *resid_velo = r;
    return CISM_OK;
}

extern "C" int cism_load_vector_gravity(cism_ctx *c, const double *sigma, const double *stagwbndsigma,
                                        double dx, double dy, const int32_t *active_cell,
                                        const int32_t *active_vertex, const double *xVertex,
                                        const double *yVertex, const double *stagusrf,
                                        const double *stagthck, const double *dusrf_dx,
                                        const double *dusrf_dy, int whichassemble_taud,
                                        double *loadu, double *loadv) {
    CTX_BEGIN(c);
    const size_t ncol = (size_t)c->g.sx * c->g.sy;
    CISM_TRY(up(c, c->sigma, sigma, (size_t)c->g.nz));
    CISM_TRY(up(c, c->wsig, stagwbndsigma, (size_t)c->g.nz + 1));
    CISM_TRY(up(c, c->ac, active_cell, (size_t)c->g.nx * c->g.ny));
    CISM_TRY(up(c, c->av, active_vertex, ncol));
    CISM_TRY(up(c, c->xV, xVertex, ncol));
    CISM_TRY(up(c, c->yV, yVertex, ncol));
    CISM_TRY(up(c, c->susrf, stagusrf, ncol));
    CISM_TRY(up(c, c->sthck, stagthck, ncol));
    CISM_TRY(up(c, c->dsdx, dusrf_dx, ncol));
    CISM_TRY(up(c, c->dsdy, dusrf_dy, ncol));
    CISM_TRY(upload_vec(c, c->load, loadu, loadv));
    CISM_TRY(load_gravity_device(c, whichassemble_taud, dx, dy));
    return download_vec(c, c->load, loadu, loadv);
}

extern "C" int cism_load_vector_lateral_bc(cism_ctx *c, const double *sigma, const int32_t *land_mask,
                                           const int32_t *ocean_mask,
                                           const int32_t *calving_front_mask,
                                           const int32_t *active_cell, const double *xVertex,
                                           const double *yVertex, const double *stagusrf,
                                           const double *stagthck, double *loadu, double *loadv) {
    CTX_BEGIN(c);
    (void)land_mask;
    const size_t ncol = (size_t)c->g.sx * c->g.sy, ncell = (size_t)c->g.nx * c->g.ny;
    CISM_TRY(up(c, c->sigma, sigma, (size_t)c->g.nz));
    CISM_TRY(up(c, c->ac, active_cell, ncell));
    CISM_TRY(up(c, c->xV, xVertex, ncol));
    CISM_TRY(up(c, c->yV, yVertex, ncol));
    CISM_TRY(up(c, c->susrf, stagusrf, ncol));
    CISM_TRY(up(c, c->sthck, stagthck, ncol));
    int32_t *masks = nullptr;
    HIP_TRY(hipMalloc(&masks, sizeof(int32_t) * 2 * ncell));
    int st = up(c, masks, ocean_mask, ncell);
    if (!st) st = up(c, masks + ncell, calving_front_mask, ncell);
    if (!st) st = upload_vec(c, c->load, loadu, loadv);
    if (!st) st = load_lateral_device(c, masks, masks + ncell);
    if (!st) st = download_vec(c, c->load, loadu, loadv);
    (void)hipFree(masks);
    return st;
}

extern "C" int cism_staggered_parallel_halo(cism_ctx *c, int lead, double *a) {
    CTX_BEGIN(c);
    const GridDev &g = c->g;
    if (lead == g.nz) {
        // one (nz, nx-1, ny-1) array: carried as the u component of the temp vector
        CISM_TRY(upload_vec(c, c->t, a, a));
        CISM_TRY(launch_halo_vec(c, c->t));
        double *tmp = (double *)malloc(sizeof(double) * g.nn);
        if (!tmp) return CISM_ERR_NOMEM;
        int st = download_vec(c, c->t, a, tmp);
        free(tmp);
        return st;
    }
    if (lead == 27 * g.nz) {
        CISM_TRY(ensure_stage(c));
        CISM_TRY(upload_block(c, 0, a));
        CISM_TRY(launch_halo_mat(c, c->A));
        return download_block(c, 0, a);
    }
    cism_set_error("halo: lead must be nz or 27*nz (got %d)", lead);
    return CISM_ERR_ARG;
}

// ---------------------------------------------------------------------------
// device-resident Picard loop (glissade_velo_higher.F90:2003-2016, 2206-3727)
// ---------------------------------------------------------------------------
extern "C" void cism_solve_opts_default(cism_solve_opts *o) {
    memset(o, 0, sizeof *o);
    o->whichbabc = 4;
    o->whichefvs = 2;
    o->whichresid = 3;
    o->whichsparse = 3;
    o->whichprecond = 2;
    o->whichapprox = 2;
    o->maxiter_nonlinear = 100;
    o->precision = 0;
    o->efvs_constant = 2336041.42829;
    o->dx = o->dy = 1000.0;
    o->inner_rtol = 1.0e-3;
    o->adaptive_gamma = 0.05;
    o->anderson_m = 0;
    o->whichassemble_beta = 1;
}

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// The Picard loop on the device (velo_higher:2206-3727), inputs already resident: the
// 3D Blatter-Pattyn system, or with dim2 the SSA system of a one-level context.
// t_dev0 > 0: the time the inputs became resident (seconds_device counts from there),
// else from this call, after the stream has drained
static int picard_device(cism_ctx *c, const cism_solve_opts *o, int dim2, double t0, double *uvel,
                         double *vvel, double *efvs, double *bu, double *bv, double *resid_u,
                         double *resid_v, cism_solve_stats *stats, double t_dev0 = 0.0) {
    const GridDev &g = c->g;
    if (o->whichassemble_beta != 0 && o->whichassemble_beta != 1) {
        cism_set_error("whichassemble_beta must be 0 (standard) or 1 (local)");
        return CISM_ERR_ARG;
    }
    if (o->anderson_m < 0 || o->anderson_m > AA_MAX) {
        cism_set_error("anderson_m must be 0..%d", AA_MAX);
        return CISM_ERR_ARG;
    }
    if (o->skip_final_residual < 0 || o->skip_final_residual > 2) {
        cism_set_error("skip_final_residual must be 0, 1 or 2");
        return CISM_ERR_ARG;
    }
    // the depth-integrated 2D solvers L1L2 (3) and DIVA (4); with a no-slip bed the SSA and
    // L1L2 velocities are zero and the loop ends at once (velo_higher:3234-3245)
    const bool depth = dim2 && (o->whichapprox == 3 || o->whichapprox == 4);
    const bool zero_2d = dim2 && o->whichbabc == 6 && o->whichapprox != 4;
    anderson_reset(c);
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (!(t_dev0 > 0.0)) t_dev0 = now_s();   // inputs resident from here on
    double crit = 1.0e10, target = 1.0e-12;
    double L2 = 1.0e20, L2rel = 1.0e20, resid_velo = 1.0;
    const bool rv_after = o->whichresid == 3 || o->whichresid == 4;
    bool b_loaded = false;   // c->b holds the load vector with its halo (3D, no prescribed node)
    int counter = 0, lin_total = 0, st = CISM_OK;
    double t_asm = 0, t_lin = 0;
    bool zero_exit = false, stop_hit = false;
    cism_solve_stats S;
    memset(&S, 0, sizeof S);
    // the mixed-precision path (precision 1) falls back to the plain FP64 solver for the
    // rest of the solve once its refinement stalls; it also rejects an initial guess
    // whose Picard residual exceeds the zero guess's (|b|) and starts from zero instead
    int prec = o->precision;
    bool guess_checked = prec != 1;
    TimerRange r_loop("glissade_vhs_nonlinear_loop");
    while (crit >= target && counter < o->maxiter_nonlinear) {
        counter++;
        HIP_TRY(hipMemcpyAsync(c->usav, c->x, sizeof(double) * 2 * g.npad, hipMemcpyDeviceToDevice,
                               c->stream));
        if (c->bp_on) {   // velo_higher:2237-2243, 2502-2517: DIVA's basal velocity is the 3D bed level
            if (depth && o->whichapprox == 4)
                CISM_TRY(calcbeta_from_device(c, c->dep_u3 + (c->dep_nz3 - 1), c->dep_v3 + (c->dep_nz3 - 1),
                                              c->dep_nz3));
            else
                CISM_TRY(calcbeta_device(c));
        }
        double ta = now_s();
        bool fuse32 = false;
        {
            // basal friction (velo_higher:3039-3052): the local assembly inside the gather, the
            // standard one after it; the 2D branch applies it for every which_ho_babc
            // (:2832-2852), DIVA with beta_eff
            const bool basal = dim2 || o->whichbabc != 6, standard = o->whichassemble_beta == 0;
            // The 3D gather sums a coefficient and its transpose over the same elements in the
            // same order from symmetric element matrices, so its matrix is symmetric as it
            // stands; when no Dirichlet row or separate basal pass changes it afterwards, the
            // mixed path takes the FP32 matrix and the nonzero count from the gather and
            // skips the symmetry pass (which would find nothing to repair)
            fuse32 = prec == 1 && !dim2 && !c->dirichlet_any && !(basal && standard);
            {
                TimerRange r_asm(dim2 ? "glissade_assemble_2d" : "glissade_assemble_3d", c->stream, true);
                if (depth)
                    CISM_TRY(assemble_depth_device(c, o->whichapprox, o->whichefvs, o->efvs_constant,
                                                   basal && !standard, o->whichbabc == 6, o->dx, o->dy));
                else if (dim2)
                    CISM_TRY(assemble2d_device(c, o->whichefvs, o->efvs_constant, basal && !standard, o->dx,
                                               o->dy));
                else {
                    if (fuse32) {
                        CISM_TRY(ensure_a32(c));
                        CISM_TRY(launch_halo_vec(c, c->x));   // the gather's residual reads x's halo
                    }
                    // with skip_final_residual the FP64 matrix is read at most once per
                    // Picard iteration, by a refinement residual that the element matrices
                    // give as well (residual_from_elements): its rows are not written
                    CISM_TRY(assemble_device(c, o->whichapprox, o->whichefvs, o->efvs_constant,
                                             basal && !standard, o->dx, o->dy, fuse32 ? c->A32 : nullptr,
                                             fuse32 ? o->skip_final_residual == 0 : 1));
                }
                if (basal && standard)
                    CISM_TRY(basal_standard_device(c, depth ? depth_basal_beta(c, o->whichapprox) : nullptr));
            }
            // b = load, as the Dirichlet update leaves it; without prescribed nodes in the 3D
            // loop nothing changes b after the first Picard iteration's copy and its halo
            const bool b_fresh = dim2 || c->dirichlet_any || !b_loaded;
            if (b_fresh)
                HIP_TRY(hipMemcpyAsync(c->b, c->load, sizeof(double) * 2 * g.npad, hipMemcpyDeviceToDevice,
                                       c->stream));
            {
                TimerRange r_d(dim2 ? "glissade_dirichlet_2d" : "glissade_dirichlet_3d", c->stream, true);
                CISM_TRY(dirichlet_device(c));
            }
            {
                TimerRange r_h("glissade_halo_Axxs", c->stream, true);
                if (c->a64_valid) CISM_TRY(launch_halo_mat(c, c->A));
                if (b_fresh) CISM_TRY(launch_halo_vec(c, c->b));
                b_loaded = true;
            }
            TimerRange r_s("glissade_chk_symmetry");   // ends at the synchronization below
            if (fuse32) {
                CISM_TRY(launch_a32_halo(c));
                CISM_TRY(launch_halo_vec(c, c->r));
            } else {
                // the mixed-precision path gets its FP32 symmetric matrix from the same pass,
                // the FP64 one its FP64 symmetric copy (used when no pair stays unrepaired)
                if (prec == 1) CISM_TRY(ensure_a32(c));
                double *as64 = prec == 0 && c->sym64_on && ensure_a64s(c) == CISM_OK ? c->A64s : nullptr;
                CISM_TRY(symmetry_device(c, prec == 1 ? c->A32 : nullptr, as64));
            }
            CISM_TRY(sync_scalars(c));
            sym64_commit(c);
        }
        if (c->sc_host->singular_flag) {
            cism_set_error("Jacobian matrix is singular");
            return CISM_ERR_SINGULAR;
        }
        if (c->bp_on && c->sc_host->beta_bad) {
            cism_set_error("Invalid beta value in calcbeta (negative or NaN)");
            return CISM_ERR_NAN;
        }
        const double nnz = c->sc_host->sums[fuse32 ? 4 : 1];
        t_asm += now_s() - ta;
        if (nnz == 0.0) {
            zero_exit = true;
            L2 = 0.0;
            break;
        }
        if (zero_2d) {
            HIP_TRY(hipMemsetAsync(c->x, 0, sizeof(double) * 2 * g.npad, c->stream));
            HIP_TRY(hipMemsetAsync(c->t, 0, sizeof(double) * 2 * g.npad, c->stream));
            L2 = 0.0;
            L2rel = 0.0;
        } else {
            // the mixed path takes the first residual of its refinement from this pass
            const bool share_r = prec == 1;
            if (!fuse32) {   // the fused gather has formed it already
                TimerRange r_r("glissade_resid_vec");
                if (share_r) CISM_TRY(launch_halo_vec(c, c->x));
                CISM_TRY(residual_vector_device(c, c->t, share_r ? c->r : nullptr));
                CISM_TRY(sync_scalars(c));
            }
            L2 = sqrt(c->sc_host->sums[0]);
            const double rhs = sqrt(c->sc_host->sums[1]);
            L2rel = rhs > 0.0 ? L2 / rhs : 0.0;
            if (!guess_checked) {
                guess_checked = true;
                // the guess is kept if it is closer than the zero guess to the solution of this
                // first Picard system in that system's energy norm or as close, i.e. J(x) =
                // x.A x / 2 - b.x <= J(0) = 0 (the zero guess itself passes); otherwise (a NaN
                // included) it is dropped and this Picard iteration starts over from zero.  (The residual norm is no such test: a guess
                // within 10 % of the solution can have a residual far above |b|, the zero
                // guess's, because the operator amplifies its rough error components.)
                double twoJ = 0.0;
                CISM_TRY(guess_energy_device(c, c->t, &twoJ));
                if (!(twoJ <= 0.0)) {
                    HIP_TRY(hipMemsetAsync(c->x, 0, sizeof(double) * 2 * g.npad, c->stream));
                    S.guess_rejected = 1;
                    counter--;
                    t_asm += now_s() - ta;
                    continue;
                }
            }
            double tl = now_s(), err = 0.0;
            int niters = 0;
            // the reference path solves every linear system to the fixed tolerance
            double lin_tol = c->tolerance;
            if (o->precision == 1 || o->adaptive_linear) {
                // The Picard test on the absolute L2 norm (target 1e-4, velo_higher:2007) ends
                // the loop only once |A(x) x - b| < 1e-4, and near the fixed point that norm is
                // the last linear solve's residual.  Relative to |b| (1e4 .. 1e6 on ice sheets)
                // the target is 1e-8 .. 1e-10, tighter than the linear tolerance: the fast path
                // solves to half the target, which its FP64 refinement reaches, so that its
                // loop can end by the reference's own test.
                if (o->whichresid == 3 && rhs > 0.0) {
                    const double t_abs = o->target_override > 0 ? o->target_override : 1.0e-4;
                    lin_tol = fmin(lin_tol, 0.5 * t_abs / rhs);
                }
            }
            if (o->adaptive_linear) {
                // inexact Picard: no tighter than gamma times the current relative residual
                const double gam = o->adaptive_gamma > 0 ? o->adaptive_gamma : 0.05;
                lin_tol = fmin(1.0e-2, fmax(lin_tol, gam * L2rel));
            }
            roctxRangePushA("glissade_pcg_slv_struct");   // popped once the linear solve has ended
            if (prec == 1) {
                int n64 = 0, n32 = 0, stalled = 0, unconv = 0;
                const int pc = o->whichprecond == 2 && o->coarse ? 3 : o->whichprecond;
                st = solver_mixed(c, pc, c->maxiters, lin_tol, o->inner_rtol, &err, &niters,
                                  &n64, &n32, 1, &stalled, &unconv, o->skip_final_residual);
                S.unconverged_linear += unconv;
                S.fp64_matvecs += n64;
                S.fp32_matvecs += n32;
                S.refinement_steps += n64;
                if (!st && stalled) {
                    // the FP64 fallback: this Picard iteration's system again from its start
                    // velocity, with the reference's solver; the rest of the solve stays on it
                    S.fallbacks++;
                    prec = 0;
                    CISM_TRY(anderson_reset(c));
                    HIP_TRY(hipMemcpyAsync(c->x, c->usav, sizeof(double) * 2 * g.npad,
                                           hipMemcpyDeviceToDevice, c->stream));
                    if (!c->a64_valid) {
                        // the mixed path kept no FP64 rows: this Picard iteration's matrix
                        // again (the same velocity), as the reference path forms it
                        CISM_TRY(assemble_device(c, o->whichapprox, o->whichefvs, o->efvs_constant,
                                                 c->asm_add_basal, o->dx, o->dy));
                        CISM_TRY(launch_halo_mat(c, c->A));
                        CISM_TRY(symmetry_device(
                            c, nullptr, c->sym64_on && ensure_a64s(c) == CISM_OK ? c->A64s : nullptr));
                        CISM_TRY(sync_scalars(c));
                        sym64_commit(c);
                    }
                    int n2 = 0;
                    st = solver_pcg64(c, o->whichsparse == 3, o->whichprecond, c->maxiters, lin_tol, &err, &n2);
                    niters += n2;
                }
            } else {
                st = solver_pcg64(c, o->whichsparse == 3, o->whichprecond, c->maxiters, lin_tol, &err,
                                  &niters);
                S.unconverged_linear += !st && !(err < lin_tol);
            }
            roctxRangePop();
            if (st) return st;
            t_lin += now_s() - tl;
            lin_total += niters;
        }
        CISM_TRY(launch_halo_vec(c, c->x));
        if (prec == 1 && o->anderson_m > 0) {
            double fn = 0.0;
            CISM_TRY(anderson_step(c, o->anderson_m, &fn));
            CISM_TRY(launch_halo_vec(c, c->x));
        }
        // the loop's test is on L2 (whichresid 3, 4) or on the velocity residual; with the
        // former, the velocity residual is a report only and is formed once, after the loop
        if (!rv_after) {
            CISM_TRY(residual_velocity_device(c, o->whichresid));
            CISM_TRY(sync_scalars(c));
            resid_velo = c->sc_host->sums[0];
        }
        if (depth && o->whichapprox == 4) CISM_TRY(diva_post_solve_device(c));
        if (o->whichresid == 2) resid_velo = c->sc_host->sums[1] > 0 ? resid_velo / c->sc_host->sums[1] : 0.0;
        if (o->whichresid == 3) { crit = L2; target = 1.0e-4; }
        else if (o->whichresid == 4) { crit = L2rel; target = 1.0e-7; }
        else { crit = resid_velo; target = 1.0e-4; }
        if (o->target_override > 0) target = o->target_override;
        if (o->stop_rel_err > 0 && c->xref) {
            double e = 0.0;
            CISM_TRY(rel_err_device(c, &e));
            S.final_rel_err = e;
            if (e <= o->stop_rel_err) {
                stop_hit = true;
                break;
            }
        }
    }
    // converged_soln of the reference (velo_higher:3731-3745): the loop ended before
    // maxiter_nonlinear; a solve stopped on its error target, or with nothing to solve,
    // has converged too
    S.converged = counter < o->maxiter_nonlinear || stop_hit || zero_exit;
    if (rv_after && counter > 0 && !zero_exit) {
        // the last iteration's x and its start usav are still in place
        CISM_TRY(residual_velocity_device(c, o->whichresid));
        CISM_TRY(sync_scalars(c));
        resid_velo = c->sc_host->sums[0];
    }
    if (zero_exit) {
        HIP_TRY(hipMemsetAsync(c->x, 0, sizeof(double) * 2 * g.npad, c->stream));
        HIP_TRY(hipMemsetAsync(c->b, 0, sizeof(double) * 2 * g.npad, c->stream));
        HIP_TRY(hipMemsetAsync(c->t, 0, sizeof(double) * 2 * g.npad, c->stream));
    }
    if (depth && o->whichapprox == 3) {
        CISM_TRY(l1l2_velocity_device(c, o->dx, o->dy));
        CISM_TRY(sync_scalars(c));
        if (c->sc_host->singular_flag) {
            cism_set_error("Jacobian matrix is singular");
            return CISM_ERR_SINGULAR;
        }
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    S.seconds_device = now_s() - t_dev0;
    CISM_TRY(download_vec(c, c->x, uvel, vvel));
    if (efvs && dim2) {
        CISM_TRY(ssa_efvs_layers_device(c, c->ssa_flw3));
        CISM_TRY(down(c, efvs, c->ssa_flw3, (size_t)(c->ssa_nz3 - 1) * g.nx * g.ny));
    } else if (efvs) {
        CISM_TRY(down(c, efvs, c->efvs, (size_t)g.ne));
    }
    if (bu && bv) CISM_TRY(download_vec(c, c->b, bu, bv));
    if (resid_u && resid_v) CISM_TRY(download_vec(c, c->t, resid_u, resid_v));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->xref) {
        double e = 0.0;
        CISM_TRY(rel_err_device(c, &e));
        S.final_rel_err = e;
    } else {
        S.final_rel_err = -1.0;
    }
    S.picard_iters = counter;
    S.linear_iters = lin_total;
    S.final_L2_norm = L2;
    S.final_resid_velo = resid_velo;
    S.seconds_assembly = t_asm;
    S.seconds_linear = t_lin;
    S.seconds_total = now_s() - t0;
    S.status = st;
    if (stats) *stats = S;
    return CISM_OK;
}

extern "C" int cism_velo_higher_solve_3d(cism_ctx *c, const cism_solve_opts *o, const double *sigma,
                                         const int32_t *active_cell, const int32_t *active_vertex,
                                         const double *xVertex, const double *yVertex,
                                         const double *stagusrf, const double *stagthck,
                                         const double *flwafact, const double *beta_internal,
                                         const int32_t *umask_dirichlet,
                                         const int32_t *vmask_dirichlet, const double *loadu,
                                         const double *loadv, double *uvel, double *vvel,
                                         double *efvs, double *bu, double *bv, double *resid_u,
                                         double *resid_v, cism_solve_stats *stats) {
    CTX_BEGIN(c);
    if (!o) return CISM_ERR_ARG;
    CISM_TRY(check_options(o->whichapprox, o->whichefvs));
    if (o->whichsparse != 2 && o->whichsparse != 3) {
        cism_set_error("whichsparse must be 2 or 3 (native PCG)");
        return CISM_ERR_UNSUPPORTED;
    }
    const double t0 = now_s();
    const GridDev &g = c->g;
    if (g.nz < 2) {
        cism_set_error("the 3D solve needs nz >= 2 (a one-level context runs cism_velo_higher_solve_2d)");
        return CISM_ERR_ARG;
    }
    const size_t ncol = (size_t)g.sx * g.sy;
    CISM_TRY(upload_geometry(c, sigma, active_cell, xVertex, yVertex, stagusrf, stagthck, flwafact));
    CISM_TRY(up(c, c->av, active_vertex, ncol));
    CISM_TRY(up(c, c->beta, beta_internal, ncol));
    CISM_TRY(upload_dirichlet_masks(c, umask_dirichlet, vmask_dirichlet, (size_t)g.nn));
    CISM_TRY(upload_vec(c, c->load, loadu, loadv));
    if (o->initial_guess == 0) {
        CISM_TRY(upload_vec(c, c->x, uvel, vvel));
    } else if (o->initial_guess != 2) {
        cism_set_error("initial_guess must be 0 (uvel, vvel) or 2 (the context's velocity) here");
        return CISM_ERR_ARG;
    }
    HIP_TRY(hipMemsetAsync(c->sc, 0, sizeof(DevScal), c->stream));
    return picard_device(c, o, 0, t0, uvel, vvel, efvs, bu, bv, resid_u, resid_v, stats);
}

// the Dirichlet masks of glissade_velo_higher_solve (velo_higher:1269-1324) without
// kinbcmask or no-penetration boundaries: the bed level under no-slip, else none
__global__ __launch_bounds__(256) void k_noslip_masks(GridDev g, int noslip, int32_t *__restrict__ um,
                                                      int32_t *__restrict__ vm) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nn) return;
    const int m = noslip && (int)(q % g.nz) == g.nz - 1;
    um[q] = m;
    vm[q] = m;
}

// glissade_velo_higher_solve (velo_higher:654-3986, 3D branch) from the ice geometry:
// the inputs, grounded fraction, beta, loads and Dirichlet masks on the device, then the
// Picard loop; only the cell fields go in and the velocity (and efvs) come out.
extern "C" int cism_glissade_velo_higher_solve(cism_ctx *c, const cism_solve_opts *o,
                                               const cism_velo_setup *su, const double *sigma,
                                               const double *thck, const double *usrf, const double *topg,
                                               const double *flwa, const cism_basal_physics *bp,
                                               const double *beta_external, const double *effecpress_stag,
                                               double *uvel, double *vvel, double *efvs,
                                               cism_solve_stats *stats) {
    CTX_BEGIN(c);
    if (!o || !su || !sigma || !uvel || !vvel) {
        cism_set_error("glissade_velo_higher_solve: null options, setup, sigma or velocity");
        return CISM_ERR_ARG;
    }
    CISM_TRY(check_options(o->whichapprox, o->whichefvs));
    if (o->whichsparse != 2 && o->whichsparse != 3) {
        cism_set_error("whichsparse must be 2 or 3 (native PCG)");
        return CISM_ERR_UNSUPPORTED;
    }
    const GridDev &g = c->g;
    if (g.nz < 2) {
        cism_set_error("glissade_velo_higher_solve needs a 3D context (nz >= 2)");
        return CISM_ERR_ARG;
    }
    const bool noslip = o->whichbabc == 6;
    if (!noslip && !bp) {
        cism_set_error("glissade_velo_higher_solve: basal physics needed unless which_ho_babc = 6 (no-slip)");
        return CISM_ERR_ARG;
    }
    if (su->initial_guess < 0 || su->initial_guess > 2 || (su->initial_guess == 1 && !c->guess)) {
        cism_set_error("glissade_velo_higher_solve: initial_guess 0, 1 (after cism_learned_guess_load) or 2");
        return CISM_ERR_ARG;
    }
    if (su->whichassemble_taud != 0 && su->whichassemble_taud != 1) {
        cism_set_error("glissade_velo_higher_solve: whichassemble_taud must be 0 or 1");
        return CISM_ERR_UNSUPPORTED;
    }
    const double t0 = now_s();
    // sigma and stagwbndsigma (glide_setup.F90:355-363)
    std::vector<double> w(g.nz + 1);
    w[0] = 0.0;
    for (int k = 1; k < g.nz; ++k) w[k] = (sigma[k - 1] + sigma[k]) / 2.0;
    w[g.nz] = 1.0;
    // host arrays in: sigma, the cell fields, the basal fields (and the initial velocity)
    CISM_TRY(up(c, c->sigma, sigma, (size_t)g.nz));
    CISM_TRY(up(c, c->wsig, w.data(), (size_t)g.nz + 1));
    CISM_TRY(velo_inputs_upload(c, thck, usrf, topg, flwa));
    if (bp) CISM_TRY(basal_fields_upload(c, bp, beta_external, effecpress_stag));
    if (su->initial_guess == 0) CISM_TRY(upload_vec(c, c->x, uvel, vvel));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const double t_dev0 = now_s();   // inputs resident from here on
    roctxRangePushA("glissade_vhs_init");   // popped once the loop's inputs are set
    CISM_TRY(velo_inputs_compute(c, su->dx, su->dy, su->thklim, su->eus, su->max_slope));
    CISM_TRY(mask_planes_device(c));
    const long long ncell = (long long)g.nx * g.ny, ncol = (long long)g.sx * g.sy;
    if (bp) {
        CISM_TRY(basal_setup_device(c, c->planes, su->whichground, bp));
    } else {
        c->bp_on = 0;
        HIP_TRY(hipMemsetAsync(c->beta, 0, sizeof(double) * ncol, c->stream));
    }
    // gravity and lateral ice-ocean loads (velo_higher:2033-2111)
    HIP_TRY(hipMemsetAsync(c->load, 0, sizeof(double) * 2 * g.npad, c->stream));
    {
        TimerRange r_g("glissade_load_vector_gravity", c->stream, true);
        CISM_TRY(load_gravity_device(c, su->whichassemble_taud, su->dx, su->dy));
    }
    if (su->lateral_bc) {
        TimerRange r_l("glissade_load_vector_lateral_bc", c->stream, true);
        CISM_TRY(load_lateral_device(c, c->planes + 2 * ncell, c->planes + 4 * ncell));
    }
    hipLaunchKernelGGL(k_noslip_masks, dim3(nblocks_for(g.nn)), dim3(256), 0, c->stream, g, noslip ? 1 : 0,
                       c->umask, c->vmask);
    HIP_TRY(hipGetLastError());
    c->dirichlet_any = noslip ? 1 : 0;
    HIP_TRY(hipMemsetAsync(c->sc, 0, sizeof(DevScal), c->stream));
    if (su->initial_guess == 1) {
        // beta at zero velocity for the network's input (exact for laws that do not
        // depend on the velocity); the Picard loop recomputes it from its iterates
        HIP_TRY(hipMemsetAsync(c->x, 0, sizeof(double) * 2 * g.npad, c->stream));
        if (c->bp_on) CISM_TRY(calcbeta_device(c));
        CISM_TRY(learned_guess_device(c));
    }
    if (roctx_sync_enabled()) HIP_TRY(hipStreamSynchronize(c->stream));
    roctxRangePop();
    cism_solve_opts oo = *o;
    oo.initial_guess = 2;
    return picard_device(c, &oo, 0, t0, uvel, vvel, efvs, nullptr, nullptr, nullptr, nullptr, stats, t_dev0);
}

// The SSA branch of glissade_velo_higher_solve (solve_2d, velo_higher:2206-3767) on a
// one-level context: the 2D velocity starts from the bed level of uvel/vvel, the 2D
// load is the column sum of the 3D load (velo_higher:2123-2137), the Dirichlet masks are
// the bed level's, and every level of uvel/vvel takes the 2D solution at the end.
extern "C" int cism_velo_higher_solve_2d(cism_ctx *c, const cism_solve_opts *o, int nz3, const double *sigma,
                                         const int32_t *active_cell, const int32_t *active_vertex,
                                         const double *xVertex, const double *yVertex,
                                         const double *stagusrf, const double *stagthck,
                                         const double *flwafact, const double *beta_internal,
                                         const int32_t *umask_dirichlet,
                                         const int32_t *vmask_dirichlet, const double *loadu,
                                         const double *loadv, double *uvel, double *vvel,
                                         double *efvs, double *bu, double *bv, double *resid_u,
                                         double *resid_v, cism_solve_stats *stats) {
    CTX_BEGIN(c);
    (void)stagusrf;
    const GridDev &g = c->g;
    if (!o || g.nz != 1 || nz3 < 2 || nz3 > 64) {
        cism_set_error("velo_higher_solve_2d: needs a one-level context (nz = 1) and 2 <= nz3 <= 64");
        return CISM_ERR_ARG;
    }
    if (o->whichapprox != 1 || o->whichefvs < 0 || o->whichefvs > 2) {
        cism_set_error("velo_higher_solve_2d runs the SSA approximation (whichapprox = 1)");
        return CISM_ERR_UNSUPPORTED;
    }
    if (o->whichsparse != 2 && o->whichsparse != 3) {
        cism_set_error("whichsparse must be 2 or 3 (native PCG)");
        return CISM_ERR_UNSUPPORTED;
    }
    const double t0 = now_s();
    const size_t ncol = (size_t)g.sx * g.sy, ncell = (size_t)g.nx * g.ny;
    // host marshalling between the 3D arrays and the one-level system
    std::vector<double> u2(ncol), v2(ncol), lu2(ncol, 0.0), lv2(ncol, 0.0);
    std::vector<int32_t> um2(ncol), vm2(ncol);
    for (size_t col = 0; col < ncol; ++col) {
        const size_t b = (size_t)(nz3 - 1) + (size_t)nz3 * col;
        u2[col] = uvel[b];
        v2[col] = vvel[b];
        um2[col] = umask_dirichlet[b];
        vm2[col] = vmask_dirichlet[b];
        for (int k = 0; k < nz3; ++k) {
            lu2[col] = lu2[col] + loadu[k + (size_t)nz3 * col];
            lv2[col] = lv2[col] + loadv[k + (size_t)nz3 * col];
        }
    }
    CISM_TRY(ensure_ssa(c, nz3));
    CISM_TRY(up(c, c->ssa_sigma, sigma, (size_t)nz3));
    CISM_TRY(up(c, c->ssa_flw3, flwafact, (size_t)(nz3 - 1) * ncell));
    CISM_TRY(up(c, c->ac, active_cell, ncell));
    CISM_TRY(up(c, c->xV, xVertex, ncol));
    CISM_TRY(up(c, c->yV, yVertex, ncol));
    CISM_TRY(up(c, c->sthck, stagthck, ncol));
    CISM_TRY(up(c, c->av, active_vertex, ncol));
    CISM_TRY(up(c, c->beta, beta_internal, ncol));
    CISM_TRY(upload_dirichlet_masks(c, um2.data(), vm2.data(), ncol));
    CISM_TRY(upload_vec(c, c->load, lu2.data(), lv2.data()));
    CISM_TRY(upload_vec(c, c->x, u2.data(), v2.data()));
    CISM_TRY(ssa_prepare_device(c));
    HIP_TRY(hipMemsetAsync(c->sc, 0, sizeof(DevScal), c->stream));
    CISM_TRY(picard_device(c, o, 1, t0, u2.data(), v2.data(), efvs, bu, bv, resid_u, resid_v, stats));
    for (size_t col = 0; col < ncol; ++col)
        for (int k = 0; k < nz3; ++k) {
            uvel[k + (size_t)nz3 * col] = u2[col];
            vvel[k + (size_t)nz3 * col] = v2[col];
        }
    return CISM_OK;
}

// The L1L2 and DIVA branches of glissade_velo_higher_solve (solve_2d with which_ho_approx
// 3 or 4, velo_higher:1117-1173, 2206-3830) on a one-level context; see cism_gpu.h
extern "C" int cism_velo_higher_solve_depth(
    cism_ctx *c, const cism_solve_opts *o, int nz3, const double *sigma, const int32_t *ice_mask,
    const int32_t *ice_plus_land_mask, const int32_t *active_cell, const int32_t *active_vertex,
    const double *xVertex, const double *yVertex, const double *thck, const double *stagthck,
    const double *dusrf_dx, const double *dusrf_dy, const double *flwa, const double *flwafact,
    const double *beta_internal, const int32_t *umask_dirichlet, const int32_t *vmask_dirichlet,
    const double *loadu, const double *loadv, double *uvel, double *vvel, double *uvel_2d, double *vvel_2d,
    double *btractx, double *btracty, double *efvs, double *bu, double *bv, double *resid_u, double *resid_v,
    cism_solve_stats *stats) {
    CTX_BEGIN(c);
    const GridDev &g = c->g;
    if (!o || g.nz != 1 || nz3 < 2 || nz3 > 64) {
        cism_set_error("velo_higher_solve_depth: needs a one-level context (nz = 1) and 2 <= nz3 <= 64");
        return CISM_ERR_ARG;
    }
    const int approx = o->whichapprox;
    if ((approx != 3 && approx != 4) || o->whichefvs < 0 || o->whichefvs > 2) {
        cism_set_error("velo_higher_solve_depth runs L1L2 (whichapprox = 3) or DIVA (4)");
        return CISM_ERR_UNSUPPORTED;
    }
    if (o->whichsparse != 2 && o->whichsparse != 3) {
        cism_set_error("whichsparse must be 2 or 3 (native PCG)");
        return CISM_ERR_UNSUPPORTED;
    }
    if (approx == 4 && o->anderson_m != 0) {
        cism_set_error("DIVA carries its viscosity and basal traction across Picard iterations: anderson_m must be 0");
        return CISM_ERR_ARG;
    }
    if (!sigma || !active_cell || !active_vertex || !xVertex || !yVertex || !thck || !stagthck || !flwafact ||
        !beta_internal || !umask_dirichlet || !vmask_dirichlet || !loadu || !loadv || !uvel || !vvel ||
        (approx == 3 && (!ice_mask || !dusrf_dx || !dusrf_dy || !flwa)) ||
        (approx == 4 && (!ice_plus_land_mask || !uvel_2d || !vvel_2d || !efvs))) {
        cism_set_error("velo_higher_solve_depth: a required array is null");
        return CISM_ERR_ARG;
    }
    const double t0 = now_s();
    const size_t ncol = (size_t)g.sx * g.sy, ncell = (size_t)g.nx * g.ny, L = (size_t)nz3 - 1;
    // host marshalling: the 2D start velocity, bed-level masks and column-summed loads
    std::vector<double> u2(ncol), v2(ncol), lu2(ncol, 0.0), lv2(ncol, 0.0);
    std::vector<int32_t> um2(ncol), vm2(ncol);
    for (size_t col = 0; col < ncol; ++col) {
        const size_t b = (size_t)(nz3 - 1) + (size_t)nz3 * col;
        u2[col] = approx == 4 ? uvel_2d[col] : uvel[b];
        v2[col] = approx == 4 ? vvel_2d[col] : vvel[b];
        um2[col] = umask_dirichlet[b];
        vm2[col] = vmask_dirichlet[b];
        for (int k = 0; k < nz3; ++k) {
            lu2[col] = lu2[col] + loadu[k + (size_t)nz3 * col];
            lv2[col] = lv2[col] + loadv[k + (size_t)nz3 * col];
        }
    }
    CISM_TRY(ensure_depth(c, nz3));
    CISM_TRY(up(c, c->ssa_sigma, sigma, (size_t)nz3));
    CISM_TRY(up(c, c->ssa_flw3, flwafact, L * ncell));
    if (flwa) CISM_TRY(up(c, c->dep_flwa, flwa, L * ncell));
    CISM_TRY(up(c, c->dep_thck, thck, ncell));
    if (ice_mask) CISM_TRY(up(c, c->dep_ice, ice_mask, ncell));
    if (ice_plus_land_mask) CISM_TRY(up(c, c->dep_ipl, ice_plus_land_mask, ncell));
    CISM_TRY(up(c, c->ac, active_cell, ncell));
    CISM_TRY(up(c, c->xV, xVertex, ncol));
    CISM_TRY(up(c, c->yV, yVertex, ncol));
    CISM_TRY(up(c, c->sthck, stagthck, ncol));
    if (dusrf_dx) CISM_TRY(up(c, c->dsdx, dusrf_dx, ncol));
    if (dusrf_dy) CISM_TRY(up(c, c->dsdy, dusrf_dy, ncol));
    CISM_TRY(up(c, c->av, active_vertex, ncol));
    CISM_TRY(up(c, c->beta, beta_internal, ncol));
    CISM_TRY(up(c, c->dep_u3, uvel, (size_t)nz3 * ncol));
    CISM_TRY(up(c, c->dep_v3, vvel, (size_t)nz3 * ncol));
    if (approx == 4) {
        CISM_TRY(up(c, c->dep_efvs3, efvs, L * ncell));
        if (btractx && btracty) {
            CISM_TRY(up(c, c->dep_btr, btractx, ncol));
            CISM_TRY(up(c, c->dep_btr + ncol, btracty, ncol));
        } else {
            HIP_TRY(hipMemsetAsync(c->dep_btr, 0, sizeof(double) * 2 * ncol, c->stream));
        }
    }
    CISM_TRY(upload_dirichlet_masks(c, um2.data(), vm2.data(), ncol));
    CISM_TRY(upload_vec(c, c->load, lu2.data(), lv2.data()));
    CISM_TRY(upload_vec(c, c->x, u2.data(), v2.data()));
    CISM_TRY(depth_prepare_device(c, approx));
    HIP_TRY(hipMemsetAsync(c->sc, 0, sizeof(DevScal), c->stream));
    CISM_TRY(picard_device(c, o, 1, t0, u2.data(), v2.data(), nullptr, bu, bv, resid_u, resid_v, stats));
    if (uvel_2d && vvel_2d) {
        memcpy(uvel_2d, u2.data(), sizeof(double) * ncol);
        memcpy(vvel_2d, v2.data(), sizeof(double) * ncol);
    }
    CISM_TRY(down(c, uvel, c->dep_u3, (size_t)nz3 * ncol));
    CISM_TRY(down(c, vvel, c->dep_v3, (size_t)nz3 * ncol));
    if (efvs) CISM_TRY(down(c, efvs, c->dep_efvs3, L * ncell));
    if (approx == 4 && btractx && btracty) {
        CISM_TRY(down(c, btractx, c->dep_btr, ncol));
        CISM_TRY(down(c, btracty, c->dep_btr + ncol, ncol));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CISM_OK;
}

// ---------------------------------------------------------------------------
// reference velocity and kernel timing (measurement support)
// ---------------------------------------------------------------------------
extern "C" int cism_set_reference_velocity(cism_ctx *c, const double *uref, const double *vref) {
    CTX_BEGIN(c);
    if (!uref || !vref) {
        if (c->xref) HIP_TRY(hipFree(c->xref));
        c->xref = nullptr;
        return CISM_OK;
    }
    if (!c->xref) CISM_TRY(dalloc(&c->xref, 2 * (size_t)c->g.npad));
    CISM_TRY(upload_vec(c, c->xref, uref, vref));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CISM_OK;
}

__global__ __launch_bounds__(256) void k_scale_to_f32(long long n, const double *__restrict__ s,
                                                      float *__restrict__ out, int *__restrict__ bad) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double v = s[t];
    if (!(v > 0.0) || !(v < 1.0e30)) bad[0] = 1;
    out[t] = (float)v;
}

extern "C" int cism_set_fp64_symmetric(cism_ctx *c, int enable) {
    CTX_BEGIN(c);
    if (enable != 0 && enable != 1) {
        cism_set_error("cism_set_fp64_symmetric: enable must be 0 or 1");
        return CISM_ERR_ARG;
    }
    c->sym64_on = enable;
    if (!enable) c->a64s_valid = 0;
    return CISM_OK;
}

extern "C" int cism_set_precond_scale(cism_ctx *c, const double *scale_u, const double *scale_v) {
    CTX_BEGIN(c);
    if (!scale_u || !scale_v) {
        c->cscale_on = 0;
        coarse_invalidate(c);   // the coarse setup's SIA factors carry the scale
        return CISM_OK;
    }
    const long long ncol = (long long)c->g.sx * c->g.sy;
    if (!c->cscale) CISM_TRY(dalloc(&c->cscale, 2 * (size_t)ncol));
    double *st = c->vstage;   // 2*npad >= 2*ncol doubles
    HIP_TRY(hipMemcpyAsync(st, scale_u, sizeof(double) * ncol, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(st + ncol, scale_v, sizeof(double) * ncol, hipMemcpyHostToDevice, c->stream));
    int *bad = &c->sc->nan_flag;
    HIP_TRY(hipMemsetAsync(bad, 0, sizeof(int), c->stream));
    hipLaunchKernelGGL(k_scale_to_f32, dim3(nblocks_for(2 * ncol)), dim3(256), 0, c->stream, 2 * ncol, st,
                       c->cscale, bad);
    HIP_TRY(hipGetLastError());
    CISM_TRY(sync_scalars(c));
    if (c->sc_host->nan_flag) {
        c->cscale_on = 0;
        cism_set_error("preconditioner scales must be finite and positive");
        return CISM_ERR_ARG;
    }
    c->cscale_on = 1;
    coarse_invalidate(c);
    return CISM_OK;
}

extern "C" int cism_column_scale_target(cism_ctx *c, double *scale_u, double *scale_v) {
    CTX_BEGIN(c);
    if (!scale_u || !scale_v) return CISM_ERR_ARG;
    const long long ncol = (long long)c->g.sx * c->g.sy;
    double *out = c->vstage;
    CISM_TRY(column_scale_device(c, out));
    HIP_TRY(hipMemcpyAsync(scale_u, out, sizeof(double) * ncol, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(scale_v, out + ncol, sizeof(double) * ncol, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CISM_OK;
}

// ---------------------------------------------------------------------------
// post-solve diagnostics
// ---------------------------------------------------------------------------
static int download_tau(cism_ctx *c, const double *tau, double *txz, double *tyz, double *txx,
                        double *tyy, double *txy, double *teff) {
    const long long ne = c->g.ne;
    double *outs[6] = {txz, tyz, txx, tyy, txy, teff};
    for (int q = 0; q < 6; ++q)
        if (outs[q]) CISM_TRY(down(c, outs[q], tau + q * ne, (size_t)ne));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CISM_OK;
}

extern "C" int cism_compute_internal_stress(cism_ctx *c, const double *sigma, const int32_t *active_cell,
                                            const double *xVertex, const double *yVertex,
                                            const double *stagusrf, const double *stagthck,
                                            const double *flwafact, const double *efvs, int whichefvs,
                                            double efvs_constant, int whichapprox, const double *uvel,
                                            const double *vvel, double *tau_xz, double *tau_yz,
                                            double *tau_xx, double *tau_yy, double *tau_xy,
                                            double *tau_eff) {
    CTX_BEGIN(c);
    if (c->g.nz < 2) {
        cism_set_error("compute_internal_stress needs a 3D context (nz >= 2)");
        return CISM_ERR_ARG;
    }
    (void)efvs;   // used by the reference for L1L2 only, which this path does not run
    CISM_TRY(check_options(whichapprox, whichefvs));
    CISM_TRY(upload_geometry(c, sigma, active_cell, xVertex, yVertex, stagusrf, stagthck, flwafact));
    CISM_TRY(upload_vec(c, c->x, uvel, vvel));
    CISM_TRY(ensure_ke(c));
    HIP_TRY(hipMemsetAsync(c->sc, 0, sizeof(DevScal), c->stream));
    CISM_TRY(internal_stress_device(c, whichapprox, whichefvs, efvs_constant, c->Ke));
    CISM_TRY(check_flags(c));
    return download_tau(c, c->Ke, tau_xz, tau_yz, tau_xx, tau_yy, tau_xy, tau_eff);
}

extern "C" int cism_compute_basal_friction_heatflx(cism_ctx *c, const int32_t *active_cell,
                                                   const int32_t *active_vertex, const double *xVertex,
                                                   const double *yVertex, const double *ubas,
                                                   const double *vbas, const double *beta,
                                                   int whichassemble_bfric, double *bfricflx) {
    CTX_BEGIN(c);
    (void)xVertex, (void)yVertex;   // the reference's quadrature weights are 1: no geometry enters
    const GridDev &g = c->g;
    const size_t ncol = (size_t)g.sx * g.sy;
    CISM_TRY(up(c, c->ac, active_cell, (size_t)g.nx * g.ny));
    CISM_TRY(up(c, c->av, active_vertex, ncol));
    CISM_TRY(up(c, c->beta, beta, ncol));
    // interleave the basal velocity (u, v per vertex) into the staging buffer
    double *xb = c->vstage + 2 * ncol;   // vstage holds 2*npad >= 4*ncol doubles
    HIP_TRY(hipMemcpyAsync(c->vstage, ubas, sizeof(double) * ncol, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->vstage + ncol, vbas, sizeof(double) * ncol, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_interleave, dim3(nblocks_for(2 * (long long)ncol)), dim3(256), 0, c->stream,
                       (long long)ncol, c->vstage, xb);
    HIP_TRY(hipGetLastError());
    CISM_TRY(ensure_ke(c));
    CISM_TRY(basal_diagnostics_device(c, whichassemble_bfric, c->Ke, nullptr, xb, 2));
    CISM_TRY(down(c, bfricflx, c->Ke, (size_t)g.nx * g.ny));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CISM_OK;
}

extern "C" int cism_compute_basal_traction(cism_ctx *c, const double *beta, const double *uvel,
                                           const double *vvel, double *btractx, double *btracty) {
    CTX_BEGIN(c);
    const size_t ncol = (size_t)c->g.sx * c->g.sy;
    CISM_TRY(up(c, c->beta, beta, ncol));
    CISM_TRY(upload_vec(c, c->x, uvel, vvel));
    CISM_TRY(ensure_ke(c));
    CISM_TRY(basal_diagnostics_device(c, 1, nullptr, c->Ke));
    CISM_TRY(down(c, btractx, c->Ke, ncol));
    CISM_TRY(down(c, btracty, c->Ke + ncol, ncol));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CISM_OK;
}

extern "C" int cism_velo_diagnostics(cism_ctx *c, int whichapprox, int whichefvs, double efvs_constant,
                                     int whichassemble_bfric, double *tau_xz, double *tau_yz,
                                     double *tau_xx, double *tau_yy, double *tau_xy, double *tau_eff,
                                     double *bfricflx, double *btractx, double *btracty) {
    CTX_BEGIN(c);
    const GridDev &g = c->g;
    CISM_TRY(ensure_ke(c));
    if (tau_xz || tau_yz || tau_xx || tau_yy || tau_xy || tau_eff) {
        if (g.nz < 2) {
            cism_set_error("the stress components need a 3D context (nz >= 2)");
            return CISM_ERR_ARG;
        }
        CISM_TRY(check_options(whichapprox, whichefvs));
        HIP_TRY(hipMemsetAsync(c->sc, 0, sizeof(DevScal), c->stream));
        CISM_TRY(internal_stress_device(c, whichapprox, whichefvs, efvs_constant, c->Ke));
        CISM_TRY(check_flags(c));
        CISM_TRY(download_tau(c, c->Ke, tau_xz, tau_yz, tau_xx, tau_yy, tau_xy, tau_eff));
    }
    if (bfricflx || btractx || btracty) {
        const size_t ncol = (size_t)g.sx * g.sy;
        double *bf = c->Ke, *bt = c->Ke + (size_t)g.nx * g.ny;
        CISM_TRY(basal_diagnostics_device(c, whichassemble_bfric, bf, bt));
        if (bfricflx) CISM_TRY(down(c, bfricflx, bf, (size_t)g.nx * g.ny));
        if (btractx) CISM_TRY(down(c, btractx, bt, ncol));
        if (btracty) CISM_TRY(down(c, btracty, bt + ncol, ncol));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return CISM_OK;
}

// plain copy of n doubles: the calibration kernel for hardware byte counters
__global__ __launch_bounds__(256) void k_copy_calib(long long n, const double2 *__restrict__ a,
                                                    double2 *__restrict__ b) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n / 2) b[t] = a[t];
}

// ---------------------------------------------------------------------------
// coarse-correction diagnostics (coarse.hip)
// ---------------------------------------------------------------------------
extern "C" int cism_coarse_build(cism_ctx *c, const int32_t *active_vertex, const int32_t *umask_dirichlet,
                                 const int32_t *vmask_dirichlet, const double *Auu, const double *Auv,
                                 const double *Avu, const double *Avv, int *nlevels) {
    CTX_BEGIN(c);
    CISM_TRY(up(c, c->av, active_vertex, (size_t)c->g.sx * c->g.sy));
    CISM_TRY(upload_dirichlet_masks(c, umask_dirichlet, vmask_dirichlet, (size_t)c->g.nn));
    CISM_TRY(upload_matrix(c, Auu, Auv, Avu, Avv));
    c->a_zero_ok = 0;   // A is not the assembly's
    c->as_ok = 0;
    CISM_TRY(ensure_a32(c));
    CISM_TRY(symmetry_device(c, c->A32));
    CISM_TRY(coarse_setup(c, c->A32, c->F32, c->cscale_on ? c->cscale : nullptr));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (nlevels) *nlevels = c->coarse->nlev;
    return CISM_OK;
}

extern "C" int cism_coarse_level(cism_ctx *c, int level, int *ax, int *ay, double *blocks) {
    CTX_BEGIN(c);
    if (!c->coarse || !c->coarse->ok) {
        cism_set_error("no coarse hierarchy: build one (cism_coarse_build or a solve with coarse = 1)");
        return CISM_ERR_ARG;
    }
    if (level < 1 || level > c->coarse->nlev) {
        cism_set_error("coarse level %d out of 1..%d", level, c->coarse->nlev);
        return CISM_ERR_ARG;
    }
    const CoarseLevel &L = c->coarse->lev[level - 1];
    if (ax) *ax = L.ax;
    if (ay) *ay = L.ay;
    if (blocks) {
        HIP_TRY(hipMemcpyAsync(blocks, L.A64, sizeof(double) * 36 * (size_t)L.ax * L.ay, hipMemcpyDeviceToHost,
                               c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return CISM_OK;
}

__global__ void k_d2f(long long n, const double *__restrict__ a, float *__restrict__ b) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) b[q] = (float)a[q];
}
__global__ void k_f2d(long long n, const float *__restrict__ a, double *__restrict__ b) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) b[q] = (double)a[q];
}

extern "C" int cism_coarse_correction(cism_ctx *c, const double *ru, const double *rv, double *zu, double *zv) {
    CTX_BEGIN(c);
    if (!c->coarse || !c->coarse->ok || !c->A32) {
        cism_set_error("no coarse hierarchy: build one (cism_coarse_build or a solve with coarse = 1)");
        return CISM_ERR_ARG;
    }
    const long long n2 = 2 * c->g.nn;
    CISM_TRY(upload_vec(c, c->t, ru, rv));
    hipLaunchKernelGGL(k_d2f, dim3(nblocks_for(n2)), dim3(256), 0, c->stream, n2, c->t, c->r32);
    HIP_TRY(hipMemsetAsync(c->z32, 0, sizeof(float) * n2, c->stream));
    CISM_TRY(coarse_colsum(c, c->r32));
    CISM_TRY(coarse_apply(c, c->z32));
    hipLaunchKernelGGL(k_f2d, dim3(nblocks_for(n2)), dim3(256), 0, c->stream, n2, c->z32, c->t);
    HIP_TRY(hipGetLastError());
    return download_vec(c, c->t, zu, zv);
}

extern "C" int cism_time_kernel(cism_ctx *c, int which, int reps, double *ms_per_launch) {
    CTX_BEGIN(c);
    if (reps < 1 || which < 0 || which > 6) return CISM_ERR_ARG;
    if ((which == 1 || which == 4) && !c->A32) {
        cism_set_error("no FP32 matrix yet: run a mixed-precision solve first");
        return CISM_ERR_ARG;
    }
    const bool fp64_matvec = which == 0 || which == 5 || which == 6;
    if (fp64_matvec && !c->a64_valid) {
        cism_set_error("no FP64 matrix: the last mixed-precision solve (skip_final_residual) kept none");
        return CISM_ERR_ARG;
    }
    if (((which == 0 && c->sym64_on) || which == 6) && !c->a64s_valid) {
        // the FP64 solver's symmetric copy, as its Picard iteration forms it: a symmetry pass
        if (ensure_a64s(c) == CISM_OK) {
            CISM_TRY(symmetry_device(c, nullptr, c->A64s));
            CISM_TRY(sync_scalars(c));
            sym64_commit(c);
        }
        if (which == 6 && !c->a64s_valid) {
            cism_set_error("the FP64 matrix has no exactly symmetric copy (allocation or unrepaired pairs)");
            return CISM_ERR_ARG;
        }
    }
    // which 3: 1 GiB read + 1 GiB write copy, buffers owned by this call
    const long long ncopy = 1LL << 27;
    double *ca = nullptr, *cb = nullptr;
    if (which == 3) {
        HIP_TRY(hipMalloc(&ca, sizeof(double) * ncopy));
        if (hipMalloc(&cb, sizeof(double) * ncopy) != hipSuccess) {
            (void)hipFree(ca);
            cism_set_error("copy calibration: allocation failed");
            return CISM_ERR_NOMEM;
        }
        HIP_TRY(hipMemsetAsync(ca, 0, sizeof(double) * ncopy, c->stream));
    }
    auto launch = [&]() -> int {
        if (which == 2) return assemble_device(c, 2, 2, 0.0, 0, 0.0, 0.0);
        if (which == 3) {
            hipLaunchKernelGGL(k_copy_calib, dim3(nblocks_for(ncopy / 2)), dim3(256), 0, c->stream, ncopy,
                               (const double2 *)ca, (double2 *)cb);
            HIP_TRY(hipGetLastError());
            return CISM_OK;
        }
        if (which == 4) return launch_cg_update32(c);
        return launch_cg_matvec(c, which);
    };
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    // one untimed launch to settle clocks and caches
    int st = launch();
    if (!st) st = hipEventRecord(e0, c->stream) == hipSuccess ? CISM_OK : CISM_ERR_HIP;
    for (int r = 0; r < reps && !st; ++r) st = launch();
    float ms = 0.f;
    if (!st && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        st = CISM_ERR_HIP;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ca) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(ca);
        (void)hipFree(cb);
    }
    if (st) return st;
    if (ms_per_launch) *ms_per_launch = (double)ms / reps;
    return CISM_OK;
}

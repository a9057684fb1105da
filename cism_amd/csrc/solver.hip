// solver.hip -- structured 27-point block matvec, preconditioners and the two
// FP64 PCG iterations of glissade_velo_higher_pcg.F90, plus the FP32 inner
// solver of the mixed-precision refinement.
//
// All vectors live on the full staggered grid including the halo, and every
// kernel covers the index ranges of the reference loop it replaces, so halo
// values evolve exactly as in the reference.  Halo updates
// (staggered_parallel_halo) are device-local copies from the periodic image.
#include "internal.h"

#include <math.h>
#include <stdio.h>

// ---------------------------------------------------------------------------
// final reduction + CG scalar updates (single block, fixed order)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(CISM_BLOCK) void k_reduce(const double *__restrict__ partials,
                                                       int nb, int nv, int mode,
                                                       DevScal *__restrict__ sc) {
    __shared__ double red[CISM_BLOCK];
    for (int c = 0; c < nv; ++c) {
        double s;
        if (mode == RED_MAX) {
            s = 0.0;
            for (int b = threadIdx.x; b < nb; b += blockDim.x) s = fmax(s, partials[(size_t)b * nv + c]);
        } else {
            s = 0.0;
            for (int b = threadIdx.x; b < nb; b += blockDim.x) s += partials[(size_t)b * nv + c];
        }
        red[threadIdx.x] = s;
        __syncthreads();
        for (int off = blockDim.x / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) {
                if (mode == RED_MAX) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
                else red[threadIdx.x] += red[threadIdx.x + off];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) sc->sums[c] = red[0];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double s0 = sc->sums[0], s1 = sc->sums[1];
    switch (mode) {
    case RED_BB:
        sc->L2_rhs = sqrt(s0);
        break;
    case RED_STD_ETA1:
        sc->eta1 = s0;
        if (s0 != s0) sc->nan_flag = 1;
        sc->beta = sc->eta1 / sc->eta0;
        break;
    case RED_STD_ETA2:
        sc->eta2 = s0;
        sc->eta0 = sc->eta1;
        sc->alpha = sc->eta1 / sc->eta2;
        if (sc->alpha != sc->alpha) sc->nan_flag = 1;
        break;
    case RED_CG_FIRST:
        sc->rho_old = s0;
        sc->sigma = s1;
        sc->alpha = sc->rho_old / sc->sigma;
        if (sc->alpha != sc->alpha) sc->nan_flag = 1;
        break;
    case RED_CG_ITER: {
        const double rho = s0, delta = s1;
        sc->beta = rho / sc->rho_old;
        sc->sigma = delta - sc->beta * rho / sc->alpha;
        sc->alpha = rho / sc->sigma;
        sc->rho_old = rho;
        if (sc->alpha != sc->alpha) sc->nan_flag = 1;
        break;
    }
    case RED_ERR:
        sc->err = sqrt(s0) / sc->L2_rhs;
        break;
    default:
        break;
    }
}

int launch_reduce(cism_ctx *c, int nb, int nv, int mode) {
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(CISM_BLOCK), 0, c->stream, c->partials, nb, nv,
                       mode, c->sc);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// ---------------------------------------------------------------------------
// 27-point row product (pcg:2607-2645): slots in indxA order (kA, jA, iA
// nested, iA fastest); per slot yu += Auu*xu + Auv*xv, yv += Avu*xu + Avv*xv.
// Only k+kA needs a bounds test: owned rows never touch i, j out of range.
// ---------------------------------------------------------------------------
template <typename T, typename TA>
__device__ inline void row_product(const GridDev &g, const TA *__restrict__ A,
                                   const T *__restrict__ x, long long q, int k,
                                   T &yu, T &yv) {
    const long long np = g.npad;
    const long long cst = g.nz, rst = (long long)g.nz * g.sx;
    T su = yu, sv = yv;
#pragma unroll
    for (int kA = -1; kA <= 1; ++kA) {
        if (k + kA < 0 || k + kA >= g.nz) continue;
#pragma unroll
        for (int jA = -1; jA <= 1; ++jA)
#pragma unroll
            for (int iA = -1; iA <= 1; ++iA) {
                const int m = slot_of(iA, jA, kA);
                const long long qn = q + kA + iA * cst + jA * rst;
                const T xu = x[qn], xv = x[np + qn];
                su = su + (T)A[(0 * 27 + m) * np + q] * xu + (T)A[(1 * 27 + m) * np + q] * xv;
                sv = sv + (T)A[(2 * 27 + m) * np + q] * xu + (T)A[(3 * 27 + m) * np + q] * xv;
            }
    }
    yu = su;
    yv = sv;
}

// y = A x on owned active rows, 0 elsewhere (pcg:2533-2647)
template <typename T, typename TA>
__global__ __launch_bounds__(CISM_BLOCK) void k_matvec(GridDev g, const TA *__restrict__ A,
                                                       const int32_t *__restrict__ av,
                                                       const T *__restrict__ x, T *__restrict__ y) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nn) return;
    const int k = (int)(q % g.nz);
    const long long col = q / g.nz;
    const int i = (int)(col % g.sx), j = (int)(col / g.sx);
    T yu = 0, yv = 0;
    if (is_owned(g, i, j) && av[col]) row_product<T, TA>(g, A, x, q, k, yu, yv);
    y[q] = yu;
    y[g.npad + q] = yv;
}

int launch_matvec64(cism_ctx *c, const double *A, const double *x, double *y) {
    hipLaunchKernelGGL((k_matvec<double, double>), dim3(nblocks_for(c->g.nn)), dim3(CISM_BLOCK), 0,
                       c->stream, c->g, A, c->av, x, y);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// y = A x plus per-block partial dot products over owned nodes:
//   NDOT = 1: s0 = (x, y)            [standard PCG (d, Ad); Chron-Gear first pass]
//   NDOT = 2: s0 = (r, x), s1 = (y, x)  [Chron-Gear (r, z), (Az, z)]
template <typename T, typename TA, int NDOT>
__global__ __launch_bounds__(CISM_BLOCK) void k_matvec_dots(GridDev g, const TA *__restrict__ A,
                                                            const int32_t *__restrict__ av,
                                                            const T *__restrict__ x, T *__restrict__ y,
                                                            const T *__restrict__ r,
                                                            double *__restrict__ partials) {
    double acc[NDOT];
#pragma unroll
    for (int c = 0; c < NDOT; ++c) acc[c] = 0.0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < g.nn; q += stride) {
        const int k = (int)(q % g.nz);
        const long long col = q / g.nz;
        const int i = (int)(col % g.sx), j = (int)(col / g.sx);
        T yu = 0, yv = 0;
        const bool own = is_owned(g, i, j);
        if (own && av[col]) row_product<T, TA>(g, A, x, q, k, yu, yv);
        y[q] = yu;
        y[g.npad + q] = yv;
        if (own) {
            const double xu = x[q], xv = x[g.npad + q];
            if (NDOT == 1) {
                acc[0] += xu * (double)yu + xv * (double)yv;
            } else {
                acc[0] += (double)r[q] * xu + (double)r[g.npad + q] * xv;
                acc[NDOT - 1] += (double)yu * xu + (double)yv * xv;
            }
        }
    }
    block_reduce_store<NDOT>(acc, partials);
}

// r = b - A x on the full grid (A x = 0 off owned active rows), s0 = (r, r) over
// owned nodes.  halo_r: the halo takes the periodic image of r afterwards
// (Chron-Gear, pcg:1592-1595), computed here from the image row directly.
template <typename T, typename TA>
__global__ __launch_bounds__(CISM_BLOCK) void k_residual(GridDev g, const TA *__restrict__ A,
                                                         const int32_t *__restrict__ av,
                                                         const T *__restrict__ x,
                                                         const T *__restrict__ b, T *__restrict__ r,
                                                         int halo_r, double *__restrict__ partials) {
    double acc[1] = {0.0};
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < g.nn; q += stride) {
        const int k = (int)(q % g.nz);
        const long long col = q / g.nz;
        int i = (int)(col % g.sx), j = (int)(col / g.sx);
        const bool own = is_owned(g, i, j);
        long long qs = q;     // row whose value this entry takes
        bool zero = false;
        if (halo_r && !own) {
            const int ii = halo_image(i, g.sx, g.periodic), jj = halo_image(j, g.sy, g.periodic);
            if (ii < 0 || jj < 0) zero = true;
            else { i = ii; j = jj; qs = k + (long long)g.nz * (i + (long long)g.sx * j); }
        }
        T ru = 0, rv = 0;
        if (!zero) {
            T yu = 0, yv = 0;
            const long long cs = qs / g.nz;
            if (is_owned(g, i, j) && av[cs]) row_product<T, TA>(g, A, x, qs, k, yu, yv);
            ru = b[qs] - yu;
            rv = b[g.npad + qs] - yv;
        }
        r[q] = ru;
        r[g.npad + q] = rv;
        if (own) acc[0] += (double)ru * ru + (double)rv * rv;
    }
    block_reduce_store<1>(acc, partials);
}

// s0 = (a, b) over owned nodes (global_sum_staggered_3d_real8, pcg:2333-2379)
template <typename T>
__global__ __launch_bounds__(CISM_BLOCK) void k_dot(GridDev g, const T *__restrict__ a,
                                                    const T *__restrict__ b,
                                                    double *__restrict__ partials) {
    double acc[1] = {0.0};
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < g.nn; q += stride) {
        const long long col = q / g.nz;
        const int i = (int)(col % g.sx), j = (int)(col / g.sx);
        if (is_owned(g, i, j))
            acc[0] += (double)a[q] * (double)b[q] + (double)a[g.npad + q] * (double)b[g.npad + q];
    }
    block_reduce_store<1>(acc, partials);
}

// ---------------------------------------------------------------------------
// preconditioners (setup_preconditioner_3d pcg:2150-2265 and the apply
// branches pcg:300-365 / 1416-1454; easy_sia_solver + tridiag_solver
// pcg:2759-2904).  Coefficients are read straight from the matrix slots.
// ---------------------------------------------------------------------------
template <typename T, typename TA>
__global__ __launch_bounds__(CISM_BLOCK) void k_precond_diag(GridDev g, const TA *__restrict__ A,
                                                             const T *__restrict__ r,
                                                             T *__restrict__ z) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nn) return;
    const long long np = g.npad;
    const T du = (T)A[(0 * 27 + CISM_DIAG) * np + q], dv = (T)A[(3 * 27 + CISM_DIAG) * np + q];
    z[q] = du != (T)0 ? r[q] / du : (T)0;
    z[np + q] = dv != (T)0 ? r[np + q] / dv : (T)0;
}

template <typename T>
__global__ __launch_bounds__(CISM_BLOCK) void k_copy2(long long nn, long long np,
                                                      const T *__restrict__ a, T *__restrict__ b) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nn) return;
    b[q] = a[q];
    b[np + q] = a[np + q];
}

// One lane per (column, component): Thomas algorithm without pivoting on the
// vertical couplings (slots indxA(0,0,-1|0|+1)) of Auu (c=0) or Avv (c=1),
// for every active vertex including the halo; inactive columns keep z.
// The column of r is staged through LDS with unit-stride loads.
#define SIA_COLS 64
template <typename T, typename TA, int NZMAX>
__global__ __launch_bounds__(SIA_COLS) void k_precond_sia(GridDev g, const TA *__restrict__ A,
                                                          const int32_t *__restrict__ av,
                                                          const T *__restrict__ r,
                                                          T *__restrict__ z) {
    const int nz = g.nz;
    const long long ncol = (long long)g.sx * g.sy;
    const int comp = blockIdx.y;
    const long long col0 = (long long)blockIdx.x * SIA_COLS;
    const long long np = g.npad;
    const TA *Am = A + (long long)(comp == 0 ? 0 : 3) * 27 * np;
    const T *rc = r + comp * np;
    T *zc = z + comp * np;
    __shared__ T sr[SIA_COLS * NZMAX];
    const long long base = col0 * nz;
    const long long span = (col0 + SIA_COLS <= ncol ? SIA_COLS : ncol - col0) * nz;
    for (int t = threadIdx.x; t < span; t += SIA_COLS) sr[t] = rc[base + t];
    __syncthreads();
    const long long col = col0 + threadIdx.x;
    const bool act = col < ncol && av[col];
    if (act) {
        T gam[NZMAX];
        T *s = sr + threadIdx.x * nz;
        const long long qb = col * nz;
        // top layer: sub = 0; bottom: super = 0 (pcg:2803-2829)
        T beta = (T)Am[CISM_DIAG * np + qb];
        s[0] = s[0] / beta;
        for (int k = 1; k < nz; ++k) {
            const T sp = (T)Am[(CISM_DIAG + 9) * np + qb + k - 1];   // A(+1) of row k-1
            const T sb = (T)Am[(CISM_DIAG - 9) * np + qb + k];       // A(-1) of row k
            gam[k] = sp / beta;
            beta = (T)Am[CISM_DIAG * np + qb + k] - sb * gam[k];
            s[k] = (s[k] - sb * s[k - 1]) / beta;
        }
        for (int k = nz - 2; k >= 0; --k) s[k] = s[k] - gam[k + 1] * s[k + 1];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < span; t += SIA_COLS) {
        const long long cc = col0 + t / nz;
        if (av[cc]) zc[base + t] = sr[t];
    }
}

template <typename T, typename TA>
static int launch_precond(cism_ctx *c, int precond, const TA *A, const T *r, T *z) {
    const GridDev &g = c->g;
    if (precond == 0) {
        hipLaunchKernelGGL(k_copy2<T>, dim3(nblocks_for(g.nn)), dim3(CISM_BLOCK), 0, c->stream,
                           g.nn, g.npad, r, z);
    } else if (precond == 1) {
        hipLaunchKernelGGL((k_precond_diag<T, TA>), dim3(nblocks_for(g.nn)), dim3(CISM_BLOCK), 0,
                           c->stream, g, A, r, z);
    } else {
        const long long ncol = (long long)g.sx * g.sy;
        dim3 grid((unsigned)((ncol + SIA_COLS - 1) / SIA_COLS), 2);
        if (g.nz <= 16)
            hipLaunchKernelGGL((k_precond_sia<T, TA, 16>), grid, dim3(SIA_COLS), 0, c->stream, g, A,
                               c->av, r, z);
        else if (g.nz <= 32)
            hipLaunchKernelGGL((k_precond_sia<T, TA, 32>), grid, dim3(SIA_COLS), 0, c->stream, g, A,
                               c->av, r, z);
        else {
            cism_set_error("SIA preconditioner supports nz <= 32 (nz = %d)", g.nz);
            return CISM_ERR_UNSUPPORTED;
        }
    }
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// ---------------------------------------------------------------------------
// halo copies (staggered_parallel_halo_real8_3d/_4d)
// ---------------------------------------------------------------------------
// halo vertices: rows j in {0,1,sy-1} (all i) and columns i in {0,1,sx-1} of rows 2..sy-2
__device__ inline bool halo_vertex_of(const GridDev &g, long long h, int &i, int &j) {
    const long long nrow = 3LL * g.sx;
    if (h < nrow) {
        const int rr = (int)(h / g.sx);
        i = (int)(h % g.sx);
        j = rr < 2 ? rr : g.sy - 1;
        return true;
    }
    h -= nrow;
    const long long ncolh = 3LL * (g.sy - 3);
    if (h >= ncolh) return false;
    const int cc = (int)(h % 3);
    j = 2 + (int)(h / 3);
    i = cc < 2 ? cc : g.sx - 1;
    return true;
}

template <typename T>
__global__ __launch_bounds__(CISM_BLOCK) void k_halo_planes(GridDev g, T *__restrict__ v,
                                                            long long plane, int nplanes) {
    const long long nh = (3LL * g.sx + 3LL * (g.sy - 3)) * g.nz;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nh * nplanes) return;
    const int p = (int)(t / nh);
    const long long hv = t % nh;
    const int k = (int)(hv % g.nz);
    int i, j;
    if (!halo_vertex_of(g, hv / g.nz, i, j)) return;
    const int ii = halo_image(i, g.sx, g.periodic), jj = halo_image(j, g.sy, g.periodic);
    T *vp = v + p * plane;
    const long long qd = k + (long long)g.nz * (i + (long long)g.sx * j);
    if (ii < 0 || jj < 0) vp[qd] = (T)0;
    else vp[qd] = vp[k + (long long)g.nz * (ii + (long long)g.sx * jj)];
}

template <typename T>
static int launch_halo_planes(cism_ctx *c, T *v, int nplanes) {
    const GridDev &g = c->g;
    const long long nh = (3LL * g.sx + 3LL * (g.sy - 3)) * g.nz * nplanes;
    hipLaunchKernelGGL(k_halo_planes<T>, dim3(nblocks_for(nh)), dim3(CISM_BLOCK), 0, c->stream, g, v,
                       g.npad, nplanes);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

int launch_halo_vec(cism_ctx *c, double *v) { return launch_halo_planes<double>(c, v, 2); }
int launch_halo_mat(cism_ctx *c, double *A) { return launch_halo_planes<double>(c, A, 108); }

// ---------------------------------------------------------------------------
// vector updates
// ---------------------------------------------------------------------------
// standard PCG: d = z + beta*d (pcg:383-393)
template <typename T>
__global__ __launch_bounds__(CISM_BLOCK) void k_std_d(long long nn, long long np,
                                                      const T *__restrict__ z, T *__restrict__ d,
                                                      const DevScal *__restrict__ sc) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nn) return;
    const T beta = (T)sc->beta;
    d[q] = z[q] + beta * d[q];
    d[np + q] = z[np + q] + beta * d[np + q];
}

// standard PCG: x += alpha*d, r -= alpha*q (pcg:466-472)
template <typename T>
__global__ __launch_bounds__(CISM_BLOCK) void k_std_xr(long long nn, long long np,
                                                       const T *__restrict__ d, const T *__restrict__ qv,
                                                       T *__restrict__ x, T *__restrict__ r,
                                                       const DevScal *__restrict__ sc) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nn) return;
    const T alpha = (T)sc->alpha;
    x[q] = x[q] + alpha * d[q];
    x[np + q] = x[np + q] + alpha * d[np + q];
    r[q] = r[q] - alpha * qv[q];
    r[np + q] = r[np + q] - alpha * qv[np + q];
}

// Chron-Gear (pcg:1517-1536) with the halo update of Az (pcg:1494-1497) folded
// in: halo entries read Az at their periodic image (or 0, outflow).
template <typename T>
__global__ __launch_bounds__(CISM_BLOCK) void k_cg_update(GridDev g, const T *__restrict__ z,
                                                          const T *__restrict__ Az,
                                                          T *__restrict__ d, T *__restrict__ qv,
                                                          T *__restrict__ x, T *__restrict__ r,
                                                          const DevScal *__restrict__ sc, int first) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nn) return;
    const long long np = g.npad;
    const T alpha = (T)sc->alpha, beta = (T)sc->beta;
    T azu, azv;
    {
        const int k = (int)(q % g.nz);
        const long long col = q / g.nz;
        const int i = (int)(col % g.sx), j = (int)(col / g.sx);
        if (is_owned(g, i, j)) {
            azu = Az[q];
            azv = Az[np + q];
        } else {
            const int ii = halo_image(i, g.sx, g.periodic), jj = halo_image(j, g.sy, g.periodic);
            if (ii < 0 || jj < 0) {
                azu = azv = (T)0;
            } else {
                const long long qs = k + (long long)g.nz * (ii + (long long)g.sx * jj);
                azu = Az[qs];
                azv = Az[np + qs];
            }
        }
    }
    T du, dv, qu, qvv;
    if (first) {   // first pass: d = z, q = A d (already halo-mapped), pcg:1351-1405
        du = z[q]; dv = z[np + q];
        qu = azu; qvv = azv;
    } else {
        du = z[q] + beta * d[q];
        dv = z[np + q] + beta * d[np + q];
        qu = azu + beta * qv[q];
        qvv = azv + beta * qv[np + q];
    }
    d[q] = du; d[np + q] = dv;
    qv[q] = qu; qv[np + q] = qvv;
    x[q] = x[q] + alpha * du;
    x[np + q] = x[np + q] + alpha * dv;
    r[q] = r[q] - alpha * qu;
    r[np + q] = r[np + q] - alpha * qvv;
}


int launch_halo_n(cism_ctx *c, double *v, int nplanes) { return launch_halo_planes<double>(c, v, nplanes); }

// ---------------------------------------------------------------------------
// PCG drivers
// ---------------------------------------------------------------------------
__global__ void k_init_scalars(DevScal *sc) {
    sc->rho_old = sc->alpha = sc->beta = sc->sigma = sc->eta1 = sc->eta2 = 0.0;
    sc->eta0 = 1.0;
    sc->err = 0.0;
    sc->nan_flag = 0;
}

int sync_scalars(cism_ctx *c) {
    HIP_TRY(hipMemcpyAsync(c->sc_host, c->sc, sizeof(DevScal), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CISM_OK;
}

template <typename T>
__global__ __launch_bounds__(CISM_BLOCK) void k_zero2(long long n2, T *__restrict__ a, T *__restrict__ b,
                                                      T *__restrict__ c3) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n2) return;
    a[q] = (T)0;
    b[q] = (T)0;
    c3[q] = (T)0;
}

template <typename T>
struct CGVecs {
    T *x, *b, *r, *d, *z, *q, *Az;
};

// The reference iterations on a generic precision.  check = 0: every fifth
// iteration the true residual b - A x decides convergence against tol (the
// reference rule, pcg:478-537, 1542-1597).  check = 1 (inner solves of the
// mixed-precision refinement): the recursive residual r is measured instead,
// with no extra matvec, against tol * |r0|.
template <typename T, typename TA>
static int run_cg(cism_ctx *c, const TA *A, CGVecs<T> v, int chrongear, int precond, int maxiters,
                  double tol, int check, double *err_out, int *niters_out) {
    const GridDev &g = c->g;
    const int nb = red_blocks_for(g.nn), nbe = nblocks_for(g.nn);
    const long long np = g.npad;
    hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(1), 0, c->stream, c->sc);
    hipLaunchKernelGGL(k_zero2<T>, dim3(nblocks_for(2 * np)), dim3(CISM_BLOCK), 0, c->stream, 2 * np,
                       v.d, v.z, v.q);
    // (b, b) -> L2_rhs
    hipLaunchKernelGGL(k_dot<T>, dim3(nb), dim3(CISM_BLOCK), 0, c->stream, g, v.b, v.b, c->partials);
    CISM_TRY(launch_reduce(c, nb, 1, RED_BB));
    int niters = maxiters;
    double err = 0.0;
    auto check_conv = [&](int n, bool &done) -> int {
        if (check == 0) {
            hipLaunchKernelGGL((k_residual<T, TA>), dim3(nb), dim3(CISM_BLOCK), 0, c->stream, g, A,
                               c->av, v.x, v.b, v.r, chrongear, c->partials);
        } else {
            hipLaunchKernelGGL(k_dot<T>, dim3(nb), dim3(CISM_BLOCK), 0, c->stream, g, v.r, v.r,
                               c->partials);
        }
        CISM_TRY(launch_reduce(c, nb, 1, RED_ERR));
        CISM_TRY(sync_scalars(c));
        if (c->sc_host->nan_flag) {
            cism_set_error("PCG solver has failed, alpha = NaN");
            return CISM_ERR_NAN;
        }
        err = c->sc_host->err;
        done = err < tol;
        if (done) niters = n;
        return CISM_OK;
    };
    bool done = false;
    if (!chrongear) {
        // pcg:253-304: halo x, r = b - A x (no halo on r)
        CISM_TRY(launch_halo_planes<T>(c, v.x, 2));
        hipLaunchKernelGGL((k_residual<T, TA>), dim3(nb), dim3(CISM_BLOCK), 0, c->stream, g, A, c->av,
                           v.x, v.b, v.r, 0, c->partials);
        for (int n = 1; n <= maxiters && !done; ++n) {
            CISM_TRY(launch_precond<T, TA>(c, precond, A, v.r, v.z));
            hipLaunchKernelGGL(k_dot<T>, dim3(nb), dim3(CISM_BLOCK), 0, c->stream, g, v.r, v.z,
                               c->partials);
            CISM_TRY(launch_reduce(c, nb, 1, RED_STD_ETA1));
            hipLaunchKernelGGL(k_std_d<T>, dim3(nbe), dim3(CISM_BLOCK), 0, c->stream, g.nn, np, v.z,
                               v.d, c->sc);
            CISM_TRY(launch_halo_planes<T>(c, v.d, 2));
            hipLaunchKernelGGL((k_matvec_dots<T, TA, 1>), dim3(nb), dim3(CISM_BLOCK), 0, c->stream, g,
                               A, c->av, v.d, v.q, (const T *)nullptr, c->partials);
            CISM_TRY(launch_reduce(c, nb, 1, RED_STD_ETA2));
            hipLaunchKernelGGL(k_std_xr<T>, dim3(nbe), dim3(CISM_BLOCK), 0, c->stream, g.nn, np, v.d,
                               v.q, v.x, v.r, c->sc);
            if (n % 5 == 0) {
                if (check == 0) CISM_TRY(launch_halo_planes<T>(c, v.x, 2));
                CISM_TRY(check_conv(n, done));
            }
        }
    } else {
        // pcg:1258-1405: halo x; r = b - A x; halo r; z = M^-1 r; d = z; q = A d
        CISM_TRY(launch_halo_planes<T>(c, v.x, 2));
        hipLaunchKernelGGL((k_residual<T, TA>), dim3(nb), dim3(CISM_BLOCK), 0, c->stream, g, A, c->av,
                           v.x, v.b, v.r, 1, c->partials);
        CISM_TRY(launch_precond<T, TA>(c, precond, A, v.r, v.z));
        hipLaunchKernelGGL((k_matvec_dots<T, TA, 2>), dim3(nb), dim3(CISM_BLOCK), 0, c->stream, g, A,
                           c->av, v.z, v.Az, v.r, c->partials);
        CISM_TRY(launch_reduce(c, nb, 2, RED_CG_FIRST));
        hipLaunchKernelGGL(k_cg_update<T>, dim3(nbe), dim3(CISM_BLOCK), 0, c->stream, g, v.z, v.Az,
                           v.d, v.q, v.x, v.r, c->sc, 1);
        for (int n = 2; n <= maxiters && !done; ++n) {
            CISM_TRY(launch_precond<T, TA>(c, precond, A, v.r, v.z));
            hipLaunchKernelGGL((k_matvec_dots<T, TA, 2>), dim3(nb), dim3(CISM_BLOCK), 0, c->stream,
                               g, A, c->av, v.z, v.Az, v.r, c->partials);
            CISM_TRY(launch_reduce(c, nb, 2, RED_CG_ITER));
            hipLaunchKernelGGL(k_cg_update<T>, dim3(nbe), dim3(CISM_BLOCK), 0, c->stream, g, v.z, v.Az,
                               v.d, v.q, v.x, v.r, c->sc, 0);
            if (n % 5 == 0) CISM_TRY(check_conv(n, done));
        }
    }
    CISM_TRY(sync_scalars(c));
    if (c->sc_host->nan_flag) {
        cism_set_error("PCG solver has failed, alpha = NaN");
        return CISM_ERR_NAN;
    }
    *err_out = err;
    *niters_out = niters;
    return CISM_OK;
}

// Solves A x = b for c->x (initial guess in place) with the FP64 matrix c->A.
int solver_pcg64(cism_ctx *c, int chrongear, int precond, int maxiters, double tol, double *err,
                 int *niters) {
    CGVecs<double> v = {c->x, c->b, c->r, c->d, c->z, c->q, c->Az};
    return run_cg<double, double>(c, c->A, v, chrongear, precond, maxiters, tol, 0, err, niters);
}

// ---------------------------------------------------------------------------
// mixed precision: FP32 matrix and inner Chron-Gear, FP64 refinement
// ---------------------------------------------------------------------------
int ensure_a32(cism_ctx *c) {
    if (c->A32) return CISM_OK;
    const size_t v2 = 2 * (size_t)c->g.npad;
    hipError_t e = hipMalloc(&c->A32, sizeof(float) * 108 * (size_t)c->g.npad);
    float **vs[] = {&c->x32, &c->b32, &c->r32, &c->d32, &c->z32, &c->q32, &c->Az32};
    for (float **p : vs)
        if (e == hipSuccess) e = hipMalloc(p, sizeof(float) * v2);
    if (e != hipSuccess) {
        cism_set_error("FP32 matrix allocation failed: %s", hipGetErrorString(e));
        return CISM_ERR_NOMEM;
    }
    return CISM_OK;
}

__global__ __launch_bounds__(CISM_BLOCK) void k_to_f32(long long n, const double *__restrict__ a,
                                                       float *__restrict__ b) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) b[q] = (float)a[q];
}

// r32 = r64 / s (all entries), x32 = 0
__global__ __launch_bounds__(CISM_BLOCK) void k_rhs32(long long n2, const double *__restrict__ r,
                                                      const DevScal *__restrict__ sc,
                                                      float *__restrict__ r32, float *__restrict__ x32) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n2) return;
    const double s = sc->sums[2];
    r32[q] = (float)(s > 0.0 ? r[q] / s : 0.0);
    x32[q] = 0.0f;
}

// x += s * e (all entries)
__global__ __launch_bounds__(CISM_BLOCK) void k_refine(long long n2, const float *__restrict__ e,
                                                       const DevScal *__restrict__ sc,
                                                       double *__restrict__ x) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n2) return;
    x[q] = x[q] + sc->sums[2] * (double)e[q];
}

__global__ void k_save_scale(DevScal *sc) { sc->sums[2] = sqrt(sc->sums[0]); }

int convert_a32(cism_ctx *c) {
    CISM_TRY(ensure_a32(c));
    const long long n = 108LL * c->g.npad;
    hipLaunchKernelGGL(k_to_f32, dim3(nblocks_for(n)), dim3(CISM_BLOCK), 0, c->stream, n, c->A, c->A32);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// Iterative refinement: r = b - A64 x (FP64), solve A32 e = r/|r| with FP32
// Chron-Gear to relative tolerance inner_rtol, x += |r| e, until the FP64
// relative residual is below tol.  The FP32 matrix must be current (convert_a32).
int solver_mixed(cism_ctx *c, int precond, int maxiters, double tol, double inner_rtol, double *err_out,
                 int *niters_out, int *n64, int *n32) {
    const GridDev &g = c->g;
    const int nb = red_blocks_for(g.nn);
    const long long n2 = 2 * g.npad;
    CISM_TRY(ensure_a32(c));
    hipLaunchKernelGGL(k_dot<double>, dim3(nb), dim3(CISM_BLOCK), 0, c->stream, g, c->b, c->b, c->partials);
    CISM_TRY(launch_reduce(c, nb, 1, RED_SUMS));
    CISM_TRY(sync_scalars(c));
    const double bnorm = sqrt(c->sc_host->sums[0]);
    int total = 0, outer = 0;
    double err = 0.0;
    CGVecs<float> v = {c->x32, c->b32, c->r32, c->d32, c->z32, c->q32, c->Az32};
    while (true) {
        CISM_TRY(launch_halo_vec(c, c->x));
        hipLaunchKernelGGL((k_residual<double, double>), dim3(nb), dim3(CISM_BLOCK), 0, c->stream, g, c->A,
                           c->av, c->x, c->b, c->r, 1, c->partials);
        CISM_TRY(launch_reduce(c, nb, 1, RED_SUMS));
        hipLaunchKernelGGL(k_save_scale, dim3(1), dim3(1), 0, c->stream, c->sc);
        CISM_TRY(sync_scalars(c));
        ++*n64;
        err = bnorm > 0 ? sqrt(c->sc_host->sums[0]) / bnorm : 0.0;
        if (err < tol || total >= maxiters * 4 || outer >= 50) break;
        hipLaunchKernelGGL(k_rhs32, dim3(nblocks_for(n2)), dim3(CISM_BLOCK), 0, c->stream, n2, c->r, c->sc,
                           c->b32, c->x32);
        double e_in = 0.0;
        int it = 0;
        // the inner reductions use sums[0..1]; the scale |r| stays in sums[2]
        CISM_TRY(run_cg<float, float>(c, c->A32, v, 1, precond, maxiters, inner_rtol, 1, &e_in, &it));
        hipLaunchKernelGGL(k_refine, dim3(nblocks_for(n2)), dim3(CISM_BLOCK), 0, c->stream, n2, c->x32, c->sc,
                           c->x);
        HIP_TRY(hipGetLastError());
        total += it;
        *n32 += it;
        ++outer;
    }
    *err_out = err;
    *niters_out = total;
    return CISM_OK;
}

// ---------------------------------------------------------------------------
// measurement helpers
// ---------------------------------------------------------------------------
// one launch of the solver's matvec exactly as the Chron-Gear iteration issues it
int launch_cg_matvec(cism_ctx *c, int fp32) {
    const GridDev &g = c->g;
    const int nb = red_blocks_for(g.nn);
    if (fp32)
        hipLaunchKernelGGL((k_matvec_dots<float, float, 2>), dim3(nb), dim3(CISM_BLOCK), 0, c->stream, g,
                           c->A32, c->av, c->z32, c->Az32, c->r32, c->partials);
    else
        hipLaunchKernelGGL((k_matvec_dots<double, double, 2>), dim3(nb), dim3(CISM_BLOCK), 0, c->stream,
                           g, c->A, c->av, c->z, c->Az, c->r, c->partials);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// sums[0] = |x - xref|^2, sums[1] = |xref|^2 over owned nodes
__global__ __launch_bounds__(CISM_BLOCK) void k_err_sums(GridDev g, const double *__restrict__ x,
                                                         const double *__restrict__ xr,
                                                         double *__restrict__ partials) {
    double acc[2] = {0.0, 0.0};
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < g.nn; q += stride) {
        const long long col = q / g.nz;
        const int i = (int)(col % g.sx), j = (int)(col / g.sx);
        if (!is_owned(g, i, j)) continue;
        const double du = x[q] - xr[q], dv = x[g.npad + q] - xr[g.npad + q];
        acc[0] += du * du + dv * dv;
        acc[1] += xr[q] * xr[q] + xr[g.npad + q] * xr[g.npad + q];
    }
    block_reduce_store<2>(acc, partials);
}

int rel_err_device(cism_ctx *c, double *rel) {
    const int nb = red_blocks_for(c->g.nn);
    hipLaunchKernelGGL(k_err_sums, dim3(nb), dim3(CISM_BLOCK), 0, c->stream, c->g, c->x, c->xref,
                       c->partials);
    HIP_TRY(hipGetLastError());
    CISM_TRY(launch_reduce(c, nb, 2, RED_SUMS));
    CISM_TRY(sync_scalars(c));
    const double n = sqrt(c->sc_host->sums[1]);
    *rel = n > 0 ? sqrt(c->sc_host->sums[0]) / n : sqrt(c->sc_host->sums[0]);
    return CISM_OK;
}

// solver.hip -- structured 27-point block matvec, preconditioners and the two
// FP64 PCG iterations of glissade_velo_higher_pcg.F90, plus the FP32 inner
// solver of the mixed-precision refinement.
//
// All vectors live on the full staggered grid including the halo, and every
// kernel covers the index ranges of the reference loop it replaces, so halo
// values evolve exactly as in the reference.  Halo updates
// (staggered_parallel_halo) are device-local copies from the periodic image.
#include "internal.h"
#include "refine_policy.h"

#include <math.h>
#include <type_traits>
#include <stdio.h>

// ---------------------------------------------------------------------------
// final reductions + CG scalar updates (fixed order)
// ---------------------------------------------------------------------------
// The scalar step that follows a reduction (mode), on sc->sums.
__device__ inline void scalar_update(int mode, DevScal *__restrict__ sc) {
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
        if (sc->zero_stop && s0 == 0.0) {
            sc->alpha = 0.0;
            break;
        }
        sc->alpha = sc->rho_old / sc->sigma;
        if (sc->alpha != sc->alpha) sc->nan_flag = 1;
        break;
    case RED_CG_ITER: {
        const double rho = s0, delta = s1;
        if (sc->zero_stop && (rho == 0.0 || sc->rho_old == 0.0)) {
            sc->alpha = 0.0;
            sc->beta = 0.0;
            sc->rho_old = rho;
            break;
        }
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

// k_reduce, for any width: one block of 1024 lanes, each with 8 independent
// accumulators over a fixed set of partials (lane t, chain u takes blocks
// t + 1024 (8 r + u)), combined in a fixed order: the partials of a 70,000-block
// grid are read with 8 loads in flight per lane instead of one dependent chain, and
// the result does not depend on timing.
#define RED_THREADS 1024
__global__ __launch_bounds__(RED_THREADS) void k_reduce(const double *__restrict__ partials,
                                                        int nb, int nv, int mode,
                                                        DevScal *__restrict__ sc) {
    __shared__ double red[RED_THREADS];
    const bool mx = mode == RED_MAX;
    for (int c = 0; c < nv; ++c) {
        double a[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] = 0.0;
        int b = threadIdx.x;
        for (; b + 7 * RED_THREADS < nb; b += 8 * RED_THREADS) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = partials[(size_t)(b + u * RED_THREADS) * nv + c];
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] = mx ? fmax(a[u], v[u]) : a[u] + v[u];
        }
        for (int u = 0; b < nb; b += RED_THREADS, ++u) {   // fewer than 8 strides remain
            const double v = partials[(size_t)b * nv + c];
            a[u] = mx ? fmax(a[u], v) : a[u] + v;
        }
        double s = a[0];
#pragma unroll
        for (int u = 1; u < 8; ++u) s = mx ? fmax(s, a[u]) : s + a[u];
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
    scalar_update(mode, sc);
}

// Two-pass reduction for the common widths: RED_GROUPS workgroups each combine a
// contiguous range of the partials (lane-strided, then a fixed shuffle tree), then
// one wave combines the group results in group order and applies the scalar step.
// Deterministic like k_reduce, and the first pass is spread over the chip.
#define RED_GROUPS 64
template <int NV>
__global__ __launch_bounds__(CISM_BLOCK) void k_reduce_groups(const double *__restrict__ partials, int nb,
                                                              int mode, double *__restrict__ groups) {
    const bool mx = mode == RED_MAX;
    const int chunk = (nb + RED_GROUPS - 1) / RED_GROUPS;
    const int b0 = blockIdx.x * chunk, b1 = min(nb, b0 + chunk);
    double a[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) a[c] = 0.0;
    for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x)
#pragma unroll
        for (int c = 0; c < NV; ++c) {
            const double v = partials[(size_t)b * NV + c];
            a[c] = mx ? fmax(a[c], v) : a[c] + v;
        }
    if (mx) block_max_store<NV>(a, groups);
    else block_reduce_store<NV>(a, groups);
}

// out = null: the sums go to sc->sums and the scalar step (mode) follows; else
// they go to out
template <int NV>
__global__ void k_reduce_final(const double *__restrict__ groups, int mode, DevScal *__restrict__ sc,
                               double *__restrict__ out) {
    const bool mx = mode == RED_MAX;
    if (threadIdx.x < NV) {
        double s = groups[threadIdx.x];
        for (int g = 1; g < RED_GROUPS; ++g) {
            const double v = groups[g * NV + threadIdx.x];
            s = mx ? fmax(s, v) : s + v;
        }
        if (out) out[threadIdx.x] = s;
        else sc->sums[threadIdx.x] = s;
    }
    if (out) return;
    __syncthreads();
    if (threadIdx.x == 0) scalar_update(mode, sc);
}

template <int NV>
static void launch_reduce2(cism_ctx *c, int nb, int mode, double *out = nullptr) {
    hipLaunchKernelGGL(k_reduce_groups<NV>, dim3(RED_GROUPS), dim3(CISM_BLOCK), 0, c->stream, c->partials, nb,
                       mode, c->red_groups);
    hipLaunchKernelGGL(k_reduce_final<NV>, dim3(1), dim3(64), 0, c->stream, c->red_groups, mode, c->sc, out);
}

// sums of NV values per block of the partials into out[0..NV), nothing else
template <int NV>
static int launch_reduce_to(cism_ctx *c, int nb, double *out) {
    launch_reduce2<NV>(c, nb, RED_SUMS, out);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

int launch_reduce(cism_ctx *c, int nb, int nv, int mode) {
    if (nv == 1) launch_reduce2<1>(c, nb, mode);
    else if (nv == 2) launch_reduce2<2>(c, nb, mode);
    else if (nv == 4) launch_reduce2<4>(c, nb, mode);
    else if (nv == 5) launch_reduce2<5>(c, nb, mode);
    else
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(RED_THREADS), 0, c->stream, c->partials, nb, nv, mode, c->sc);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// ---------------------------------------------------------------------------
// 27-point row product (pcg:2607-2645): slots in indxA order (kA, jA, iA
// nested, iA fastest); per slot yu += Auu*xu + Auv*xv, yv += Avu*xu + Avv*xv.
// Only k+kA needs a bounds test: owned rows never touch i, j out of range.
// The four blocks of a slot come with one 16-byte load (FP32) or two (FP64);
// x arrives as (u, v) pairs.
// ---------------------------------------------------------------------------
template <typename T> struct Vec2;
template <> struct Vec2<float> { using t = float2; };
template <> struct Vec2<double> { using t = double2; };


// Matrix accessors.  blocks(q, m, qn): the four blocks of slot m of row q, whose
// neighbour is qn; entry(g, b, m, q): one block of one slot.
template <typename T>
struct MatFull {   // all 27 slots stored (aidx)
    static constexpr bool xcd = false;
    static constexpr bool node_partials = false;
    const T *A;
    __device__ inline void blocks(long long q, int m, long long qn, T &uu, T &uv, T &vu, T &vv) const {
        (void)qn;
        if constexpr (std::is_same<T, double>::value) load_blocks_nt(A, q, m, uu, uv, vu, vv);
        else load_blocks(A, q, m, uu, uv, vu, vv);
    }
    __device__ inline T entry(const GridDev &g, int b, int m, long long q) const {
        (void)g;
        return A[aidx(b, m, q)];
    }
};

struct MatSym32 {   // FP32 symmetric storage (sidx); matvec launched XCD-aware
    static constexpr bool xcd = true;
    static constexpr bool node_partials = false;   // partials in launch order
    const float *A;
    __device__ inline void blocks(long long q, int m, long long qn, float &uu, float &uv, float &vu,
                                  float &vv) const {
        if (m >= CISM_DIAG) {
            const float4 a = *reinterpret_cast<const float4 *>(A + sidx(0, m - CISM_DIAG, q));
            uu = a.x; uv = a.y; vu = a.z; vv = a.w;
        } else {
            const float4 a = *reinterpret_cast<const float4 *>(A + sidx(0, CISM_DIAG - m, qn));
            uu = a.x; uv = a.z; vu = a.y; vv = a.w;
        }
    }
    __device__ inline float entry(const GridDev &g, int b, int m, long long q) const {
        if (m >= CISM_DIAG) return A[sidx(b, m - CISM_DIAG, q)];
        const int iA = m % 3 - 1, jA = (m / 3) % 3 - 1, kA = m / 9 - 1;
        const long long qn = q + kA + (long long)g.nz * (iA + (long long)g.sx * jA);
        const int bt = b == 1 ? 2 : (b == 2 ? 1 : b);
        return A[sidx(bt, CISM_DIAG - m, qn)];
    }
};

// FP64 symmetric storage (s64idx) of an exactly symmetric FP64 matrix: the same products
// as MatFull on it, bit for bit, from 448 instead of 864 stored bytes per node.  XCD: the
// matvec's launch order (see xcd_block); its dot partials keep the node-block order of the
// plain launch either way (partial_block), so every sum is the one MatFull gives.
#define CISM_SYM64_XCD true
template <bool XCD>
struct MatSym64 {
    static constexpr bool xcd = XCD;
    static constexpr bool node_partials = true;
    const double *A;
    __device__ inline void blocks(long long q, int m, long long qn, double &uu, double &uv, double &vu,
                                  double &vv) const {
        if (m >= CISM_DIAG) {
            const double2 a = *reinterpret_cast<const double2 *>(A + s64idx(0, m - CISM_DIAG, q));
            const double2 b = *reinterpret_cast<const double2 *>(A + s64idx(2, m - CISM_DIAG, q));
            uu = a.x; uv = a.y; vu = b.x; vv = b.y;
        } else {
            const double2 a = *reinterpret_cast<const double2 *>(A + s64idx(0, CISM_DIAG - m, qn));
            const double2 b = *reinterpret_cast<const double2 *>(A + s64idx(2, CISM_DIAG - m, qn));
            uu = a.x; uv = b.x; vu = a.y; vv = b.y;
        }
    }
    __device__ inline double entry(const GridDev &g, int b, int m, long long q) const {
        if (m >= CISM_DIAG) return A[s64idx(b, m - CISM_DIAG, q)];
        const int iA = m % 3 - 1, jA = (m / 3) % 3 - 1, kA = m / 9 - 1;
        const long long qn = q + kA + (long long)g.nz * (iA + (long long)g.sx * jA);
        const int bt = b == 1 ? 2 : (b == 2 ? 1 : b);
        return A[s64idx(bt, CISM_DIAG - m, qn)];
    }
};

template <typename T, typename M>
__device__ inline void row_product(const GridDev &g, const M &A, const T *__restrict__ x, long long q,
                                   int k, T &yu, T &yv) {
    using V2 = typename Vec2<T>::t;
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
                const V2 xx = *reinterpret_cast<const V2 *>(x + 2 * qn);
                T uu, uv, vu, vv;
                A.blocks(q, m, qn, uu, uv, vu, vv);
                su = su + uu * xx.x + uv * xx.y;
                sv = sv + vu * xx.x + vv * xx.y;
            }
    }
    yu = su;
    yv = sv;
}

template <typename T>
__device__ inline void store2(T *__restrict__ y, long long q, T a, T b) {
    using V2 = typename Vec2<T>::t;
    V2 v;
    v.x = a;
    v.y = b;
    *reinterpret_cast<V2 *>(y + 2 * q) = v;
}

// y = A x on owned active rows, 0 elsewhere (pcg:2533-2647)
template <typename T>
__global__ __launch_bounds__(CISM_BLOCK) void k_matvec(GridDev g, const T *__restrict__ A,
                                                       const int32_t *__restrict__ av,
                                                       const T *__restrict__ x, T *__restrict__ y) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nn) return;
    const int k = (int)(q % g.nz);
    const long long col = q / g.nz;
    const int i = (int)(col % g.sx), j = (int)(col / g.sx);
    T yu = 0, yv = 0;
    if (is_owned(g, i, j) && av[col]) row_product<T>(g, MatFull<T>{A}, x, q, k, yu, yv);
    store2<T>(y, q, yu, yv);
}

int launch_matvec64(cism_ctx *c, const double *A, const double *x, double *y) {
    hipLaunchKernelGGL((k_matvec<double>), dim3(nblocks_for(c->g.nn)), dim3(CISM_BLOCK), 0,
                       c->stream, c->g, A, c->av, x, y);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// y = A x plus per-block partial dot products over owned nodes:
//   NDOT = 1: s0 = (x, y)               [standard PCG (d, Ad)]
//   NDOT = 2: s0 = (r, x), s1 = (y, x)  [Chron-Gear (r, z), (Az, z)]
// One lane per node, one pass over the grid; block partials go to a second
// deterministic pass (k_reduce).
template <typename T, int NDOT, typename M>
__global__ __launch_bounds__(CISM_BLOCK) void k_matvec_dots(GridDev g, const M A,
                                                            const int32_t *__restrict__ av,
                                                            const T *__restrict__ x, T *__restrict__ y,
                                                            const T *__restrict__ r,
                                                            double *__restrict__ partials) {
    using V2 = typename Vec2<T>::t;
    double acc[NDOT];
#pragma unroll
    for (int c = 0; c < NDOT; ++c) acc[c] = 0.0;
    const long long lb = M::xcd ? xcd_block() : (long long)blockIdx.x;
    const long long q = lb * blockDim.x + threadIdx.x;
    if (q < g.nn) {
        const int k = (int)(q % g.nz);
        const long long col = q / g.nz;
        const int i = (int)(col % g.sx), j = (int)(col / g.sx);
        T yu = 0, yv = 0;
        const bool own = is_owned(g, i, j);
        if (own && av[col]) row_product<T>(g, A, x, q, k, yu, yv);
        store2<T>(y, q, yu, yv);
        if (own) {
            const V2 xx = *reinterpret_cast<const V2 *>(x + 2 * q);
            if (NDOT == 1) {
                acc[0] = (double)xx.x * (double)yu + (double)xx.y * (double)yv;
            } else {
                const V2 rr = *reinterpret_cast<const V2 *>(r + 2 * q);
                acc[0] = (double)rr.x * (double)xx.x + (double)rr.y * (double)xx.y;
                acc[NDOT - 1] = (double)yu * (double)xx.x + (double)yv * (double)xx.y;
            }
        }
    }
    // node_partials: the partials of node block lb sit where the plain launch puts them
    block_reduce_store<NDOT>(acc, partials, M::node_partials ? lb : -1);
}

// r = b - A x on the full grid (A x = 0 off owned active rows), s0 = (r, r) over
// owned nodes.  halo_r: the halo takes the periodic image of r afterwards
// (Chron-Gear, pcg:1592-1595), computed here from the image row directly.
template <typename T, typename M>
__global__ __launch_bounds__(CISM_BLOCK) void k_residual(GridDev g, const M A,
                                                         const int32_t *__restrict__ av,
                                                         const T *__restrict__ x,
                                                         const T *__restrict__ b, T *__restrict__ r,
                                                         int halo_r, double *__restrict__ partials) {
    double acc[1] = {0.0};
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < g.nn) {
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
            if (is_owned(g, i, j) && av[cs]) row_product<T>(g, A, x, qs, k, yu, yv);
            ru = b[2 * qs] - yu;
            rv = b[2 * qs + 1] - yv;
        }
        store2<T>(r, q, ru, rv);
        if (own) acc[0] = (double)ru * ru + (double)rv * rv;
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
            acc[0] += (double)a[2 * q] * (double)b[2 * q] + (double)a[2 * q + 1] * (double)b[2 * q + 1];
    }
    block_reduce_store<1>(acc, partials);
}

// ---------------------------------------------------------------------------
// preconditioners (setup_preconditioner_3d pcg:2150-2265 and the apply
// branches pcg:300-365 / 1416-1454; easy_sia_solver + tridiag_solver
// pcg:2759-2904).  Coefficients are read straight from the matrix slots.
// ---------------------------------------------------------------------------
template <typename T, typename M>
__global__ __launch_bounds__(CISM_BLOCK) void k_precond_diag(GridDev g, const M A,
                                                             const T *__restrict__ r,
                                                             T *__restrict__ z) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nn) return;
    const T du = A.entry(g, 0, CISM_DIAG, q), dv = A.entry(g, 3, CISM_DIAG, q);
    store2<T>(z, q, du != (T)0 ? r[2 * q] / du : (T)0, dv != (T)0 ? r[2 * q + 1] / dv : (T)0);
}

template <typename T>
__global__ __launch_bounds__(CISM_BLOCK) void k_copy(long long n, const T *__restrict__ a,
                                                     T *__restrict__ b) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) b[q] = a[q];
}

// Column-tridiagonal (SIA) preconditioner.  Setup, once per linear solve (the
// role of setup_preconditioner_3d): each (column, component) lane runs the
// forward part of the Thomas algorithm of tridiag_solver on the vertical
// couplings (slots indxA(0,0,-1|0|+1) of Auu or Avv) and stores, for every
// level k, the sub-diagonal sb(k), the pivot beta(k) and gamma(k) =
// super(k-1)/beta(k-1).  The factors are stored level-major across columns,
// F[((f*2 + comp)*nz + k)*ncol + col], so lanes of neighbouring columns read
// neighbouring values.  Apply, every iteration: the same forward/back
// substitution as tridiag_solver, with identical operations, on r staged
// through LDS.  Inactive columns keep z, as easy_sia_solver leaves them.
//
// An optional per-column scale s (cs[comp*ncol + col], s > 0) turns the
// preconditioner into z = s M^-1 r by storing sb/s and beta/s (gamma is a ratio
// and unchanged): still symmetric positive definite, and free in the apply.
template <typename T, typename M, int NZMAX>
__global__ __launch_bounds__(CISM_BLOCK) void k_sia_setup(GridDev g, const M A,
                                                          const int32_t *__restrict__ av,
                                                          const float *__restrict__ cs,
                                                          T *__restrict__ F) {
    const long long ncol = (long long)g.sx * g.sy;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * ncol) return;
    const int comp = (int)(t / ncol);
    const long long col = t - comp * ncol;
    if (!av[col]) return;
    const int nz = g.nz, blk = comp == 0 ? 0 : 3;
    const long long qb = col * nz;
    auto Fi = [&](int f, int k) { return ((long long)(f * 2 + comp) * nz + k) * ncol + col; };
    const T inv = cs ? (T)(1.0f / cs[t]) : (T)1;
    // the column's three diagonals, all loaded before the first store (F could alias
    // the matrix as far as the compiler knows, and each level would wait on its loads)
    T dg[NZMAX], up[NZMAX], lo[NZMAX];
#pragma unroll
    for (int k = 0; k < NZMAX; ++k)
        if (k < nz) {
            dg[k] = A.entry(g, blk, CISM_DIAG, qb + k);
            up[k] = k + 1 < nz ? A.entry(g, blk, CISM_DIAG + 9, qb + k) : (T)0;
            lo[k] = k > 0 ? A.entry(g, blk, CISM_DIAG - 9, qb + k) : (T)0;
        }
    T beta = dg[0];
    F[Fi(0, 0)] = (T)0;
    F[Fi(1, 0)] = cs ? beta * inv : beta;
    F[Fi(2, 0)] = (T)0;
#pragma unroll
    for (int k = 1; k < NZMAX; ++k)
        if (k < nz) {
            const T gam = up[k - 1] / beta;   // super(k-1) / beta(k-1)
            const T sb = lo[k];
            beta = dg[k] - sb * gam;
            F[Fi(0, k)] = cs ? sb * inv : sb;
            F[Fi(1, k)] = cs ? beta * inv : beta;
            F[Fi(2, k)] = gam;
        }
}

// k_sia_setup for the FP32 symmetric matrix: the same factors, from the diagonal
// slot and the slot above (whose entry at level k-1 is the coupling below level k)
// of 64 columns at a time, staged through LDS so that every global load is a
// contiguous 1 KB run over the block instead of one scattered word per level.
template <int NZMAX>
__global__ __launch_bounds__(128) void k_sia_setup_sym32(GridDev g, const float *__restrict__ As,
                                                         const int32_t *__restrict__ av,
                                                         const float *__restrict__ cs,
                                                         float *__restrict__ F) {
    const int nz = g.nz;
    const long long ncol = (long long)g.sx * g.sy;
    const long long col0 = (long long)blockIdx.x * 64;
    __shared__ float4 sd[64 * NZMAX], su[64 * NZMAX];
    const int ncols = (int)(col0 + 64 <= ncol ? 64 : ncol - col0);
    const long long Q0 = col0 * nz;
    const int nn = ncols * nz;
    for (int t = threadIdx.x; t < nn; t += blockDim.x) {
        sd[t] = *reinterpret_cast<const float4 *>(As + sidx(0, 0, Q0 + t));
        su[t] = *reinterpret_cast<const float4 *>(As + sidx(0, 9, Q0 + t));   // slot DIAG + 9
    }
    __syncthreads();
    const int comp = threadIdx.x >> 6, lc = threadIdx.x & 63;
    const long long col = col0 + lc;
    if (lc >= ncols || !av[col]) return;
    auto pick = [&](const float4 &v) { return comp == 0 ? v.x : v.w; };   // Auu or Avv
    auto Fi = [&](int f, int k) { return ((long long)(f * 2 + comp) * nz + k) * ncol + col; };
    const float inv = cs ? 1.0f / cs[comp * ncol + col] : 1.0f;
    const float4 *d = sd + lc * nz, *u = su + lc * nz;
    float beta = pick(d[0]);
    F[Fi(0, 0)] = 0.0f;
    F[Fi(1, 0)] = cs ? beta * inv : beta;
    F[Fi(2, 0)] = 0.0f;
    for (int k = 1; k < nz; ++k) {
        const float gam = pick(u[k - 1]) / beta;   // super(k-1) / beta(k-1)
        const float sb = pick(u[k - 1]);            // sub(k) = super(k-1), the symmetric pair
        beta = pick(d[k]) - sb * gam;
        F[Fi(0, k)] = cs ? sb * inv : sb;
        F[Fi(1, k)] = cs ? beta * inv : beta;
        F[Fi(2, k)] = gam;
    }
}

// One column and component of the SIA preconditioner: the Thomas substitutions
// on s (level k at s[2k]) with the factors precomputed by k_sia_setup.  For
// nz <= 12 the factors are first loaded into registers (sia_factors), so that a
// caller can issue those loads early; the substitutions would otherwise wait on
// one global load per level.
template <typename T, int NZMAX>
struct SiaFactors {
    T fs[NZMAX], fb[NZMAX], fg[NZMAX];
};

template <typename T, int NZMAX>
__device__ __forceinline__ void sia_factors(int nz, long long ncol, const T *__restrict__ F, long long col,
                                            int comp, SiaFactors<T, NZMAX> &f) {
    const T *sb = F + (long long)(0 * 2 + comp) * nz * ncol + col;
    const T *be = F + (long long)(1 * 2 + comp) * nz * ncol + col;
    const T *ga = F + (long long)(2 * 2 + comp) * nz * ncol + col;
#pragma unroll
    for (int k = 0; k < NZMAX; ++k)
        if (k < nz) {
            f.fs[k] = sb[k * ncol];
            f.fb[k] = be[k * ncol];
            f.fg[k] = ga[k * ncol];
        }
}

template <typename T, int NZMAX>
__device__ __forceinline__ void sia_solve(int nz, const SiaFactors<T, NZMAX> &f, T *s) {
    T prev = s[0] / f.fb[0];
    s[0] = prev;
#pragma unroll
    for (int k = 1; k < NZMAX; ++k)
        if (k < nz) {
            prev = (s[2 * k] - f.fs[k] * prev) / f.fb[k];
            s[2 * k] = prev;
        }
#pragma unroll
    for (int k = NZMAX - 2; k >= 0; --k)
        if (k <= nz - 2) {
            prev = s[2 * k] - f.fg[k + 1] * prev;
            s[2 * k] = prev;
        }
}

template <typename T, int NZMAX>
__device__ __forceinline__ void sia_column(int nz, long long ncol, const T *__restrict__ F, long long col,
                                           int comp, T *s) {
    if constexpr (NZMAX <= 12) {
        SiaFactors<T, NZMAX> f;
        sia_factors<T, NZMAX>(nz, ncol, F, col, comp, f);
        sia_solve<T, NZMAX>(nz, f, s);
    } else {
        const T *sb = F + (long long)(0 * 2 + comp) * nz * ncol + col;
        const T *be = F + (long long)(1 * 2 + comp) * nz * ncol + col;
        const T *ga = F + (long long)(2 * 2 + comp) * nz * ncol + col;
        T prev = s[0] / be[0];
        s[0] = prev;
        for (int k = 1; k < nz; ++k) {
            prev = (s[2 * k] - sb[k * ncol] * prev) / be[k * ncol];
            s[2 * k] = prev;
        }
        for (int k = nz - 2; k >= 0; --k) {
            prev = s[2 * k] - ga[(k + 1) * ncol] * prev;
            s[2 * k] = prev;
        }
    }
}

// z = M^-1 r, M the SIA (column tridiagonal) preconditioner: one block per 64
// columns, one lane per column and component; r staged in LDS for coalescing
// colsum (optional): the (u, v) sums of r over the levels of owned active columns, as the
// fused update leaves them (prescribed components left out with um/vm), for the coarse
// correction's restriction
template <typename T, int NZMAX>
__global__ __launch_bounds__(128) void k_sia_apply(GridDev g, const T *__restrict__ F,
                                                   const int32_t *__restrict__ av,
                                                   const T *__restrict__ r, T *__restrict__ z,
                                                   T *__restrict__ colsum = nullptr,
                                                   const int32_t *__restrict__ um = nullptr,
                                                   const int32_t *__restrict__ vm = nullptr) {
    const int nz = g.nz;
    const long long ncol = (long long)g.sx * g.sy;
    const long long col0 = (long long)blockIdx.x * 64;
    __shared__ T sr[2 * 64 * NZMAX];
    const int ncols = (int)(col0 + 64 <= ncol ? 64 : ncol - col0);
    const long long Q0 = col0 * nz;
    const int n2 = 2 * ncols * nz;
    for (int t = threadIdx.x; t < n2; t += blockDim.x) sr[t] = r[2 * Q0 + t];
    __syncthreads();
    const int comp = threadIdx.x >> 6, lc = threadIdx.x & 63;
    const long long col = col0 + lc;
    if (lc < ncols && av[col]) {
        if (colsum) {
            const int jc = (int)(col / g.sx), ic = (int)(col - (long long)jc * g.sx);
            if (is_owned(g, ic, jc)) {
                const int32_t *dm = comp == 0 ? um : vm;
                T cs = (T)0;
                for (int k = 0; k < nz; ++k)
                    if (!dm || dm[k + (long long)nz * col] != 1) cs += sr[2 * (lc * nz + k) + comp];
                colsum[2 * col + comp] = cs;
            }
        }
        sia_column<T, NZMAX>(nz, ncol, F, col, comp, sr + 2 * lc * nz + comp);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n2; t += blockDim.x)
        if (av[col0 + t / (2 * nz)]) z[2 * Q0 + t] = sr[t];
}

template <typename T, typename M>
static int launch_sia_setup(cism_ctx *c, const M A, const float *cs, T *F) {
    const long long ncol = (long long)c->g.sx * c->g.sy;
    if constexpr (std::is_same<M, MatSym32>::value) {
        const dim3 grid((unsigned)((ncol + 63) / 64));
        if (c->g.nz <= 12)
            hipLaunchKernelGGL((k_sia_setup_sym32<12>), grid, dim3(128), 0, c->stream, c->g, A.A, c->av, cs, F);
        else if (c->g.nz <= 24)
            hipLaunchKernelGGL((k_sia_setup_sym32<24>), grid, dim3(128), 0, c->stream, c->g, A.A, c->av, cs, F);
        else if (c->g.nz <= 48)
            hipLaunchKernelGGL((k_sia_setup_sym32<48>), grid, dim3(128), 0, c->stream, c->g, A.A, c->av, cs, F);
        else {
            cism_set_error("SIA preconditioner supports nz <= 48 (nz = %d)", c->g.nz);
            return CISM_ERR_UNSUPPORTED;
        }
        HIP_TRY(hipGetLastError());
        return CISM_OK;
    }
    const dim3 grid(nblocks_for(2 * ncol));
    if (c->g.nz <= 12)
        hipLaunchKernelGGL((k_sia_setup<T, M, 12>), grid, dim3(CISM_BLOCK), 0, c->stream, c->g, A, c->av, cs, F);
    else if (c->g.nz <= 24)
        hipLaunchKernelGGL((k_sia_setup<T, M, 24>), grid, dim3(CISM_BLOCK), 0, c->stream, c->g, A, c->av, cs, F);
    else if (c->g.nz <= 48)
        hipLaunchKernelGGL((k_sia_setup<T, M, 48>), grid, dim3(CISM_BLOCK), 0, c->stream, c->g, A, c->av, cs, F);
    else {
        cism_set_error("SIA preconditioner supports nz <= 48 (nz = %d)", c->g.nz);
        return CISM_ERR_UNSUPPORTED;
    }
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// precond 3 (FP32 symmetric matrix only): the SIA preconditioner plus the aggregation
// coarse correction (coarse.hip), whose V-cycle and prolongation follow the SIA apply or
// the fused update
#define PRECOND_SIA_COARSE 3

template <typename T, typename M>
static int launch_precond(cism_ctx *c, int precond, const M A, const T *F, const T *r, T *z) {
    const GridDev &g = c->g;
    if (precond == 0) {
        hipLaunchKernelGGL(k_copy<T>, dim3(nblocks_for(2 * g.nn)), dim3(CISM_BLOCK), 0, c->stream,
                           2 * g.nn, r, z);
    } else if (precond == 1) {
        hipLaunchKernelGGL((k_precond_diag<T, M>), dim3(nblocks_for(g.nn)), dim3(CISM_BLOCK), 0,
                           c->stream, g, A, r, z);
    } else {
        const long long ncol = (long long)g.sx * g.sy;
        const dim3 grid((unsigned)((ncol + 63) / 64));
        T *cs = nullptr;
        if constexpr (std::is_same<M, MatSym32>::value && std::is_same<T, float>::value)
            if (precond == PRECOND_SIA_COARSE) cs = c->coarse->colsum;
        const int32_t *um = cs && c->dirichlet_any ? c->umask : nullptr;
        const int32_t *vm = cs && c->dirichlet_any ? c->vmask : nullptr;
        if (g.nz <= 12)
            hipLaunchKernelGGL((k_sia_apply<T, 12>), grid, dim3(128), 0, c->stream, g, F, c->av, r, z, cs, um, vm);
        else if (g.nz <= 24)
            hipLaunchKernelGGL((k_sia_apply<T, 24>), grid, dim3(128), 0, c->stream, g, F, c->av, r, z, cs, um, vm);
        else if (g.nz <= 48)
            hipLaunchKernelGGL((k_sia_apply<T, 48>), grid, dim3(128), 0, c->stream, g, F, c->av, r, z, cs, um, vm);
        else {
            cism_set_error("SIA preconditioner supports nz <= 48 (nz = %d)", g.nz);
            return CISM_ERR_UNSUPPORTED;
        }
        HIP_TRY(hipGetLastError());
        if constexpr (std::is_same<M, MatSym32>::value && std::is_same<T, float>::value)
            if (precond == PRECOND_SIA_COARSE) CISM_TRY(coarse_apply(c, z));   // column sums: k_sia_apply
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

// node halo copies: NV values per node, stored per node with stride `vstride`
// (vectors: NV = 2 interleaved components) or through aidx (matrix: 108)
template <typename T, bool MAT>
__global__ __launch_bounds__(CISM_BLOCK) void k_halo(GridDev g, T *__restrict__ v) {
    const long long nh = (3LL * g.sx + 3LL * (g.sy - 3)) * g.nz;
    const int NV = MAT ? 108 : 2;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nh * NV) return;
    const int p = (int)(t / nh);           // value index (block*27+slot, or component)
    const long long hv = t % nh;
    const int k = (int)(hv % g.nz);
    int i, j;
    if (!halo_vertex_of(g, hv / g.nz, i, j)) return;
    const int ii = halo_image(i, g.sx, g.periodic), jj = halo_image(j, g.sy, g.periodic);
    const long long qd = k + (long long)g.nz * (i + (long long)g.sx * j);
    const long long qs = k + (long long)g.nz * (ii + (long long)g.sx * jj);
    const long long dst = MAT ? aidx(p / 27, p % 27, qd) : 2 * qd + p;
    const long long src = MAT ? aidx(p / 27, p % 27, qs) : 2 * qs + p;
    v[dst] = (ii < 0 || jj < 0) ? (T)0 : v[src];
}

template <typename T, bool MAT>
static int launch_halo_t(cism_ctx *c, T *v) {
    const GridDev &g = c->g;
    const long long nh = (3LL * g.sx + 3LL * (g.sy - 3)) * g.nz * (MAT ? 108 : 2);
    hipLaunchKernelGGL((k_halo<T, MAT>), dim3(nblocks_for(nh)), dim3(CISM_BLOCK), 0, c->stream, g, v);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// the FP32 symmetric rows (slots CISM_DIAG .. 26) of the halo nodes: their images' rows
// (zero without an image), as the FP64 halo update makes them, for an FP32 matrix written
// by the gather
__global__ __launch_bounds__(CISM_BLOCK) void k_a32_halo(GridDev g, float *__restrict__ As) {
    const long long nh = (3LL * g.sx + 3LL * (g.sy - 3)) * g.nz;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nh * CISM_SYM_SLOTS) return;
    const int s = (int)(t / nh);
    const long long hv = t % nh;
    const int k = (int)(hv % g.nz);
    int i, j;
    if (!halo_vertex_of(g, hv / g.nz, i, j)) return;
    const int ii = halo_image(i, g.sx, g.periodic), jj = halo_image(j, g.sy, g.periodic);
    const long long q = k + (long long)g.nz * (i + (long long)g.sx * j);
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (ii >= 0 && jj >= 0) v = *reinterpret_cast<const float4 *>(As + sidx(0, s, k + (long long)g.nz * (ii + (long long)g.sx * jj)));
    *reinterpret_cast<float4 *>(As + sidx(0, s, q)) = v;
}

int launch_a32_halo(cism_ctx *c) {
    const GridDev &g = c->g;
    const long long n = (3LL * g.sx + 3LL * (g.sy - 3)) * g.nz * CISM_SYM_SLOTS;
    hipLaunchKernelGGL(k_a32_halo, dim3(nblocks_for(n)), dim3(CISM_BLOCK), 0, c->stream, g, c->A32);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

int launch_halo_vec(cism_ctx *c, double *v) { return launch_halo_t<double, false>(c, v); }
int launch_halo_mat(cism_ctx *c, double *A) {
    if (A == c->A) c->a64s_valid = 0;
    return launch_halo_t<double, true>(c, A);
}

// ---------------------------------------------------------------------------
// vector updates (elementwise over the 2*nn interleaved entries)
// ---------------------------------------------------------------------------
// standard PCG: d = z + beta*d (pcg:383-393)
template <typename T>
__global__ __launch_bounds__(CISM_BLOCK) void k_std_d(long long n2, const T *__restrict__ z,
                                                      T *__restrict__ d,
                                                      const DevScal *__restrict__ sc) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n2) return;
    const T beta = (T)sc->beta;
    d[t] = z[t] + beta * d[t];
}

// standard PCG: x += alpha*d, r -= alpha*q (pcg:466-472)
template <typename T>
__global__ __launch_bounds__(CISM_BLOCK) void k_std_xr(long long n2, const T *__restrict__ d,
                                                       const T *__restrict__ qv, T *__restrict__ x,
                                                       T *__restrict__ r,
                                                       const DevScal *__restrict__ sc) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n2) return;
    const T alpha = (T)sc->alpha;
    x[t] = x[t] + alpha * d[t];
    r[t] = r[t] - alpha * qv[t];
}

// Chron-Gear (pcg:1517-1536) of node q with the halo update of Az (pcg:1494-1497)
// folded in: halo entries read Az at their periodic image (or 0, outflow).
// Returns 0 for a skipped node, 1 for an updated halo node, 2 for an updated owned one.
//
// skip_av (diagonal or SIA preconditioner only): owned nodes of inactive columns are
// left alone.  There z = 0 (zeroed at the start, never written by those
// preconditioners), A z = 0 and d = q = 0, so their update would rewrite d = q = 0
// and leave x and r unchanged: skipping it changes no value and saves the traffic.
template <typename T>
__device__ __forceinline__ int cg_update_node(const GridDev &g, long long q, int k, long long col, int i,
                                               int j, const T *__restrict__ z, const T *__restrict__ Az,
                                               T *__restrict__ d, T *__restrict__ qv, T *__restrict__ x,
                                               T *__restrict__ r, T alpha, T beta, int first,
                                               const int32_t *__restrict__ skip_av,
                                               typename Vec2<T>::t &rr) {
    // one lane per node (k, i, j) of column col: both components through one 8- or
    // 16-byte access per vector
    using V2 = typename Vec2<T>::t;
    V2 az;
    bool own;
    {
        own = is_owned(g, i, j);
        if (own && skip_av && !skip_av[col]) return 0;
        if (own) {
            az = reinterpret_cast<const V2 *>(Az)[q];
        } else {
            const int ii = halo_image(i, g.sx, g.periodic), jj = halo_image(j, g.sy, g.periodic);
            if (ii < 0 || jj < 0) {
                az.x = (T)0;
                az.y = (T)0;
            } else {
                az = reinterpret_cast<const V2 *>(Az)[k + (long long)g.nz * (ii + (long long)g.sx * jj)];
            }
        }
    }
    const V2 zz = reinterpret_cast<const V2 *>(z)[q];
    V2 dd, qq;
    if (first) {   // first pass: d = z, q = A d (already halo-mapped), pcg:1351-1405
        dd = zz;
        qq = az;
    } else {
        const V2 d0 = reinterpret_cast<const V2 *>(d)[q], q0 = reinterpret_cast<const V2 *>(qv)[q];
        dd.x = zz.x + beta * d0.x;
        dd.y = zz.y + beta * d0.y;
        qq.x = az.x + beta * q0.x;
        qq.y = az.y + beta * q0.y;
    }
    reinterpret_cast<V2 *>(d)[q] = dd;
    reinterpret_cast<V2 *>(qv)[q] = qq;
    V2 xx = reinterpret_cast<const V2 *>(x)[q];
    rr = reinterpret_cast<const V2 *>(r)[q];
    xx.x = xx.x + alpha * dd.x;
    xx.y = xx.y + alpha * dd.y;
    rr.x = rr.x - alpha * qq.x;
    rr.y = rr.y - alpha * qq.y;
    reinterpret_cast<V2 *>(x)[q] = xx;
    reinterpret_cast<V2 *>(r)[q] = rr;
    return own ? 2 : 1;
}

template <typename T>
__global__ __launch_bounds__(CISM_BLOCK) void k_cg_update(GridDev g, const T *__restrict__ z,
                                                          const T *__restrict__ Az,
                                                          T *__restrict__ d, T *__restrict__ qv,
                                                          T *__restrict__ x, T *__restrict__ r,
                                                          const DevScal *__restrict__ sc, int first,
                                                          const int32_t *__restrict__ skip_av) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nn) return;
    const int k = (int)(q % g.nz);
    const long long col = q / g.nz;
    const int i = (int)(col % g.sx), j = (int)(col / g.sx);
    typename Vec2<T>::t rr;
    cg_update_node<T>(g, q, k, col, i, j, z, Az, d, qv, x, r, (T)sc->alpha, (T)sc->beta, first, skip_av, rr);
}

// The Chron-Gear update followed by the SIA preconditioner of the next iteration
// (z = M^-1 r, pcg:1462-1470), for one block of 64 columns: the block updates its
// nodes, keeps the new r in LDS, solves its columns there and writes z.  The
// update of a node reads only its own z, so z is overwritten in place.  Same
// values as k_cg_update then k_sia_apply; one pass over r fewer.  The column
// factors are loaded before the update so their latency overlaps it.  With
// want_rr the block also leaves (r, r) over its owned nodes in partials (the
// convergence check of the inner solves, pcg:1540-1560).  With colsum the block also
// leaves the (u, v) sums of the new r over the levels of its owned active columns, the
// first stage of the coarse correction's restriction (coarse.hip).
template <typename T, int NZMAX>
__global__ __launch_bounds__(256) void k_cg_update_sia(GridDev g, T *z, const T *__restrict__ Az,
                                                       T *__restrict__ d, T *__restrict__ qv,
                                                       T *__restrict__ x, T *__restrict__ r,
                                                       const DevScal *__restrict__ sc, int first,
                                                       const int32_t *__restrict__ av,
                                                       const T *__restrict__ F, int want_rr,
                                                       double *__restrict__ partials, T *__restrict__ colsum,
                                                       const int32_t *__restrict__ um,
                                                       const int32_t *__restrict__ vm) {
    const int nz = g.nz;
    const long long ncol = (long long)g.sx * g.sy;
    const long long col0 = (long long)blockIdx.x * 64;
    __shared__ T sr[2 * 64 * NZMAX];
    const int ncols = (int)(col0 + 64 <= ncol ? 64 : ncol - col0);
    const long long Q0 = col0 * nz;
    const int nn = ncols * nz;
    // the update runs on all 256 lanes; the column solves on the first 128 (column, component)
    const int comp = (threadIdx.x >> 6) & 1, lc = threadIdx.x & 63;
    const long long col = col0 + lc;
    const bool solve = threadIdx.x < 128 && lc < ncols && av[col];
    SiaFactors<T, (NZMAX <= 12 ? NZMAX : 1)> f;
    if constexpr (NZMAX <= 12)
        if (solve) sia_factors<T, NZMAX>(nz, ncol, F, col, comp, f);
    const T alpha = (T)sc->alpha, beta = (T)sc->beta;
    double acc[1] = {0.0};
    // node t of the block: level t % nz of local column t / nz (32-bit: the context
    // limits grids to 2^30 nodes)
    const int c0 = (int)col0;
    for (int t = threadIdx.x; t < nn; t += blockDim.x) {
        const int lcol = t / nz;
        const int cn = c0 + lcol;
        const int jn = cn / g.sx;
        typename Vec2<T>::t rr;
        const int st = cg_update_node<T>(g, Q0 + t, t - lcol * nz, cn, cn - jn * g.sx, jn, z, Az, d, qv, x, r,
                                         alpha, beta, first, av, rr);
        if (st) {
            sr[2 * t] = rr.x;
            sr[2 * t + 1] = rr.y;
            if (st == 2) acc[0] += (double)rr.x * rr.x + (double)rr.y * rr.y;
        }
    }
    __syncthreads();
    if (solve) {
        if (colsum) {
            const int jc = (int)(col / g.sx), ic = (int)(col - (long long)jc * g.sx);
            if (is_owned(g, ic, jc)) {
                // prescribed (Dirichlet) components stay out of the coarse space
                const int32_t *dm = comp == 0 ? um : vm;
                T cs = (T)0;
                for (int k = 0; k < nz; ++k)
                    if (!dm || dm[k + (long long)nz * col] != 1) cs += sr[2 * (lc * nz + k) + comp];
                colsum[2 * col + comp] = cs;
            }
        }
        if constexpr (NZMAX <= 12) sia_solve<T, NZMAX>(nz, f, sr + 2 * lc * nz + comp);
        else sia_column<T, NZMAX>(nz, ncol, F, col, comp, sr + 2 * lc * nz + comp);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 2 * nn; t += blockDim.x)
        if (av[col0 + t / (2 * nz)]) z[2 * Q0 + t] = sr[t];
    if (want_rr) block_reduce_store<1>(acc, partials);
}

// ---------------------------------------------------------------------------
// PCG drivers
// ---------------------------------------------------------------------------
__global__ void k_init_scalars(DevScal *sc, int zero_stop) {
    sc->rho_old = sc->alpha = sc->beta = sc->sigma = sc->eta1 = sc->eta2 = 0.0;
    sc->eta0 = 1.0;
    sc->err = 0.0;
    sc->nan_flag = 0;
    sc->zero_stop = zero_stop;
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
    T *F;              // column factors of the SIA preconditioner (6*nn)
    const float *cs;   // optional per-column preconditioner scale (2*ncol) or null
};

// blocks of the fused update (one per 64 columns)
static inline int sia_blocks(const GridDev &g) { return (int)(((long long)g.sx * g.sy + 63) / 64); }

template <typename T>
static int launch_cg_update(cism_ctx *c, int precond, CGVecs<T> &v, int first, int want_rr = 0) {
    const GridDev &g = c->g;
    if (precond >= 2) {
        // fused with the SIA preconditioner; the caller skips its own precond call
        const dim3 grid((unsigned)sia_blocks(g));
        T *cs = nullptr;
        if constexpr (std::is_same<T, float>::value)
            if (precond == PRECOND_SIA_COARSE) cs = c->coarse->colsum;
        const int32_t *um = cs && c->dirichlet_any ? c->umask : nullptr;
        const int32_t *vm = cs && c->dirichlet_any ? c->vmask : nullptr;
        if (g.nz <= 12)
            hipLaunchKernelGGL((k_cg_update_sia<T, 12>), grid, dim3(256), 0, c->stream, g, v.z, v.Az, v.d,
                               v.q, v.x, v.r, c->sc, first, c->av, v.F, want_rr, c->partials, cs, um, vm);
        else if (g.nz <= 24)
            hipLaunchKernelGGL((k_cg_update_sia<T, 24>), grid, dim3(256), 0, c->stream, g, v.z, v.Az, v.d,
                               v.q, v.x, v.r, c->sc, first, c->av, v.F, want_rr, c->partials, cs, um, vm);
        else
            hipLaunchKernelGGL((k_cg_update_sia<T, 48>), grid, dim3(256), 0, c->stream, g, v.z, v.Az, v.d,
                               v.q, v.x, v.r, c->sc, first, c->av, v.F, want_rr, c->partials, cs, um, vm);
        HIP_TRY(hipGetLastError());
        if constexpr (std::is_same<T, float>::value)
            if (precond == PRECOND_SIA_COARSE) CISM_TRY(coarse_apply(c, v.z));
        return CISM_OK;
    } else {
        hipLaunchKernelGGL(k_cg_update<T>, dim3(nblocks_for(g.nn)), dim3(CISM_BLOCK), 0, c->stream, g, v.z,
                           v.Az, v.d, v.q, v.x, v.r, c->sc, first, precond != 0 ? c->av : nullptr);
    }
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// The reference iterations on a generic precision.  check = 0: every fifth
// iteration the true residual b - A x decides convergence against tol (the
// reference rule, pcg:478-537, 1542-1597).  check = 1 (inner solves of the
// mixed-precision refinement): the recursive residual r is measured instead,
// with no extra matvec, against tol * |r0|.
template <typename T, typename M>
static int run_cg(cism_ctx *c, const M A, CGVecs<T> v, int chrongear, int precond, int maxiters,
                  double tol, int check, double *err_out, int *niters_out, int x_zero = 0) {
    const GridDev &g = c->g;
    const int nb = red_blocks_for(g.nn), nbm = nblocks_for(g.nn), nb2 = nblocks_for(2 * g.nn);
    const int nbv = M::xcd ? xcd_blocks_for(g.nn) : nbm;   // matvec grid
    // partials of a matvec: one per launched block, or per node block (node_partials; the
    // padding blocks of an XCD grid, past the nodes, are left out of the sums)
    const int nbr = M::xcd && !M::node_partials ? nbv : nbm;
    const long long np = g.npad;
    // an inner solve (check = 1) whose residual becomes exactly zero has converged; the
    // reference recurrences would divide 0 by 0 there
    hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(1), 0, c->stream, c->sc, check);
    if (precond == PRECOND_SIA_COARSE && !std::is_same<M, MatSym32>::value) {
        cism_set_error("the coarse correction is a preconditioner of the FP32 inner solves only");
        return CISM_ERR_ARG;
    }
    if (precond >= 2) {
        TimerRange r_p("pcg_precond_init", c->stream, true);
        if constexpr (std::is_same<M, MatSym32>::value && std::is_same<T, float>::value) {
            // with the coarse correction its setup also forms the SIA factors (in the pass
            // that stages the matrix's slots per column anyway), once per FP32 matrix
            if (precond == PRECOND_SIA_COARSE) CISM_TRY(coarse_setup(c, A.A, v.F, v.cs));
            else CISM_TRY((launch_sia_setup<T, M>(c, A, v.cs, v.F)));
        } else {
            CISM_TRY((launch_sia_setup<T, M>(c, A, v.cs, v.F)));
        }
    }
    if (!x_zero) {   // with x_zero the caller has zeroed z and formed L2_rhs (k_rhs32)
        hipLaunchKernelGGL(k_zero2<T>, dim3(nblocks_for(2 * np)), dim3(CISM_BLOCK), 0, c->stream, 2 * np,
                           v.d, v.z, v.q);
        // (b, b) -> L2_rhs
        hipLaunchKernelGGL(k_dot<T>, dim3(nb), dim3(CISM_BLOCK), 0, c->stream, g, v.b, v.b, c->partials);
        CISM_TRY(launch_reduce(c, nb, 1, RED_BB));
    }
    int niters = maxiters;
    double err = 0.0;
    auto check_conv = [&](int n, bool &done) -> int {
        // the true residual every fifth iteration, up to the synchronization that reads it
        TimerRange r_c("pcg_matmult_resid", c->stream, true);
        if (check == 0) {
            hipLaunchKernelGGL((k_residual<T, M>), dim3(nbm), dim3(CISM_BLOCK), 0, c->stream, g, A,
                               c->av, v.x, v.b, v.r, chrongear, c->partials);
            CISM_TRY(launch_reduce(c, nbm, 1, RED_ERR));
        } else if (chrongear && precond >= 2) {
            // (r, r) came with the fused update of this iteration
            CISM_TRY(launch_reduce(c, sia_blocks(g), 1, RED_ERR));
        } else {
            hipLaunchKernelGGL(k_dot<T>, dim3(nb), dim3(CISM_BLOCK), 0, c->stream, g, v.r, v.r,
                               c->partials);
            CISM_TRY(launch_reduce(c, nb, 1, RED_ERR));
        }
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
        CISM_TRY((launch_halo_t<T, false>(c, v.x)));
        hipLaunchKernelGGL((k_residual<T, M>), dim3(nbm), dim3(CISM_BLOCK), 0, c->stream, g, A, c->av,
                           v.x, v.b, v.r, 0, c->partials);
        for (int n = 1; n <= maxiters && !done; ++n) {
            CISM_TRY((launch_precond<T, M>(c, precond, A, v.F, v.r, v.z)));
            hipLaunchKernelGGL(k_dot<T>, dim3(nb), dim3(CISM_BLOCK), 0, c->stream, g, v.r, v.z,
                               c->partials);
            CISM_TRY(launch_reduce(c, nb, 1, RED_STD_ETA1));
            hipLaunchKernelGGL(k_std_d<T>, dim3(nb2), dim3(CISM_BLOCK), 0, c->stream, 2 * g.nn, v.z,
                               v.d, c->sc);
            CISM_TRY((launch_halo_t<T, false>(c, v.d)));
            hipLaunchKernelGGL((k_matvec_dots<T, 1, M>), dim3(nbv), dim3(CISM_BLOCK), 0, c->stream, g,
                               A, c->av, v.d, v.q, (const T *)nullptr, c->partials);
            CISM_TRY(launch_reduce(c, nbr, 1, RED_STD_ETA2));
            hipLaunchKernelGGL(k_std_xr<T>, dim3(nb2), dim3(CISM_BLOCK), 0, c->stream, 2 * g.nn, v.d,
                               v.q, v.x, v.r, c->sc);
            if (n % 5 == 0) {
                if (check == 0) CISM_TRY((launch_halo_t<T, false>(c, v.x)));
                CISM_TRY(check_conv(n, done));
            }
        }
    } else {
        // pcg:1258-1405: halo x; r = b - A x; halo r; z = M^-1 r; d = z; q = A d.  With
        // x_zero the caller has set x = 0 and r = b (halo entries included).
        if (!x_zero) {
            CISM_TRY((launch_halo_t<T, false>(c, v.x)));
            hipLaunchKernelGGL((k_residual<T, M>), dim3(nbm), dim3(CISM_BLOCK), 0, c->stream, g, A, c->av,
                               v.x, v.b, v.r, 1, c->partials);
        }
        CISM_TRY((launch_precond<T, M>(c, precond, A, v.F, v.r, v.z)));
        hipLaunchKernelGGL((k_matvec_dots<T, 2, M>), dim3(nbv), dim3(CISM_BLOCK), 0, c->stream, g, A,
                           c->av, v.z, v.Az, v.r, c->partials);
        CISM_TRY(launch_reduce(c, nbr, 2, RED_CG_FIRST));
        // with the SIA preconditioner the update also forms the next z
        CISM_TRY(launch_cg_update<T>(c, precond, v, 1));
        for (int n = 2; n <= maxiters && !done; ++n) {
            if (precond < 2) CISM_TRY((launch_precond<T, M>(c, precond, A, v.F, v.r, v.z)));
            hipLaunchKernelGGL((k_matvec_dots<T, 2, M>), dim3(nbv), dim3(CISM_BLOCK), 0, c->stream,
                               g, A, c->av, v.z, v.Az, v.r, c->partials);
            CISM_TRY(launch_reduce(c, nbr, 2, RED_CG_ITER));
            CISM_TRY(launch_cg_update<T>(c, precond, v, 0, check == 1 && n % 5 == 0));
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

// Solves A x = b for c->x (initial guess in place) with the FP64 matrix c->A, read from its
// symmetric copy when the last symmetry pass left it exactly symmetric (the same iterates)
int solver_pcg64(cism_ctx *c, int chrongear, int precond, int maxiters, double tol, double *err,
                 int *niters) {
    CGVecs<double> v = {c->x, c->b, c->r, c->d, c->z, c->q, c->Az, c->F64, nullptr};
    if (c->a64s_valid && c->sym64_on)
        return run_cg<double>(c, MatSym64<CISM_SYM64_XCD>{c->A64s}, v, chrongear, precond, maxiters, tol, 0, err,
                              niters);
    return run_cg<double>(c, MatFull<double>{c->A}, v, chrongear, precond, maxiters, tol, 0, err, niters);
}

int ensure_a64s(cism_ctx *c) {
    if (c->A64s) return CISM_OK;
    hipError_t e = hipMalloc(&c->A64s, sizeof(double) * 4 * CISM_SYM_SLOTS * (size_t)c->g.npad);
    if (e != hipSuccess) {   // not an error of the solve: the full storage serves
        c->A64s = nullptr;
        (void)hipGetLastError();
        return CISM_ERR_NOMEM;
    }
    return CISM_OK;
}

// ---------------------------------------------------------------------------
// mixed precision: FP32 matrix and inner Chron-Gear, FP64 refinement
// ---------------------------------------------------------------------------
int ensure_a32(cism_ctx *c) {
    if (c->A32) return CISM_OK;
    const size_t v2 = 2 * (size_t)c->g.npad;
    hipError_t e = hipMalloc(&c->A32, sizeof(float) * 4 * CISM_SYM_SLOTS * (size_t)c->g.npad);
    float **vs[] = {&c->x32, &c->b32, &c->r32, &c->d32, &c->z32, &c->q32, &c->Az32};
    for (float **p : vs)
        if (e == hipSuccess) e = hipMalloc(p, sizeof(float) * v2);
    if (e == hipSuccess) e = hipMalloc(&c->F32, sizeof(float) * 6 * (size_t)c->g.npad);
    if (e != hipSuccess) {
        cism_set_error("FP32 matrix allocation failed: %s", hipGetErrorString(e));
        return CISM_ERR_NOMEM;
    }
    return CISM_OK;
}

// b32 = r64 / s (all entries), and the inner solve's start: x32 = 0, so its first
// residual is b32 itself (halo entries included: r64 carries its halo images)
// Also z32 = 0 (the preconditioner writes z on active columns only, and the matvec reads
// it around them; d and q are written before they are read) and, per block, (b32, b32)
// over owned nodes: the inner solve's |b| (RED_BB), so that run_cg with x_zero needs
// neither its zeroing pass nor its dot product.
__global__ __launch_bounds__(CISM_BLOCK) void k_rhs32(GridDev g, long long n2, const double *__restrict__ r,
                                                      const DevScal *__restrict__ sc,
                                                      float *__restrict__ b32, float *__restrict__ x32,
                                                      float *__restrict__ r32, float *__restrict__ z32,
                                                      double *__restrict__ partials) {
    double acc[1] = {0.0};
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n2) {
        const double s = sc->sums[2];
        const float v = (float)(s > 0.0 ? r[q] / s : 0.0);
        b32[q] = v;
        r32[q] = v;
        x32[q] = 0.0f;
        z32[q] = 0.0f;
        const long long node = q >> 1;
        if (node < g.nn) {
            const long long col = node / g.nz;
            if (is_owned(g, (int)(col % g.sx), (int)(col / g.sx))) acc[0] = (double)v * v;
        }
    }
    block_reduce_store<1>(acc, partials);
}

// x += s * e (all entries)
__global__ __launch_bounds__(CISM_BLOCK) void k_refine(long long n2, const float *__restrict__ e,
                                                       const DevScal *__restrict__ sc,
                                                       double *__restrict__ x) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n2) return;
    x[q] = x[q] + sc->sums[2] * (double)e[q];
}

__global__ void k_save_scale(DevScal *sc, int from) { sc->sums[2] = sqrt(sc->sums[from]); }

// Iterative refinement: r = b - A64 x (FP64), solve A32 e = r/|r| with FP32
// Chron-Gear to relative tolerance inner_rtol, x += |r| e, until the FP64
// relative residual is below tol.  The FP32 matrix must be current: the symmetry pass
// of the Picard loop writes it (symmetry_device with c->A32).  With r_ready the
// first residual, its norm and |b| come from the Picard residual pass.  *stalled is set
// when the refinement is given up (refine_policy.h): the caller then falls back to the
// FP64 solver; an FP32 breakdown (alpha = NaN) is such a stall, not an error.
// *unconverged is set when the budget ran out first (an inexact solve, as the
// reference's PCG ends at maxiters).
int solver_mixed(cism_ctx *c, int precond, int maxiters, double tol, double inner_rtol, double *err_out,
                 int *niters_out, int *n64, int *n32, int r_ready, int *stalled, int *unconverged,
                 int skip_final) {
    const GridDev &g = c->g;
    const int nb = red_blocks_for(g.nn);
    const long long n2 = 2 * g.npad;
    CISM_TRY(ensure_a32(c));
    double bnorm;
    if (r_ready) {
        bnorm = sqrt(c->sc_host->sums[3]);
    } else {
        hipLaunchKernelGGL(k_dot<double>, dim3(nb), dim3(CISM_BLOCK), 0, c->stream, g, c->b, c->b,
                           c->partials);
        CISM_TRY(launch_reduce(c, nb, 1, RED_SUMS));
        CISM_TRY(sync_scalars(c));
        bnorm = sqrt(c->sc_host->sums[0]);
    }
    int total = 0, outer = 0;
    double err = 0.0;
    CGVecs<float> v = {c->x32, c->b32, c->r32, c->d32, c->z32, c->q32, c->Az32, c->F32,
                       c->cscale_on ? c->cscale : nullptr};
    RefineMonitor mon;
    refine_monitor_init(&mon);
    int inner_conv = 1;
    if (stalled) *stalled = 0;
    if (unconverged) *unconverged = 0;
    while (true) {
        double rr;
        if (r_ready && outer == 0) {
            rr = c->sc_host->sums[2];
            hipLaunchKernelGGL(k_save_scale, dim3(1), dim3(1), 0, c->stream, c->sc, 2);
        } else {
            CISM_TRY(launch_halo_vec(c, c->x));
            if (c->a64_valid) {
                hipLaunchKernelGGL((k_residual<double, MatFull<double>>), dim3(nblocks_for(g.nn)),
                                   dim3(CISM_BLOCK), 0, c->stream, g, MatFull<double>{c->A}, c->av, c->x, c->b,
                                   c->r, 1, c->partials);
                CISM_TRY(launch_reduce(c, nblocks_for(g.nn), 1, RED_SUMS));
                hipLaunchKernelGGL(k_save_scale, dim3(1), dim3(1), 0, c->stream, c->sc, 0);
                CISM_TRY(sync_scalars(c));
                rr = c->sc_host->sums[0];
            } else {
                // the assembly kept no FP64 rows: the same residual from its element matrices
                CISM_TRY(residual_from_elements(c));
                hipLaunchKernelGGL(k_save_scale, dim3(1), dim3(1), 0, c->stream, c->sc, 0);
                CISM_TRY(sync_scalars(c));
                rr = c->sc_host->sums[0];
            }
        }
        ++*n64;
        err = bnorm > 0 ? sqrt(rr) / bnorm : 0.0;
        const int verdict = refine_decide(&mon, err, tol, outer, total, maxiters, inner_conv);
        if (verdict == REFINE_CONVERGED) break;
        if (verdict == REFINE_STALLED) {
            if (stalled) *stalled = 1;
            break;
        }
        if (verdict == REFINE_EXHAUSTED) {
            if (unconverged) *unconverged = 1;
            break;
        }
        hipLaunchKernelGGL(k_rhs32, dim3(nblocks_for(n2)), dim3(CISM_BLOCK), 0, c->stream, g, n2, c->r, c->sc,
                           c->b32, c->x32, c->r32, c->z32, c->partials);
        CISM_TRY(launch_reduce(c, nblocks_for(n2), 1, RED_BB));
        double e_in = 0.0;
        int it = 0;
        // the inner reductions use sums[0..1]; the scale |r| stays in sums[2]
        const int st_in = run_cg<float>(c, MatSym32{c->A32}, v, 1, precond, maxiters, inner_rtol, 1, &e_in, &it, 1);
        if (st_in == CISM_ERR_NAN && stalled) {
            // the FP32 iteration broke down; x is untouched by this outer step
            *stalled = 1;
            *n32 += it;
            total += it;
            break;
        }
        CISM_TRY(st_in);
        inner_conv = it < maxiters;
        hipLaunchKernelGGL(k_refine, dim3(nblocks_for(n2)), dim3(CISM_BLOCK), 0, c->stream, n2, c->x32, c->sc,
                           c->x);
        HIP_TRY(hipGetLastError());
        total += it;
        *n32 += it;
        ++outer;
        // skip_final: the first step's FP32 solve reached e_in (its recursive residual
        // relative to |r|), so the FP64 relative residual is about err * e_in; when that is
        // below half the tolerance and e_in is far above FP32's rounding floor, the FP64
        // residual that would confirm it is left to the next Picard iteration's residual
        if (skip_final == 1 && outer == 1 && inner_conv && e_in >= 1.0e-4 && err * e_in <= 0.5 * tol) {
            err = err * e_in;
            break;
        }
    }
    *err_out = err;
    *niters_out = total;
    return CISM_OK;
}

// ---------------------------------------------------------------------------
// per-column preconditioner scale of a solved system: for each owned active
// column and component, s = (x_c . M_c x_c) / (x_c . (A x)_c), the ratio of the
// energy of the column's vertical (SIA) block to that of the full operator on
// the solution (A x = b - r); s < 1 where horizontal coupling stiffens the
// column.  Clamped to [1/16, 16]; 1 elsewhere.  Out: s[comp*ncol + col].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(CISM_BLOCK) void k_colscale(GridDev g, const double *__restrict__ A,
                                                         const int32_t *__restrict__ av,
                                                         const double *__restrict__ x,
                                                         const double *__restrict__ b,
                                                         const double *__restrict__ r,
                                                         double *__restrict__ out) {
    const long long ncol = (long long)g.sx * g.sy;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * ncol) return;
    const int comp = (int)(t / ncol);
    const long long col = t - comp * ncol;
    const int i = (int)(col % g.sx), j = (int)(col / g.sx);
    double s = 1.0;
    if (is_owned(g, i, j) && av[col]) {
        const int blk = comp == 0 ? 0 : 3;
        const long long qb = col * g.nz;
        double num = 0.0, den = 0.0;
        for (int k = 0; k < g.nz; ++k) {
            const long long q = qb + k;
            double mx = A[aidx(blk, CISM_DIAG, q)] * x[2 * q + comp];
            if (k > 0) mx += A[aidx(blk, CISM_DIAG - 9, q)] * x[2 * (q - 1) + comp];
            if (k < g.nz - 1) mx += A[aidx(blk, CISM_DIAG + 9, q)] * x[2 * (q + 1) + comp];
            num += x[2 * q + comp] * mx;
            den += x[2 * q + comp] * (b[2 * q + comp] - r[2 * q + comp]);
        }
        if (den > 0.0 && num > 0.0) s = fmin(16.0, fmax(1.0 / 16.0, num / den));
    }
    out[t] = s;
}

// Energy of an initial guess in the first Picard system A x = b (A symmetric positive
// definite after the Dirichlet elimination): 2 J(x) = x.A x - 2 b.x = x.(A x - b) - b.x,
// with t = A x - b from the residual pass.  J(0) = 0 and 2 (J(x) - J(x*)) is the squared
// A-norm error, so J(x) < 0 exactly when x is closer to that system's solution than the
// zero guess is, in the A-norm.  Owned active rows; sums[0] = x.t, sums[1] = b.x.
__global__ __launch_bounds__(CISM_BLOCK) void k_energy(GridDev g, const int32_t *__restrict__ av,
                                                       const double *__restrict__ x,
                                                       const double *__restrict__ t,
                                                       const double *__restrict__ b,
                                                       double *__restrict__ partials) {
    double acc[2] = {0.0, 0.0};
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < g.nn; q += stride) {
        const long long col = q / g.nz;
        const int i = (int)(col % g.sx), j = (int)(col / g.sx);
        if (is_owned(g, i, j) && av[col]) {
            acc[0] += x[2 * q] * t[2 * q] + x[2 * q + 1] * t[2 * q + 1];
            acc[1] += b[2 * q] * x[2 * q] + b[2 * q + 1] * x[2 * q + 1];
        }
    }
    block_reduce_store<2>(acc, partials);
}

int guess_energy_device(cism_ctx *c, const double *t, double *twoJ) {
    const GridDev &g = c->g;
    const int nb = red_blocks_for(g.nn);
    hipLaunchKernelGGL(k_energy, dim3(nb), dim3(CISM_BLOCK), 0, c->stream, g, c->av, c->x, t, c->b,
                       c->partials);
    CISM_TRY(launch_reduce(c, nb, 2, RED_SUMS));
    CISM_TRY(sync_scalars(c));
    *twoJ = c->sc_host->sums[0] - c->sc_host->sums[1];
    return CISM_OK;
}

int column_scale_device(cism_ctx *c, double *out) {
    const GridDev &g = c->g;
    if (!c->a64_valid) {
        cism_set_error("no FP64 matrix: the last mixed-precision solve (skip_final_residual) kept none");
        return CISM_ERR_ARG;
    }
    const long long ncol = (long long)g.sx * g.sy;
    CISM_TRY(launch_halo_vec(c, c->x));
    hipLaunchKernelGGL((k_residual<double, MatFull<double>>), dim3(nblocks_for(g.nn)), dim3(CISM_BLOCK), 0,
                       c->stream, g, MatFull<double>{c->A}, c->av, c->x, c->b, c->r, 1, c->partials);
    hipLaunchKernelGGL(k_colscale, dim3(nblocks_for(2 * ncol)), dim3(CISM_BLOCK), 0, c->stream, g, c->A,
                       c->av, c->x, c->b, c->r, out);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// ---------------------------------------------------------------------------
// Anderson acceleration (type II, Walker & Ni 2011) of the Picard fixed-point
// map x -> G(x), G(x) = the solution of A(x) y = b(x).  With f_n = G(x_n) - x_n
// and the last k <= m differences dX_i = x_{i+1} - x_i, dF_i = f_{i+1} - f_i:
//   gamma = argmin |f_n - dF gamma|_2,   x_{n+1} = G(x_n) - (dX + dF) gamma.
// The fixed point is the Picard fixed point; only the path to it changes.
// Buffers (each 2*npad doubles): dX[0..m), dF[0..m), xprev, fprev.
// ---------------------------------------------------------------------------
// Only owned nodes of active columns take part.  Elsewhere neither the linear
// solves nor these updates change x (halo entries are refreshed from their images
// afterwards), so f, dX and dF are zero there and the history is not kept.
__device__ inline bool aa_entry_active(const GridDev &g, const int32_t *__restrict__ av, long long t) {
    const long long q = t >> 1;
    if (q >= g.nn) return false;
    const long long col = q / g.nz;
    const int i = (int)(col % g.sx), j = (int)(col / g.sx);
    return is_owned(g, i, j) && av[col];
}

__global__ __launch_bounds__(CISM_BLOCK) void k_aa_form(GridDev g, const int32_t *__restrict__ av,
                                                        const double *__restrict__ x,
                                                        const double *__restrict__ xn,
                                                        double *__restrict__ dX, double *__restrict__ dF,
                                                        double *__restrict__ xprev,
                                                        double *__restrict__ fprev, int add) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!aa_entry_active(g, av, t)) return;
    const double f = x[t] - xn[t];
    if (add) {
        dF[t] = f - fprev[t];
        dX[t] = xn[t] - xprev[t];
    }
    fprev[t] = f;
    xprev[t] = xn[t];
}

// Gram matrix of the k columns of dF (upper triangle, row-major), then dF_i . f,
// then f . f; owned nodes only.  NV = 4*5/2 + 4 + 1 = 15 slots, unused ones 0.
#define AA_NV 15
__global__ __launch_bounds__(CISM_BLOCK) void k_aa_dots(GridDev g, const int32_t *__restrict__ av, int k,
                                                        const double *__restrict__ dF, long long stride,
                                                        const double *__restrict__ f,
                                                        double *__restrict__ partials) {
    // grid-stride over nodes, both components of a column entry through one 16-byte load
    double acc[AA_NV];
#pragma unroll
    for (int c = 0; c < AA_NV; ++c) acc[c] = 0.0;
    const long long gs = (long long)gridDim.x * blockDim.x;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < g.nn; q += gs) {
        const long long col = q / g.nz;
        const int i = (int)(col % g.sx), j = (int)(col / g.sx);
        if (!is_owned(g, i, j) || !av[col]) continue;
        double2 v2[AA_MAX];
#pragma unroll
        for (int a = 0; a < AA_MAX; ++a)
            v2[a] = a < k ? *reinterpret_cast<const double2 *>(dF + a * stride + 2 * q) : make_double2(0.0, 0.0);
        const double2 f2 = *reinterpret_cast<const double2 *>(f + 2 * q);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            double v[AA_MAX];
#pragma unroll
            for (int a = 0; a < AA_MAX; ++a) v[a] = h ? v2[a].y : v2[a].x;
            const double fv = h ? f2.y : f2.x;
            int s = 0;
#pragma unroll
            for (int a = 0; a < AA_MAX; ++a)
#pragma unroll
                for (int b = a; b < AA_MAX; ++b) acc[s++] += v[a] * v[b];
#pragma unroll
            for (int a = 0; a < AA_MAX; ++a) acc[s++] += v[a] * fv;
            acc[s] += fv * fv;
        }
    }
    block_reduce_store<AA_NV>(acc, partials);
}

__global__ __launch_bounds__(CISM_BLOCK) void k_aa_update(GridDev g, const int32_t *__restrict__ av, int k,
                                                          const double *__restrict__ dX,
                                                          const double *__restrict__ dF, long long stride,
                                                          double g0, double g1, double g2, double g3,
                                                          double *__restrict__ x) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!aa_entry_active(g, av, t)) return;
    const double gm[AA_MAX] = {g0, g1, g2, g3};
    double v = x[t];
    for (int a = 0; a < k; ++a) v -= gm[a] * (dX[a * stride + t] + dF[a * stride + t]);
    x[t] = v;
}

int anderson_reset(cism_ctx *c) {
    c->aa_count = 0;
    c->aa_head = 0;
    c->aa_fnorm_prev = 0.0;
    return CISM_OK;
}

// Solves the k x k system G gamma = h by Cholesky with a relative Tikhonov term.
static bool small_lsq(int k, const double *Gu, const double *h, double *gamma) {
    double G[AA_MAX][AA_MAX], L[AA_MAX][AA_MAX] = {};
    int s = 0;
    double tr = 0.0;
    for (int a = 0; a < AA_MAX; ++a)
        for (int b = a; b < AA_MAX; ++b, ++s) G[a][b] = G[b][a] = Gu[s];
    for (int a = 0; a < k; ++a) tr += G[a][a];
    if (!(tr > 0.0)) return false;
    for (int a = 0; a < k; ++a) G[a][a] += 1.0e-10 * tr;
    for (int a = 0; a < k; ++a) {
        for (int b = 0; b <= a; ++b) {
            double v = G[a][b];
            for (int q = 0; q < b; ++q) v -= L[a][q] * L[b][q];
            if (a == b) {
                if (!(v > 0.0)) return false;
                L[a][a] = sqrt(v);
            } else {
                L[a][b] = v / L[b][b];
            }
        }
    }
    double y[AA_MAX];
    for (int a = 0; a < k; ++a) {
        double v = h[a];
        for (int q = 0; q < a; ++q) v -= L[a][q] * y[q];
        y[a] = v / L[a][a];
    }
    for (int a = k - 1; a >= 0; --a) {
        double v = y[a];
        for (int q = a + 1; q < k; ++q) v -= L[q][a] * gamma[q];
        gamma[a] = v / L[a][a];
    }
    return true;
}

int anderson_step(cism_ctx *c, int m, double *fnorm) {
    const GridDev &g = c->g;
    const long long n2 = 2 * g.npad;
    if (m < 1 || m > AA_MAX) {
        cism_set_error("Anderson depth must be 1..%d", AA_MAX);
        return CISM_ERR_ARG;
    }
    if (!c->aa || c->aa_m != m) {
        if (c->aa) HIP_TRY(hipFree(c->aa));
        c->aa = nullptr;
        hipError_t e = hipMalloc(&c->aa, sizeof(double) * (size_t)(2 * m + 2) * n2);
        if (e != hipSuccess) {
            cism_set_error("Anderson history allocation failed: %s", hipGetErrorString(e));
            return CISM_ERR_NOMEM;
        }
        if (!c->aa_sums) HIP_TRY(hipMalloc(&c->aa_sums, sizeof(double) * 32));
        c->aa_m = m;
        anderson_reset(c);
    }
    double *dX = c->aa, *dF = c->aa + (size_t)m * n2, *xprev = c->aa + (size_t)2 * m * n2,
           *fprev = xprev + n2;
    const int add = c->aa_count > 0;
    const int slot = c->aa_head;
    hipLaunchKernelGGL(k_aa_form, dim3(nblocks_for(n2)), dim3(CISM_BLOCK), 0, c->stream, g, c->av, c->x,
                       c->usav, dX + (size_t)slot * n2, dF + (size_t)slot * n2, xprev, fprev, add);
    const int k = add ? (c->aa_count < m ? c->aa_count : m) : 0;
    if (add) c->aa_head = (c->aa_head + 1) % m;
    c->aa_count++;
    // dots: Gram of the k stored columns, their products with f (= fprev now), |f|^2
    const int nb = red_blocks_for(g.nn);
    hipLaunchKernelGGL(k_aa_dots, dim3(nb), dim3(CISM_BLOCK), 0, c->stream, g, c->av, k, dF, n2, fprev,
                       c->partials);
    CISM_TRY(launch_reduce_to<AA_NV>(c, nb, c->aa_sums));
    HIP_TRY(hipGetLastError());
    double hs[AA_NV];
    HIP_TRY(hipMemcpyAsync(hs, c->aa_sums, sizeof hs, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const double fn = sqrt(hs[AA_NV - 1]);
    *fnorm = fn;
    // safeguard: a growing fixed-point residual restarts the history (plain Picard step)
    if (k == 0 || (c->aa_fnorm_prev > 0.0 && fn > 2.0 * c->aa_fnorm_prev)) {
        if (k > 0) {
            anderson_reset(c);
            c->aa_count = 1;   // the current (x, f) stays as the base of the next difference
        }
        c->aa_fnorm_prev = fn;
        return CISM_OK;
    }
    c->aa_fnorm_prev = fn;
    // columns are stored in ring order; the least-squares problem does not depend on it
    double gam[AA_MAX] = {0.0, 0.0, 0.0, 0.0};
    double Gu[AA_NV], h[AA_MAX];
    int s = 0;
    for (int a = 0; a < AA_MAX; ++a)
        for (int b = a; b < AA_MAX; ++b, ++s) Gu[s] = hs[s];
    for (int a = 0; a < AA_MAX; ++a) h[a] = hs[10 + a];
    if (!small_lsq(k, Gu, h, gam)) return CISM_OK;   // degenerate history: plain step
    hipLaunchKernelGGL(k_aa_update, dim3(nblocks_for(n2)), dim3(CISM_BLOCK), 0, c->stream, g, c->av, k, dX, dF,
                       n2, gam[0], gam[1], gam[2], gam[3], c->x);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// ---------------------------------------------------------------------------
// measurement helpers
// ---------------------------------------------------------------------------
// one launch of the solver's matvec exactly as the Chron-Gear iteration issues it
// which: 1 the FP32 symmetric matvec; 0 the FP64 one the FP64 solver runs (symmetric
// storage when valid); 5 the FP64 one on the full storage; 6 the FP64 symmetric one in the
// other launch order
int launch_cg_matvec(cism_ctx *c, int which) {
    const GridDev &g = c->g;
    const int nb = nblocks_for(g.nn);
    if (which == 1)
        hipLaunchKernelGGL((k_matvec_dots<float, 2, MatSym32>), dim3(xcd_blocks_for(g.nn)), dim3(CISM_BLOCK), 0,
                           c->stream, g,
                           MatSym32{c->A32}, c->av, c->z32, c->Az32, c->r32, c->partials);
    else if (((which == 0 && c->sym64_on) || which == 6) && c->a64s_valid) {
        if ((which == 0) == CISM_SYM64_XCD)
            hipLaunchKernelGGL((k_matvec_dots<double, 2, MatSym64<true>>), dim3(xcd_blocks_for(g.nn)),
                               dim3(CISM_BLOCK), 0, c->stream, g, MatSym64<true>{c->A64s}, c->av, c->z, c->Az,
                               c->r, c->partials);
        else
            hipLaunchKernelGGL((k_matvec_dots<double, 2, MatSym64<false>>), dim3(nb), dim3(CISM_BLOCK), 0,
                               c->stream, g, MatSym64<false>{c->A64s}, c->av, c->z, c->Az, c->r, c->partials);
    } else
        hipLaunchKernelGGL((k_matvec_dots<double, 2, MatFull<double>>), dim3(nb), dim3(CISM_BLOCK), 0,
                           c->stream, g, MatFull<double>{c->A}, c->av, c->z, c->Az, c->r, c->partials);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// one launch of the FP32 fused Chron-Gear update with the column preconditioner, as the
// inner solves issue it (vectors, factors and scalars of the last mixed solve)
int launch_cg_update32(cism_ctx *c) {
    CGVecs<float> v = {c->x32, c->b32, c->r32, c->d32, c->z32, c->q32, c->Az32, c->F32, nullptr};
    return launch_cg_update<float>(c, 2, v, 0, 1);
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
        const double du = x[2 * q] - xr[2 * q], dv = x[2 * q + 1] - xr[2 * q + 1];
        acc[0] += du * du + dv * dv;
        acc[1] += xr[2 * q] * xr[2 * q] + xr[2 * q + 1] * xr[2 * q + 1];
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

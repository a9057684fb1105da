// coarse.hip -- aggregation coarse correction of the mixed-precision inner solves.
//
// The SIA (column tridiagonal) preconditioner resolves each column's vertical coupling
// but not the membrane-stress coupling between columns, whose long-wavelength modes the
// Chron-Gear iteration then has to propagate one neighbour per iteration: its iteration
// count grows with the grid.  The coarse correction adds to z = M_sia^-1 r the
// prolongation of an approximate solution of the Galerkin system on depth-uniform
// aggregates,
//     z = M_sia^-1 r + P V(P^T r),
// P the piecewise-constant prolongation from aggregates of CG_AGG x CG_AGG owned active
// columns (every level of a column, u and v separately), V one symmetric V-cycle over a
// hierarchy of CG_AGG x CG_AGG aggregations of aggregates with l1-Jacobi pre- and
// post-smoothing and the dense inverse of the coarsest level (at most CG_DENSE
// aggregates).  The l1-Jacobi smoother (1 / sum_j |a_ij|) converges for any symmetric
// positive definite level matrix, so V is symmetric positive definite and z remains a
// valid Chron-Gear preconditioner.  Prescribed (Dirichlet) components are left out of P:
// their rows and columns are the identity's, decoupled from the rest, and the SIA
// preconditioner solves them exactly; a coarse correction there would put an error into
// them that their unit diagonal hides from the residual norm.
//
// Level matrices are 9-point stencils of 2 x 2 blocks, A_l[a][d][b]: direction d =
// (di+1) + 3 (dj+1) to the aggregate (ai + di, aj + dj), wrapped on periodic grids, and
// block b = uu, uv, vu, vv.  Level 1 is P^T A32 P, summed in FP64 from the FP32 symmetric
// matrix (each stored coupling and its transpose once); level l+1 sums the blocks of each
// parent's children.  The V-cycle runs on FP32 copies.  The hierarchy is rebuilt once
// per FP32 matrix (the symmetry pass that writes A32 marks it stale).
//
// Periodic wrap: a coupling from aggregate a to aggregate b lies in direction
// cg_wrap(b - a) per axis, the difference taken modulo the aggregate count into
// {-1, 0, 1}; the direction from b to a is its negation, which keeps the stencil
// symmetric (A_l[a][d] = A_l[a+d][-d]^T) also when a grid of one or two aggregates per
// axis makes two directions name the same neighbour.
//
// Only the mixed-precision path uses it (cism_solve_opts.coarse): the FP64 solver keeps
// the reference's preconditioners (glissade_velo_higher_pcg.F90).
#include "internal.h"

#define CG_AGG 4
#define CG_DENSE 32   // aggregates of the coarsest level; its dense inverse is (2 CG_DENSE)^2
static_assert(2 * CG_DENSE <= 64, "k_cg_dense maps one lane column to each of at most 64 unknowns");

__host__ __device__ inline int cg_ceil(int n, int d) { return (n + d - 1) / d; }

__device__ inline int cg_wrap(int d, int n) { return d > 1 ? d - n : (d < -1 ? d + n : d); }

// the aggregate (ai + di, aj + dj) of an ax x ay level, or -1 outside a non-periodic grid
__device__ inline int cg_nb(int ai, int aj, int ax, int ay, int di, int dj, int periodic) {
    int ni = ai + di, nj = aj + dj;
    if (periodic) {
        ni = ni < 0 ? ni + ax : (ni >= ax ? ni - ax : ni);
        nj = nj < 0 ? nj + ay : (nj >= ay ? nj - ay : nj);
    } else if (ni < 0 || ni >= ax || nj < 0 || nj >= ay) {
        return -1;
    }
    return ni + ax * nj;
}

// the level-1 aggregate of every column: owned active columns their own, halo columns of
// a periodic grid their image's, every other column -1
__global__ __launch_bounds__(CISM_BLOCK) void k_cg_cidx(GridDev g, const int32_t *__restrict__ av, int ax,
                                                        int32_t *__restrict__ cidx) {
    const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (long long)g.sx * g.sy) return;
    int i = (int)(col % g.sx), j = (int)(col / g.sx);
    if (!is_owned(g, i, j)) {
        if (!g.periodic) {
            cidx[col] = -1;
            return;
        }
        i = halo_image(i, g.sx, 1);
        j = halo_image(j, g.sy, 1);
        if (i < 0 || j < 0 || !is_owned(g, i, j)) {
            cidx[col] = -1;
            return;
        }
    }
    cidx[col] = av[i + (long long)g.sx * j] ? (i - g.ilo) / CG_AGG + ax * ((j - g.jlo) / CG_AGG) : -1;
}

// Kernels over a parent level run one 16-lane group per parent, lane t on child
// (CG_AGG pi + t % 4, CG_AGG pj + t / 4), and sum over the group with butterfly shuffles
// (a fixed order).  Every lane of a group reaches the shuffles.
static_assert(CG_AGG * CG_AGG == 16, "parent groups are 16 lanes");
__device__ inline float cg_sum16(float x) {
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) x += __shfl_xor(x, off, 16);
    return x;
}
__device__ inline double cg_sum16(double x) {
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) x += __shfl_xor(x, off, 16);
    return x;
}

// Level 1, first pass: the column blocks.  One workgroup per 64 consecutive columns
// stages each stored slot of their nodes through LDS (contiguous loads, 1 KB per wave
// instruction) and lane c sums slot s over the levels of column c into
// F[c][m % 9] (m = CISM_DIAG + s: the column direction (iA+1) + 3 (jA+1) of the
// coupling) or, for the diagonal slot, F[c][9].  The sums run along one column, where the
// entries do not cancel, so FP32 keeps them to its rounding; the cancellation of the
// Galerkin sums happens in the FP64 passes after this one.
#define CG_NZMAX 48
//
// Sf (optional): the same block also forms the SIA preconditioner's factors of its columns
// (k_sia_setup_sym32's operations: the diagonal slot kept from the first round, the slot
// above at round 9), scaled by cs when given, so that the coarse setup replaces that pass.
template <int NZMAX>
__global__ __launch_bounds__(256) void k_cg_colblocks(GridDev g, const float *__restrict__ As,
                                                      const int32_t *__restrict__ av,
                                                      const int32_t *__restrict__ um,
                                                      const int32_t *__restrict__ vm, float *__restrict__ F,
                                                      const float *__restrict__ cs, float *__restrict__ Sf) {
    const int nz = g.nz;
    const long long ncol = (long long)g.sx * g.sy;
    const long long col0 = (long long)blockIdx.x * 64;
    const int ncols = (int)(col0 + 64 <= ncol ? 64 : ncol - col0);
    const long long Q0 = col0 * nz;
    const int nn = ncols * nz;
    __shared__ float4 st[64 * NZMAX];
    __shared__ float4 sd[64 * NZMAX];   // the diagonal slot, for the SIA factors
    const int c = threadIdx.x;
    const bool mine = c < ncols && av[col0 + c];
    float4 acc[10];
#pragma unroll
    for (int d = 0; d < 10; ++d) acc[d] = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int s = 0; s < CISM_SYM_SLOTS; ++s) {
        const int m = CISM_DIAG + s, kA = m / 9 - 1, d = s == 0 ? 9 : m % 9;
        __syncthreads();
        for (int t = threadIdx.x; t < nn; t += blockDim.x) {
            const float4 a = *reinterpret_cast<const float4 *>(As + sidx(0, s, Q0 + t));
            st[t] = a;
            if (s == 0 && Sf) sd[t] = a;
        }
        __syncthreads();
        if (s == 9 && Sf) {
            // SIA factors (k_sia_setup_sym32): lane (comp, column) of the first 128
            const int comp = threadIdx.x >> 6, lc = threadIdx.x & 63;
            const long long colc = col0 + lc;
            if (threadIdx.x < 128 && lc < ncols && av[colc]) {
                auto pick = [&](const float4 &v) { return comp == 0 ? v.x : v.w; };
                auto Fi = [&](int f, int k) { return ((long long)(f * 2 + comp) * nz + k) * ncol + colc; };
                const float inv = cs ? 1.0f / cs[comp * ncol + colc] : 1.0f;
                const float4 *d = sd + lc * nz, *u = st + lc * nz;
                float beta = pick(d[0]);
                Sf[Fi(0, 0)] = 0.0f;
                Sf[Fi(1, 0)] = cs ? beta * inv : beta;
                Sf[Fi(2, 0)] = 0.0f;
                for (int k = 1; k < nz; ++k) {
                    const float gam = pick(u[k - 1]) / beta;
                    const float sb = pick(u[k - 1]);
                    beta = pick(d[k]) - sb * gam;
                    Sf[Fi(0, k)] = cs ? sb * inv : sb;
                    Sf[Fi(1, k)] = cs ? beta * inv : beta;
                    Sf[Fi(2, k)] = gam;
                }
            }
        }
        if (mine) {
            for (int k = 0; k + kA < nz; ++k) {
                float4 B = st[c * nz + k];
                if (s == 0 && um) {   // a prescribed component's unit diagonal (its couplings are zero)
                    const long long q = Q0 + c * nz + k;
                    if (um[q] == 1) B.x = 0.0f;
                    if (vm[q] == 1) B.w = 0.0f;
                }
                acc[d].x += B.x;
                acc[d].y += B.y;
                acc[d].z += B.z;
                acc[d].w += B.w;
            }
        }
    }
    if (mine) {
        float4 *out = reinterpret_cast<float4 *>(F) + (col0 + c) * 10;
#pragma unroll
        for (int d = 0; d < 10; ++d) out[d] = acc[d];
    }
}

// Level 1, second pass: one 16-lane group per aggregate, a lane per owned active column of
// it, sums the column blocks by the direction of the neighbour column's aggregate: S[a][d]
// for d = 0..8, the diagonal blocks in S[a][9] (FP64, a fixed shuffle order).
__global__ __launch_bounds__(CISM_BLOCK) void k_cg_gal1(GridDev g, const float *__restrict__ F,
                                                        const int32_t *__restrict__ av,
                                                        const int32_t *__restrict__ cidx, int ax, int ay,
                                                        double *__restrict__ S) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long a = t >> 4;
    const int sub = (int)(t & 15);
    const bool oka = a < (long long)ax * ay;
    const int ai = oka ? (int)(a % ax) : 0, aj = oka ? (int)(a / ax) : 0;
    const int i = g.ilo + CG_AGG * ai + (sub & 3), j = g.jlo + CG_AGG * aj + (sub >> 2);
    double acc[10][4];
#pragma unroll
    for (int d = 0; d < 10; ++d)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[d][b] = 0.0;
    if (oka && i <= g.ihi && j <= g.jhi && av[i + (long long)g.sx * j]) {
        const long long col = i + (long long)g.sx * j;
        const float4 *Fc = reinterpret_cast<const float4 *>(F) + col * 10;
#pragma unroll
        for (int dc = 0; dc < 9; ++dc) {
            const int b = cidx[(i + dc % 3 - 1) + (long long)g.sx * (j + dc / 3 - 1)];
            if (b < 0) continue;
            const int dir = (cg_wrap(b % ax - ai, ax) + 1) + 3 * (cg_wrap(b / ax - aj, ay) + 1);
            const float4 B = Fc[dc];
#pragma unroll
            for (int d = 0; d < 9; ++d)
                if (d == dir) {
                    acc[d][0] += B.x;
                    acc[d][1] += B.y;
                    acc[d][2] += B.z;
                    acc[d][3] += B.w;
                }
        }
        const float4 D = Fc[9];
        acc[9][0] = D.x;
        acc[9][1] = D.y;
        acc[9][2] = D.z;
        acc[9][3] = D.w;
    }
#pragma unroll
    for (int d = 0; d < 10; ++d)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[d][b] = cg_sum16(acc[d][b]);
    if (oka && sub == 0) {
#pragma unroll
        for (int d = 0; d < 10; ++d)
#pragma unroll
            for (int b = 0; b < 4; ++b) S[a * 40 + 4 * d + b] = acc[d][b];
    }
}

// a level row: FP64 and FP32 blocks and the l1 diagonal (1 / l1 norm of the u and v rows)
__device__ inline void cg_store_row(long long a, const double (&o)[9][4], double *__restrict__ A64,
                                    float *__restrict__ A, float *__restrict__ dl1) {
    double lu = 0.0, lv = 0.0;
#pragma unroll
    for (int d = 0; d < 9; ++d) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            A64[a * 36 + 4 * d + b] = o[d][b];
            A[a * 36 + 4 * d + b] = (float)o[d][b];
        }
        lu += fabs(o[d][0]) + fabs(o[d][1]);
        lv += fabs(o[d][2]) + fabs(o[d][3]);
    }
    dl1[2 * a] = lu > 0.0 ? (float)(1.0 / lu) : 0.0f;
    dl1[2 * a + 1] = lv > 0.0 ? (float)(1.0 / lv) : 0.0f;
}

// Level 1, second pass: A[a][d] = S[a][d] + S[a+d][-d]^T, the centre S[a][9] + S[a][4] +
// S[a][4]^T (a coupling inside the aggregate is stored once, at its first node)
__global__ __launch_bounds__(CISM_BLOCK) void k_cg_combine1(int ax, int ay, int periodic,
                                                            const double *__restrict__ S,
                                                            double *__restrict__ A64, float *__restrict__ A,
                                                            float *__restrict__ dl1) {
    const long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= (long long)ax * ay) return;
    const int ai = (int)(a % ax), aj = (int)(a / ax);
    double o[9][4];
#pragma unroll
    for (int d = 0; d < 9; ++d) {
        const double *s = S + a * 40 + 4 * d;
        if (d == 4) {
            const double *sd = S + a * 40 + 36;
            o[d][0] = sd[0] + 2.0 * s[0];
            o[d][1] = sd[1] + (s[1] + s[2]);
            o[d][2] = sd[2] + (s[2] + s[1]);
            o[d][3] = sd[3] + 2.0 * s[3];
        } else {
            o[d][0] = s[0];
            o[d][1] = s[1];
            o[d][2] = s[2];
            o[d][3] = s[3];
            const int nb = cg_nb(ai, aj, ax, ay, d % 3 - 1, d / 3 - 1, periodic);
            if (nb >= 0) {
                const double *t = S + (long long)nb * 40 + 4 * (8 - d);
                o[d][0] += t[0];
                o[d][1] += t[2];
                o[d][2] += t[1];
                o[d][3] += t[3];
            }
        }
    }
    cg_store_row(a, o, A64, A, dl1);
}

// level l+1 from level l: each parent sums its children's blocks by the direction of the
// neighbour's parent
__global__ __launch_bounds__(CISM_BLOCK) void k_cg_galup(int axc, int ayc, int axp, int ayp, int periodic,
                                                         const double *__restrict__ Ac,
                                                         double *__restrict__ A64, float *__restrict__ A,
                                                         float *__restrict__ dl1) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long P = t >> 4;
    const int sub = (int)(t & 15);
    const bool okP = P < (long long)axp * ayp;
    const int pi = okP ? (int)(P % axp) : 0, pj = okP ? (int)(P / axp) : 0;
    const int ci = CG_AGG * pi + (sub & 3), cj = CG_AGG * pj + (sub >> 2);
    double o[9][4];
#pragma unroll
    for (int d = 0; d < 9; ++d)
#pragma unroll
        for (int b = 0; b < 4; ++b) o[d][b] = 0.0;
    if (okP && ci < axc && cj < ayc) {
        const double *src = Ac + (long long)(ci + axc * cj) * 36;
#pragma unroll
        for (int d = 0; d < 9; ++d) {
            const int nb = cg_nb(ci, cj, axc, ayc, d % 3 - 1, d / 3 - 1, periodic);
            if (nb < 0) continue;
            const int Di = cg_wrap((nb % axc) / CG_AGG - pi, axp), Dj = cg_wrap((nb / axc) / CG_AGG - pj, ayp);
            const int D = (Di + 1) + 3 * (Dj + 1);
#pragma unroll
            for (int e = 0; e < 9; ++e)
                if (e == D) {
                    o[e][0] += src[4 * d];
                    o[e][1] += src[4 * d + 1];
                    o[e][2] += src[4 * d + 2];
                    o[e][3] += src[4 * d + 3];
                }
        }
    }
#pragma unroll
    for (int d = 0; d < 9; ++d)
#pragma unroll
        for (int b = 0; b < 4; ++b) o[d][b] = cg_sum16(o[d][b]);
    if (okP && sub == 0) cg_store_row(P, o, A64, A, dl1);
}

// The coarsest level's dense matrix (2 nL unknowns, u and v of each aggregate) inverted
// by Gauss-Jordan elimination without pivoting (symmetric positive definite) in LDS.  An
// aggregate with no unknowns (zero rows) gets a unit diagonal.  A pivot that is not
// positive (a matrix made indefinite by rounding) switches the coarsest correction off:
// the inverse is written as zero.
__global__ __launch_bounds__(256) void k_cg_dense(int ax, int ay, int periodic, const double *__restrict__ A64,
                                                  float *__restrict__ Ainv) {
    __shared__ double M[2 * CG_DENSE][2 * CG_DENSE + 1];
    __shared__ double colk[2 * CG_DENSE];
    const int nL = ax * ay, n = 2 * nL, t = threadIdx.x;
    for (int w = t; w < 2 * CG_DENSE * (2 * CG_DENSE + 1); w += blockDim.x) (&M[0][0])[w] = 0.0;
    __syncthreads();
    if (t < n) {   // row t: aggregate t / 2, component t % 2 (one writer per row)
        const int a = t >> 1, cu = t & 1, ai = a % ax, aj = a / ax;
        for (int d = 0; d < 9; ++d) {
            const int nb = cg_nb(ai, aj, ax, ay, d % 3 - 1, d / 3 - 1, periodic);
            if (nb < 0) continue;
            M[t][2 * nb] += A64[(long long)a * 36 + 4 * d + 2 * cu];
            M[t][2 * nb + 1] += A64[(long long)a * 36 + 4 * d + 2 * cu + 1];
        }
        if (M[t][t] == 0.0) M[t][t] = 1.0;
    }
    __syncthreads();
    bool bad = false;
    for (int k = 0; k < n; ++k) {
        const double piv = M[k][k];
        if (!(piv > 0.0) || isinf(piv)) {   // uniform: every lane reads the same pivot
            bad = true;
            break;
        }
        if (t < n) colk[t] = M[t][k];
        __syncthreads();
        if (t < n) M[k][t] = (t == k ? 1.0 : M[k][t]) / piv;
        __syncthreads();
        // lane t: column t & 63, rows t >> 6, (t >> 6) + 4, ... (n <= 64, 256 lanes)
        {
            const int j = t & 63;
            if (j < n) {
                const double mkj = M[k][j];
                for (int i = t >> 6; i < n; i += 4)
                    if (i != k) M[i][j] = (j == k ? 0.0 : M[i][j]) - colk[i] * mkj;
            }
        }
        __syncthreads();
    }
    for (int w = t; w < n * n; w += blockDim.x) {
        const int i = w / n, j = w - i * n;
        Ainv[w] = bad ? 0.0f : (float)M[i][j];
    }
}

// ---------------------------------------------------------------------------
// the V-cycle
// ---------------------------------------------------------------------------
// column sums of r over the levels of owned active columns (the restriction's first
// stage when z comes from the stand-alone SIA apply; the fused update writes them itself)
__global__ __launch_bounds__(CISM_BLOCK) void k_cg_colsum(GridDev g, const int32_t *__restrict__ av,
                                                          const int32_t *__restrict__ um,
                                                          const int32_t *__restrict__ vm,
                                                          const float *__restrict__ r, float *__restrict__ colsum) {
    const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (long long)g.sx * g.sy) return;
    const int i = (int)(col % g.sx), j = (int)(col / g.sx);
    if (!is_owned(g, i, j) || !av[col]) return;
    const float2 *r2 = reinterpret_cast<const float2 *>(r) + col * g.nz;
    float su = 0.0f, sv = 0.0f;
    for (int k = 0; k < g.nz; ++k) {
        const float2 v = r2[k];
        const long long q = k + (long long)g.nz * col;
        if (!um || um[q] != 1) su += v.x;
        if (!vm || vm[q] != 1) sv += v.y;
    }
    reinterpret_cast<float2 *>(colsum)[col] = make_float2(su, sv);
}

// level 1: r1 = P^T r from the column sums (one 16-lane group per aggregate, a lane per
// column); with smooth, e1 = D_l1^-1 r1
__global__ __launch_bounds__(CISM_BLOCK) void k_cg_restrict1(GridDev g, const int32_t *__restrict__ av,
                                                             const float *__restrict__ colsum, int ax, int ay,
                                                             int smooth, const float *__restrict__ dl1,
                                                             float *__restrict__ r, float *__restrict__ e) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long a = t >> 4;
    const int sub = (int)(t & 15);
    const bool oka = a < (long long)ax * ay;
    float su = 0.0f, sv = 0.0f;
    if (oka) {
        const int ai = (int)(a % ax), aj = (int)(a / ax);
        const int i = g.ilo + CG_AGG * ai + (sub & 3), j = g.jlo + CG_AGG * aj + (sub >> 2);
        if (i <= g.ihi && j <= g.jhi) {
            const long long col = i + (long long)g.sx * j;
            if (av[col]) {
                const float2 v = reinterpret_cast<const float2 *>(colsum)[col];
                su = v.x;
                sv = v.y;
            }
        }
    }
    su = cg_sum16(su);
    sv = cg_sum16(sv);
    if (oka && sub == 0) {
        reinterpret_cast<float2 *>(r)[a] = make_float2(su, sv);
        if (smooth) reinterpret_cast<float2 *>(e)[a] = make_float2(dl1[2 * a] * su, dl1[2 * a + 1] * sv);
    }
}

// restriction of the pre-smoothed residual: r_{l+1}[P] = sum over children c of
// (r_l - A_l e_l)[c]; with smooth, e_{l+1} = D_l1^-1 r_{l+1}
__global__ __launch_bounds__(CISM_BLOCK) void k_cg_down(int axc, int ayc, int axp, int ayp, int periodic,
                                                        const float *__restrict__ A, const float *__restrict__ rc,
                                                        const float *__restrict__ ec, int smooth,
                                                        const float *__restrict__ dl1p, float *__restrict__ rp,
                                                        float *__restrict__ ep) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long P = t >> 4;
    const int sub = (int)(t & 15);
    const bool okP = P < (long long)axp * ayp;
    float ru = 0.0f, rv = 0.0f;
    if (okP) {
        const int ci = CG_AGG * (int)(P % axp) + (sub & 3), cj = CG_AGG * (int)(P / axp) + (sub >> 2);
        if (ci < axc && cj < ayc) {
            const long long c = ci + (long long)axc * cj;
            const float2 *e2 = reinterpret_cast<const float2 *>(ec);
            const float2 rr = reinterpret_cast<const float2 *>(rc)[c];
            ru = rr.x;
            rv = rr.y;
            const float4 *Ab = reinterpret_cast<const float4 *>(A + c * 36);
#pragma unroll
            for (int d = 0; d < 9; ++d) {
                const int nb = cg_nb(ci, cj, axc, ayc, d % 3 - 1, d / 3 - 1, periodic);
                if (nb < 0) continue;
                const float4 B = Ab[d];
                const float2 en = e2[nb];
                ru -= B.x * en.x + B.y * en.y;
                rv -= B.z * en.x + B.w * en.y;
            }
        }
    }
    ru = cg_sum16(ru);
    rv = cg_sum16(rv);
    if (okP && sub == 0) {
        reinterpret_cast<float2 *>(rp)[P] = make_float2(ru, rv);
        if (smooth) reinterpret_cast<float2 *>(ep)[P] = make_float2(dl1p[2 * P] * ru, dl1p[2 * P + 1] * rv);
    }
}

// the coarsest level: e = Ainv r (one workgroup, n <= 2 CG_DENSE)
__global__ __launch_bounds__(64) void k_cg_dense_apply(int n, const float *__restrict__ Ainv,
                                                       const float *__restrict__ r, float *__restrict__ e) {
    const int i = threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int j = 0; j < n; ++j) s += Ainv[i * n + j] * r[j];
    e[i] = s;
}

// prolongation and post-smoothing on level l: with e' = e_l + Q e_{l+1} (Q the parent
// injection), eo = e' + D_l1^-1 (r_l - A_l e')
__global__ __launch_bounds__(CISM_BLOCK) void k_cg_up(int ax, int ay, int axp, int periodic,
                                                      const float *__restrict__ A, const float *__restrict__ r,
                                                      const float *__restrict__ e, const float *__restrict__ ep,
                                                      const float *__restrict__ dl1, float *__restrict__ eo) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long long)ax * ay) return;
    const int ci = (int)(c % ax), cj = (int)(c / ax);
    const float2 *e2 = reinterpret_cast<const float2 *>(e), *p2 = reinterpret_cast<const float2 *>(ep);
    auto corr = [&](int a) {
        const float2 x = e2[a], y = p2[(a % ax) / CG_AGG + axp * ((a / ax) / CG_AGG)];
        return make_float2(x.x + y.x, x.y + y.y);
    };
    const float2 ec = corr((int)c);
    const float2 rr = reinterpret_cast<const float2 *>(r)[c];
    float ru = rr.x, rv = rr.y;
    const float4 *Ab = reinterpret_cast<const float4 *>(A + c * 36);
#pragma unroll
    for (int d = 0; d < 9; ++d) {
        const int nb = cg_nb(ci, cj, ax, ay, d % 3 - 1, d / 3 - 1, periodic);
        if (nb < 0) continue;
        const float4 B = Ab[d];
        const float2 en = corr(nb);
        ru -= B.x * en.x + B.y * en.y;
        rv -= B.z * en.x + B.w * en.y;
    }
    reinterpret_cast<float2 *>(eo)[c] = make_float2(ec.x + dl1[2 * c] * ru, ec.y + dl1[2 * c + 1] * rv);
}

// z += P e1 on every node of an owned active column and of its periodic halo images
__global__ __launch_bounds__(CISM_BLOCK) void k_cg_prolong(GridDev g, const int32_t *__restrict__ cidx,
                                                           const int32_t *__restrict__ um,
                                                           const int32_t *__restrict__ vm,
                                                           const float *__restrict__ e1, float *__restrict__ z) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nn) return;
    const int a = cidx[q / g.nz];
    if (a < 0) return;
    float2 ea = reinterpret_cast<const float2 *>(e1)[a];
    if (um) {   // halo nodes carry their image's masks
        if (um[q] == 1) ea.x = 0.0f;
        if (vm[q] == 1) ea.y = 0.0f;
    }
    float2 *z2 = reinterpret_cast<float2 *>(z) + q;
    const float2 v = *z2;
    *z2 = make_float2(v.x + ea.x, v.y + ea.y);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int coarse_alloc(cism_ctx *c) {
    if (c->coarse) return CISM_OK;
    const GridDev &g = c->g;
    CoarseHier *h = (CoarseHier *)calloc(1, sizeof(CoarseHier));
    if (!h) {
        cism_set_error("coarse hierarchy: host allocation failed");
        return CISM_ERR_NOMEM;
    }
    c->coarse = h;
    int ax = cg_ceil(g.ihi - g.ilo + 1, CG_AGG), ay = cg_ceil(g.jhi - g.jlo + 1, CG_AGG);
    h->nlev = 1;
    h->lev[0].ax = ax;
    h->lev[0].ay = ay;
    while ((long long)ax * ay > CG_DENSE) {
        if (h->nlev == CG_MAXLEV) {
            cism_set_error("coarse hierarchy: more than %d levels", CG_MAXLEV);
            return CISM_ERR_UNSUPPORTED;
        }
        ax = cg_ceil(ax, CG_AGG);
        ay = cg_ceil(ay, CG_AGG);
        h->lev[h->nlev].ax = ax;
        h->lev[h->nlev].ay = ay;
        h->nlev++;
    }
    const size_t ncol = (size_t)g.sx * g.sy;
    hipError_t e = hipMalloc(&h->cidx, sizeof(int32_t) * ncol);
    if (e == hipSuccess) e = hipMalloc(&h->colsum, sizeof(float) * 2 * ncol);
    if (e == hipSuccess) e = hipMalloc(&h->colF, sizeof(float) * 40 * ncol);
    const size_t n1 = (size_t)h->lev[0].ax * h->lev[0].ay;
    if (e == hipSuccess) e = hipMalloc(&h->S, sizeof(double) * 40 * n1);
    if (e == hipSuccess) e = hipMalloc(&h->Ainv, sizeof(float) * 4 * CG_DENSE * CG_DENSE);
    for (int l = 0; l < h->nlev && e == hipSuccess; ++l) {
        CoarseLevel &L = h->lev[l];
        const size_t n = (size_t)L.ax * L.ay;
        e = hipMalloc(&L.A64, sizeof(double) * 36 * n);
        if (e == hipSuccess) e = hipMalloc(&L.A, sizeof(float) * 36 * n);
        float **vs[] = {&L.dl1, &L.r, &L.e, &L.eo};
        for (float **p : vs)
            if (e == hipSuccess) e = hipMalloc(p, sizeof(float) * 2 * n);
    }
    if (e != hipSuccess) {
        cism_set_error("coarse hierarchy allocation failed: %s", hipGetErrorString(e));
        return CISM_ERR_NOMEM;
    }
    return CISM_OK;
}

void coarse_free(cism_ctx *c) {
    CoarseHier *h = c->coarse;
    if (!h) return;
    void *ptrs[] = {h->cidx, h->colsum, h->colF, h->S, h->Ainv};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (int l = 0; l < CG_MAXLEV; ++l) {
        CoarseLevel &L = h->lev[l];
        void *lp[] = {L.A64, L.A, L.dl1, L.r, L.e, L.eo};
        for (void *p : lp)
            if (p) (void)hipFree(p);
    }
    free(h);
    c->coarse = nullptr;
}

void coarse_invalidate(cism_ctx *c) {
    if (c->coarse) c->coarse->ok = 0;
}

int coarse_setup(cism_ctx *c, const float *As, float *Sf, const float *cs) {
    if (c->g.nz > CG_NZMAX) {
        cism_set_error("coarse correction supports nz <= %d (nz = %d)", CG_NZMAX, c->g.nz);
        return CISM_ERR_UNSUPPORTED;
    }
    CISM_TRY(coarse_alloc(c));
    CoarseHier *h = c->coarse;
    if (h->ok) return CISM_OK;
    const GridDev &g = c->g;
    const long long ncol = (long long)g.sx * g.sy;
    CoarseLevel *lv = h->lev;
    hipLaunchKernelGGL(k_cg_cidx, dim3(nblocks_for(ncol)), dim3(CISM_BLOCK), 0, c->stream, g, c->av, lv[0].ax,
                       h->cidx);
    const long long n1 = (long long)lv[0].ax * lv[0].ay;
    const int32_t *um = c->dirichlet_any ? c->umask : nullptr, *vm = c->dirichlet_any ? c->vmask : nullptr;
    const dim3 gcol((unsigned)((ncol + 63) / 64));
    if (g.nz <= 12)
        hipLaunchKernelGGL(k_cg_colblocks<12>, gcol, dim3(256), 0, c->stream, g, As, c->av, um, vm, h->colF, cs, Sf);
    else if (g.nz <= 24)
        hipLaunchKernelGGL(k_cg_colblocks<24>, gcol, dim3(256), 0, c->stream, g, As, c->av, um, vm, h->colF, cs, Sf);
    else
        hipLaunchKernelGGL(k_cg_colblocks<CG_NZMAX>, gcol, dim3(256), 0, c->stream, g, As, c->av, um, vm, h->colF, cs,
                           Sf);
    hipLaunchKernelGGL(k_cg_gal1, dim3(nblocks_for(16 * n1)), dim3(CISM_BLOCK), 0, c->stream, g, h->colF, c->av,
                       h->cidx, lv[0].ax, lv[0].ay, h->S);
    hipLaunchKernelGGL(k_cg_combine1, dim3(nblocks_for(n1)), dim3(CISM_BLOCK), 0, c->stream, lv[0].ax, lv[0].ay,
                       g.periodic, h->S, lv[0].A64, lv[0].A, lv[0].dl1);
    for (int l = 0; l + 1 < h->nlev; ++l) {
        const long long np = (long long)lv[l + 1].ax * lv[l + 1].ay;
        hipLaunchKernelGGL(k_cg_galup, dim3(nblocks_for(16 * np)), dim3(CISM_BLOCK), 0, c->stream, lv[l].ax, lv[l].ay,
                           lv[l + 1].ax, lv[l + 1].ay, g.periodic, lv[l].A64, lv[l + 1].A64, lv[l + 1].A,
                           lv[l + 1].dl1);
    }
    const CoarseLevel &C = lv[h->nlev - 1];
    hipLaunchKernelGGL(k_cg_dense, dim3(1), dim3(256), 0, c->stream, C.ax, C.ay, g.periodic, C.A64, h->Ainv);
    HIP_TRY(hipGetLastError());
    h->ok = 1;
    return CISM_OK;
}

int coarse_colsum(cism_ctx *c, const float *r) {
    const GridDev &g = c->g;
    const int32_t *um = c->dirichlet_any ? c->umask : nullptr, *vm = c->dirichlet_any ? c->vmask : nullptr;
    hipLaunchKernelGGL(k_cg_colsum, dim3(nblocks_for((long long)g.sx * g.sy)), dim3(CISM_BLOCK), 0, c->stream, g,
                       c->av, um, vm, r, c->coarse->colsum);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

int coarse_apply(cism_ctx *c, float *z) {
    CoarseHier *h = c->coarse;
    if (!h || !h->ok) {
        cism_set_error("coarse correction applied before its setup");
        return CISM_ERR_ARG;
    }
    const GridDev &g = c->g;
    CoarseLevel *lv = h->lev;
    const int L = h->nlev;
    const long long n1 = (long long)lv[0].ax * lv[0].ay;
    hipLaunchKernelGGL(k_cg_restrict1, dim3(nblocks_for(16 * n1)), dim3(CISM_BLOCK), 0, c->stream, g, c->av, h->colsum,
                       lv[0].ax, lv[0].ay, L > 1, lv[0].dl1, lv[0].r, lv[0].e);
    for (int l = 0; l + 1 < L; ++l) {
        const long long np = (long long)lv[l + 1].ax * lv[l + 1].ay;
        hipLaunchKernelGGL(k_cg_down, dim3(nblocks_for(16 * np)), dim3(CISM_BLOCK), 0, c->stream, lv[l].ax, lv[l].ay,
                           lv[l + 1].ax, lv[l + 1].ay, g.periodic, lv[l].A, lv[l].r, lv[l].e, l + 2 < L,
                           lv[l + 1].dl1, lv[l + 1].r, lv[l + 1].e);
    }
    const CoarseLevel &C = lv[L - 1];
    hipLaunchKernelGGL(k_cg_dense_apply, dim3(1), dim3(64), 0, c->stream, 2 * C.ax * C.ay, h->Ainv, C.r, C.eo);
    for (int l = L - 2; l >= 0; --l) {
        const long long n = (long long)lv[l].ax * lv[l].ay;
        hipLaunchKernelGGL(k_cg_up, dim3(nblocks_for(n)), dim3(CISM_BLOCK), 0, c->stream, lv[l].ax, lv[l].ay,
                           lv[l + 1].ax, g.periodic, lv[l].A, lv[l].r, lv[l].e, lv[l + 1].eo, lv[l].dl1, lv[l].eo);
    }
    const int32_t *um = c->dirichlet_any ? c->umask : nullptr, *vm = c->dirichlet_any ? c->vmask : nullptr;
    hipLaunchKernelGGL(k_cg_prolong, dim3(nblocks_for(g.nn)), dim3(CISM_BLOCK), 0, c->stream, g, h->cidx, um, vm,
                       lv[0].eo, z);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

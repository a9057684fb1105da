// ssa.hip -- the 2D (SSA) velocity solve (glissade_velo_higher.F90, solve_2d branch):
// assembly of the depth-integrated 9-point blocks on a one-level context (nz = 1).
//
// The 2D system is stored in the 3D layout of a one-level grid: the nine couplings
// (iA, jA) of a vertex occupy the kA = 0 slots slot_of(iA, jA, 0) of the 27-slot
// tiles, so the Dirichlet, halo, symmetry, nnz, residual, matvec and PCG code of the
// 3D path runs the 2D system unchanged (with nz = 1 every kA = +-1 slot is out of
// range and skipped), exactly as pcg_solver_*_2d mirror the 3D solvers.
//
// Assembly in two atomic-free kernels, as in 3D:
//  1. k_element_2d: one lane per cell evaluates the four quadrature points
//     (get_basis_function_derivatives_2d, velo_higher:6583-6751), the SSA viscosity
//     from the column-mean flow factor (compute_effective_viscosity, :7199-7378),
//     the thickness at the point, and the element blocks (compute_element_matrix,
//     SSA branch, :7868-7986): 10 + 10 entries of the symmetric Kuu, Kvv and the 16
//     of Kuv (Kvu = Kuv^T exactly).
//  2. k_gather_2d: one lane per vertex sums its 9 x 4 coefficients over the (up to
//     four) cells that share it, in the reference's element order (j, i ascending;
//     element_to_global_matrix_2d, :8058-8121), then adds the local basal term.
#include "internal.h"

#include <math.h>

#define KE2_ENTRIES 36   // Kuu upper 10, Kvv upper 10, Kuv 16

// packed upper-triangle index of (a, b) in a 4x4 matrix
__host__ __device__ constexpr int tri4(int a, int b) {
    return a <= b ? (a * (7 - a)) / 2 + b : (b * (7 - b)) / 2 + a;
}

// tiles of 64 cells, entry-major inside a tile
__host__ __device__ inline long long ke2idx(int ent, long long cell) {
    return (((cell >> 6) * KE2_ENTRIES + ent) << 6) + (cell & 63);
}

// node n of a cell: vertex offsets from the cell's NE vertex (ishift(3,n), jshift(3,n)):
// SW, SE, NE, NW
__constant__ int c_di2[4] = {-1, 0, 0, -1};
__constant__ int c_dj2[4] = {-1, -1, 0, 0};

// glissade_vertical_average (glissade_grid_operators.F90:1172-1236), no mask
__global__ __launch_bounds__(256) void k_vertical_average(int nx, int ny, int nz3,
                                                          const double *__restrict__ sigma,
                                                          const double *__restrict__ var,
                                                          double *__restrict__ var2) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nx * ny) return;
    double s = 0.0;
    for (int k = 0; k < nz3 - 1; ++k) s = s + var[k + (long long)(nz3 - 1) * t] * (sigma[k + 1] - sigma[k]);
    var2[t] = s;
}

__global__ __launch_bounds__(256) void k_element_2d(GridDev g, const int32_t *__restrict__ ac,
                                                    const double *__restrict__ xV,
                                                    const double *__restrict__ yV,
                                                    const double *__restrict__ vel,
                                                    const double *__restrict__ sthck,
                                                    const double *__restrict__ ff2, int whichefvs,
                                                    double efvs_constant, double *__restrict__ Ke,
                                                    double *__restrict__ efvs2, DevScal *__restrict__ sc) {
    const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= (long long)g.nx * g.ny) return;
    const int i = (int)(cell % g.nx) + 1, j = (int)(cell / g.nx) + 1;   // 1-based
    const bool ok = i >= CISM_NHALO + 1 && i <= g.nx - CISM_NHALO + 1 && j >= CISM_NHALO + 1 &&
                    j <= g.ny - CISM_NHALO + 1 && ac[cell] != 0;
    if (!ok) {
        efvs2[cell] = 0.0;
        return;
    }
    double x[4], y[4], u[4], v[4], hn[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        const long long vtx = (i - 1 + c_di2[n]) + (long long)g.sx * (j - 1 + c_dj2[n]);
        x[n] = xV[vtx];
        y[n] = yV[vtx];
        u[n] = vel[2 * vtx];
        v[n] = vel[2 * vtx + 1];
        hn[n] = sthck[vtx];
    }
    double Kuu[10], Kvv[10], Kuv[16], evsum = 0.0;
#pragma unroll
    for (int e = 0; e < 10; ++e) Kuu[e] = Kvv[e] = 0.0;
#pragma unroll
    for (int e = 0; e < 16; ++e) Kuv[e] = 0.0;
    const double r = 1.0 / sqrt(3.0);
    const double ff = ff2[cell];
    double evq[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const double xq = (p == 0 || p == 3) ? -r : r, yq = p < 2 ? -r : r;
        double dxr[4], dyr[4], phi[4];
        dxr[0] = -(1.0 - yq) / 4.0; dxr[1] = (1.0 - yq) / 4.0;
        dxr[2] = (1.0 + yq) / 4.0;  dxr[3] = -(1.0 + yq) / 4.0;
        dyr[0] = -(1.0 - xq) / 4.0; dyr[1] = -(1.0 + xq) / 4.0;
        dyr[2] = (1.0 + xq) / 4.0;  dyr[3] = (1.0 - xq) / 4.0;
        phi[0] = (1.0 - xq) * (1.0 - yq) / 4.0;
        phi[1] = (1.0 + xq) * (1.0 - yq) / 4.0;
        phi[2] = (1.0 + xq) * (1.0 + yq) / 4.0;
        phi[3] = (1.0 - xq) * (1.0 + yq) / 4.0;
        double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            J11 = J11 + dxr[n] * x[n];
            J12 = J12 + dxr[n] * y[n];
            J21 = J21 + dyr[n] * x[n];
            J22 = J22 + dyr[n] * y[n];
        }
        const double det = J11 * J22 - J12 * J21;
        double i11 = 0, i12 = 0, i21 = 0, i22 = 0;
        if (fabs(det) > 0.0) {
            i11 = J22 / det; i12 = -J12 / det; i21 = -J21 / det; i22 = J11 / det;
        } else {
            atomicOr(&sc->singular_flag, 1);
        }
        double dx[4], dy[4];
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            dx[n] = i11 * dxr[n] + i12 * dyr[n];
            dy[n] = i21 * dxr[n] + i22 * dyr[n];
        }
        double ev;
        if (whichefvs == 0) {
            ev = efvs_constant;
        } else if (whichefvs == 1) {
            const double effstrain = (500.0 / CISM_SCYR) / 200.0e3 * CISM_SCYR;
            ev = ff * pow(effstrain, (1.0 - 3.0) / 3.0);
        } else {
            double ux = 0, uy = 0, vx = 0, vy = 0;
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                ux = ux + dx[n] * u[n]; uy = uy + dy[n] * u[n];
                vx = vx + dx[n] * v[n]; vy = vy + dy[n] * v[n];
            }
            const double emin = 1.0e-8;
            const double sq = emin * emin + (ux * ux + vy * vy + ux * vy + 0.25 * ((vx + uy) * (vx + uy)));
            ev = ff * pow(sq, ((1.0 - 3.0) / 3.0) / 2);
        }
        evq[p] = ev;
        double hq = 0.0;
#pragma unroll
        for (int n = 0; n < 4; ++n) hq = hq + phi[n] * hn[n];
        const double f = hq * ev * 1.0 * det / CISM_VOL0;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = a; b < 4; ++b) {
                Kuu[tri4(a, b)] = Kuu[tri4(a, b)] + f * (4.0 * dx[a] * dx[b] + dy[a] * dy[b]);
                Kvv[tri4(a, b)] = Kvv[tri4(a, b)] + f * (4.0 * dy[a] * dy[b] + dx[a] * dx[b]);
            }
#pragma unroll
        for (int nc = 0; nc < 4; ++nc)
#pragma unroll
            for (int nr = 0; nr < 4; ++nr)
                Kuv[nr + 4 * nc] = Kuv[nr + 4 * nc] + f * (2.0 * dx[nr] * dy[nc] + dy[nr] * dx[nc]);
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) evsum = evsum + evq[p];
    efvs2[cell] = evsum / 4;
#pragma unroll
    for (int e = 0; e < 10; ++e) {
        Ke[ke2idx(e, cell)] = Kuu[e];
        Ke[ke2idx(10 + e, cell)] = Kvv[e];
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) Ke[ke2idx(20 + e, cell)] = Kuv[e];
}

__device__ inline double ke2_val(const double *__restrict__ Ke, int blk, int nr, int nc, long long cell) {
    int ent;
    if (blk == 0) ent = tri4(nr, nc);
    else if (blk == 3) ent = 10 + tri4(nr, nc);
    else if (blk == 1) ent = 20 + nr + 4 * nc;   // Kuv(nr, nc)
    else ent = 20 + nc + 4 * nr;                 // Kvu(nr, nc) = Kuv(nc, nr)
    return Ke[ke2idx(ent, cell)];
}

// node index of the vertex at offset (ox, oy) from the cell's NE vertex, or -1
__device__ inline int node2(int ox, int oy) {
    if (oy == -1) return ox == -1 ? 0 : (ox == 0 ? 1 : -1);
    if (oy == 0) return ox == 0 ? 2 : (ox == -1 ? 3 : -1);
    return -1;
}

// One lane per vertex (i, j): cells (i, j), (i+1, j), (i, j+1), (i+1, j+1) (1-based
// cell = NE vertex), the reference's element order
__global__ __launch_bounds__(256) void k_gather_2d(GridDev g, const int32_t *__restrict__ ac,
                                                   const int32_t *__restrict__ av,
                                                   const double *__restrict__ Ke,
                                                   const double *__restrict__ beta, int add_basal,
                                                   double basal_scale, double *__restrict__ A) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;   // vertex = node
    if (q >= (long long)g.sx * g.sy) return;
    const int iv = (int)(q % g.sx) + 1, jv = (int)(q / g.sx) + 1;
    long long cid[4];
    bool cv[4];
    int nr[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int dI = t & 1, dJ = t >> 1;
        const int ic = iv + dI, jc = jv + dJ;
        const bool okc = ic >= CISM_NHALO + 1 && ic <= g.nx - CISM_NHALO + 1 && jc >= CISM_NHALO + 1 &&
                         jc <= g.ny - CISM_NHALO + 1;
        cid[t] = (ic - 1) + (long long)g.nx * (jc - 1);
        cv[t] = okc && ac[cid[t]] != 0;
        nr[t] = node2(-dI, -dJ);
    }
#pragma unroll
    for (int jA = -1; jA <= 1; ++jA)
#pragma unroll
        for (int iA = -1; iA <= 1; ++iA) {
            const int m = slot_of(iA, jA, 0);
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int blk = 0; blk < 4; ++blk) {
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int dI = t & 1, dJ = t >> 1;
                    const int nc = node2(iA - dI, jA - dJ);
                    if (nc < 0 || !cv[t]) continue;
                    acc[blk] = acc[blk] + ke2_val(Ke, blk, nr[t], nc, cid[t]);
                }
                if (add_basal && iA == 0 && jA == 0 && (blk == 0 || blk == 3) && av[q])
                    acc[blk] = acc[blk] + basal_scale * beta[q];
            }
            *reinterpret_cast<double2 *>(A + aidx(0, m, q)) = make_double2(acc[0], acc[1]);
            *reinterpret_cast<double2 *>(A + aidx(2, m, q)) = make_double2(acc[2], acc[3]);
        }
}

// every element layer takes its cell's efvs (efvs_qp is vertically uniform in SSA)
__global__ __launch_bounds__(256) void k_efvs_layers(long long ncell, int nz3, const double *__restrict__ efvs2,
                                                     double *__restrict__ efvs3) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell * (nz3 - 1)) return;
    efvs3[t] = efvs2[t / (nz3 - 1)];
}

int ensure_ssa(cism_ctx *c, int nz3) {
    const GridDev &g = c->g;
    const long long ncell = (long long)g.nx * g.ny, ncpad = (ncell + 63) / 64 * 64;
    if (!c->ssa_Ke) {
        HIP_TRY(hipMalloc(&c->ssa_Ke, sizeof(double) * KE2_ENTRIES * ncpad));
        HIP_TRY(hipMalloc(&c->ssa_ff2, sizeof(double) * ncell));
        HIP_TRY(hipMalloc(&c->ssa_efvs2, sizeof(double) * ncell));
        HIP_TRY(hipMalloc(&c->ssa_sigma, sizeof(double) * 64));
    }
    if (c->ssa_nz3 != nz3) {
        if (c->ssa_flw3) HIP_TRY(hipFree(c->ssa_flw3));
        c->ssa_flw3 = nullptr;
        HIP_TRY(hipMalloc(&c->ssa_flw3, sizeof(double) * (size_t)(nz3 - 1) * ncell));
        c->ssa_nz3 = nz3;
    }
    return CISM_OK;
}

// the column-mean flow factor from the 3D one (already in ssa_flw3, sigma in ssa_sigma)
int ssa_prepare_device(cism_ctx *c) {
    const GridDev &g = c->g;
    const long long ncell = (long long)g.nx * g.ny;
    hipLaunchKernelGGL(k_vertical_average, dim3(nblocks_for(ncell)), dim3(256), 0, c->stream, g.nx, g.ny,
                       c->ssa_nz3, c->ssa_sigma, c->ssa_flw3, c->ssa_ff2);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

int assemble2d_device(cism_ctx *c, int whichefvs, double efvs_constant, int add_basal, double dx, double dy) {
    const GridDev &g = c->g;
    const long long ncell = (long long)g.nx * g.ny;
    hipLaunchKernelGGL(k_element_2d, dim3(nblocks_for(ncell)), dim3(256), 0, c->stream, g, c->ac, c->xV, c->yV,
                       c->x, c->sthck, c->ssa_ff2, whichefvs, efvs_constant, c->ssa_Ke, c->ssa_efvs2, c->sc);
    HIP_TRY(hipGetLastError());
    c->a_zero_ok = 0;   // the 3D gather's record of zero rows does not cover this matrix
    hipLaunchKernelGGL(k_gather_2d, dim3(nblocks_for((long long)g.sx * g.sy)), dim3(256), 0, c->stream, g,
                       c->ac, c->av, c->ssa_Ke, c->beta, add_basal, dx * dy / CISM_VOL0, c->A);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

int ssa_efvs_layers_device(cism_ctx *c, double *efvs3) {
    const long long ncell = (long long)c->g.nx * c->g.ny;
    hipLaunchKernelGGL(k_efvs_layers, dim3(nblocks_for(ncell * (c->ssa_nz3 - 1))), dim3(256), 0, c->stream,
                       ncell, c->ssa_nz3, c->ssa_efvs2, efvs3);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

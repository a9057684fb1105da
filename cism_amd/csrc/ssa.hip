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
    c->a_zero_ok = 0;
    c->a64s_valid = 0;   // the 3D gather's record of zero rows does not cover this matrix
    c->a64_valid = 1;
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

// ---------------------------------------------------------------------------
// L1L2 (which_ho_approx = 3) and DIVA (= 4): the depth-integrated 2D solvers
// (velo_higher:1117-1173, 2703-2830, 3655-3700, 3773-3830, 5208-6356, 7382-7864).
//
// The 2D system is the SSA one (same element matrix, compute_element_matrix's
// Blatter-Pattyn form with dphi_dz = 0, of thickness x the sigma-weighted column mean
// of the layer viscosities); what differs is the viscosity of each layer:
//   L1L2: the real root tau_parallel of x^3 + tau_perp^2 x - effstrain/A = 0 with the
//         SIA stress tau_perp = rho g depth |grad s|, efvs = 1/(2A(tau_par^2+tau_perp^2));
//   DIVA: Glen's law with the vertical shear tau_b sigma / efvs_old of the last basal
//         traction, per quadrature point (efvs_qp kept across Picard iterations), plus
//         the column integrals omega_k, omega that give beta_eff = beta/(1 + beta omega)
//         and the 3D velocity after each solve.
// L1L2 integrates its 3D velocity from the bed after the loop.  One lane per cell for
// the element kernel, one per vertex for staggering and the column integrations.
// ---------------------------------------------------------------------------
struct DepthDev {
    int nz3;
    const double *sigma, *flwa, *ff3, *thck, *dsdx, *dsdy, *btr;   // btr: btractx, btracty planes
    double *efvs3, *eqp, *omk, *om;
};

template <int APPROX>
__global__ __launch_bounds__(256) void k_element_depth(GridDev g, DepthDev d, const int32_t *__restrict__ ac,
                                                       const double *__restrict__ xV,
                                                       const double *__restrict__ yV,
                                                       const double *__restrict__ vel,
                                                       const double *__restrict__ sthck, int whichefvs,
                                                       double efvs_constant, double *__restrict__ Ke,
                                                       DevScal *__restrict__ sc) {
    const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= (long long)g.nx * g.ny) return;
    const int i = (int)(cell % g.nx) + 1, j = (int)(cell / g.nx) + 1;   // 1-based
    const int L = d.nz3 - 1;
    const long long ncol = (long long)g.sx * g.sy;
    double *ef = d.efvs3 + (long long)L * cell;
    const bool ok = i >= CISM_NHALO + 1 && i <= g.nx - CISM_NHALO + 1 && j >= CISM_NHALO + 1 &&
                    j <= g.ny - CISM_NHALO + 1 && ac[cell] != 0;
    if (!ok) {
        for (int k = 0; k < L; ++k) ef[k] = 0.0;
        if (APPROX == 4) {
            for (int k = 0; k <= L; ++k) d.omk[k + (long long)d.nz3 * cell] = 0.0;
            d.om[cell] = 0.0;
        }
        return;
    }
    double x[4], y[4], u[4], v[4], hn[4], sx[4], sy[4], bx[4], by[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        const long long vtx = (i - 1 + c_di2[n]) + (long long)g.sx * (j - 1 + c_dj2[n]);
        x[n] = xV[vtx];
        y[n] = yV[vtx];
        u[n] = vel[2 * vtx];
        v[n] = vel[2 * vtx + 1];
        hn[n] = sthck[vtx];
        if (APPROX == 3) {
            sx[n] = d.dsdx[vtx];
            sy[n] = d.dsdy[vtx];
        } else {
            bx[n] = d.btr[vtx];
            by[n] = d.btr[ncol + vtx];
        }
    }
    const double *fl = d.flwa + (long long)L * cell, *ff = d.ff3 + (long long)L * cell;
    double *eq = d.eqp + (long long)L * 4 * cell;
    for (int k = 0; k < L; ++k) ef[k] = 0.0;
    double Kuu[10], Kvv[10], Kuv[16];
#pragma unroll
    for (int e = 0; e < 10; ++e) Kuu[e] = Kvv[e] = 0.0;
#pragma unroll
    for (int e = 0; e < 16; ++e) Kuv[e] = 0.0;
    const double r = 1.0 / sqrt(3.0);
    const double p_eff = (1.0 - 3.0) / 3.0, p2 = p_eff / 2, emin = 1.0e-8;
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
        double ux = 0, uy = 0, vx = 0, vy = 0, tsx = 0, tsy = 0, th = 0, tx = 0, ty = 0;
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            ux = ux + dx[n] * u[n];
            uy = uy + dy[n] * u[n];
            vx = vx + dx[n] * v[n];
            vy = vy + dy[n] * v[n];
            if (APPROX == 3) {
                tsx = tsx + phi[n] * sx[n];
                tsy = tsy + phi[n] * sy[n];
                th = th + phi[n] * hn[n];
            } else {
                tx = tx + phi[n] * bx[n];
                ty = ty + phi[n] * by[n];
            }
        }
        const double hsq = emin * emin + ux * ux + vy * vy + ux * vy + 0.25 * ((vx + uy) * (vx + uy));
        double avg = 0.0;
        for (int k = 0; k < L; ++k) {
            double e;
            if (whichefvs == 0) {
                e = efvs_constant;
            } else if (whichefvs == 1) {
                const double effstrain = (500.0 / CISM_SCYR) / 200.0e3 * CISM_SCYR;
                e = ff[k] * pow(effstrain, p_eff);
            } else if (APPROX == 3) {
                // compute_effective_viscosity_L1L2: Cardano's root of the cubic in tau_parallel
                const double effstrain = sqrt(hsq);
                const double depth = th * d.sigma[k + 1];
                const double grads = sqrt(tsx * tsx + tsy * tsy);
                const double tau_perp = CISM_RHOI * CISM_GRAV * depth * grads;
                const double a = tau_perp * tau_perp;
                const double b = -effstrain / fl[k];
                const double c = sqrt(b * b / 4.0 + a * a * a / 27.0);
                const double rootA = pow(-b / 2.0 + c, 1.0 / 3.0);
                const double rootB = (a * a * a / 27.0 > 1.0e-6 * (b * b / 4.0))
                                         ? -pow(b / 2.0 + c, 1.0 / 3.0)
                                         : -a / (3.0 * pow(fabs(b), 1.0 / 3.0));
                const double tp = rootA + rootB;
                e = 1.0 / (2.0 * fl[k] * (tp * tp + tau_perp * tau_perp));
            } else {
                // compute_effective_viscosity_diva (no cubic): vertical shear from the old efvs
                double eo = eq[k + (long long)L * p];
                if (eo == 0.0) eo = ff[k] * pow(emin, p_eff);
                const double ss = (d.sigma[k] + d.sigma[k + 1]) / 2.0;
                const double uz = tx * ss / eo, vz = ty * ss / eo;
                const double sq = hsq + 0.25 * (uz * uz + vz * vz);
                e = ff[k] * pow(sq, p2);
            }
            if (APPROX == 4) eq[k + (long long)L * p] = e;
            ef[k] = ef[k] + e;
            avg = avg + e * (d.sigma[k + 1] - d.sigma[k]);
        }
        double hq = 0.0;
#pragma unroll
        for (int n = 0; n < 4; ++n) hq = hq + phi[n] * hn[n];
        const double f = hq * avg * 1.0 * det / CISM_VOL0;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = a; b < 4; ++b) {
                Kuu[tri4(a, b)] = Kuu[tri4(a, b)] + f * (4.0 * dx[a] * dx[b] + dy[a] * dy[b] + 0.0 * 0.0);
                Kvv[tri4(a, b)] = Kvv[tri4(a, b)] + f * (4.0 * dy[a] * dy[b] + dx[a] * dx[b] + 0.0 * 0.0);
            }
#pragma unroll
        for (int nc = 0; nc < 4; ++nc)
#pragma unroll
            for (int nr = 0; nr < 4; ++nr)
                Kuv[nr + 4 * nc] = Kuv[nr + 4 * nc] + f * (2.0 * dx[nr] * dy[nc] + dy[nr] * dx[nc]);
    }
    for (int k = 0; k < L; ++k) ef[k] = ef[k] / 4;
    if (APPROX == 4) {
        // compute_integrals_diva: omega_k from the surface down, per quadrature point, then
        // their mean; omega the trapezoid integral of omega_k
        double o0 = 0.0, o1 = 0.0, o2 = 0.0, o3 = 0.0;
        double *ok = d.omk + (long long)d.nz3 * cell;
        ok[L] = (o0 + o1 + o2 + o3) / 4;
        const double H = d.thck[cell];
        for (int k = L - 1; k >= 0; --k) {
            const double depth = 0.5 * (d.sigma[k] + d.sigma[k + 1]);
            const double dz = (d.sigma[k + 1] - d.sigma[k]) * H;
            o0 = o0 + depth / eq[k] * dz;
            o1 = o1 + depth / eq[k + L] * dz;
            o2 = o2 + depth / eq[k + 2LL * L] * dz;
            o3 = o3 + depth / eq[k + 3LL * L] * dz;
            ok[k] = (o0 + o1 + o2 + o3) / 4;
        }
        double om = 0.0;
        for (int k = 0; k < L; ++k) om = om + 0.5 * (ok[k] + ok[k + 1]) * (d.sigma[k + 1] - d.sigma[k]);
        d.om[cell] = om;
    }
#pragma unroll
    for (int e = 0; e < 10; ++e) {
        Ke[ke2idx(e, cell)] = Kuu[e];
        Ke[ke2idx(10 + e, cell)] = Kvv[e];
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) Ke[ke2idx(20 + e, cell)] = Kuv[e];
}

// parallel_halo of a cell field (parallel_slap.F90:1794-1860), nhalo = 2: periodic
// images (owned cells only are read), or zero for the outflow global boundary
__global__ __launch_bounds__(256) void k_cell_halo(int nx, int ny, int periodic, double *__restrict__ a) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nx * ny) return;
    const int i = (int)(t % nx), j = (int)(t / nx);
    const bool hi = i < CISM_NHALO || i >= nx - CISM_NHALO, hj = j < CISM_NHALO || j >= ny - CISM_NHALO;
    if (!hi && !hj) return;
    if (!periodic) {
        a[t] = 0.0;
        return;
    }
    const int w = 2 * CISM_NHALO;
    const int ii = i < CISM_NHALO ? i + nx - w : (i >= nx - CISM_NHALO ? i - (nx - w) : i);
    const int jj = j < CISM_NHALO ? j + ny - w : (j >= ny - CISM_NHALO ? j - (ny - w) : j);
    a[t] = a[ii + (long long)nx * jj];
}

// glissade_stagger with stagger_margin_in = 1 (glissade_grid_operators.F90:60-148) at
// vertex (i, j) (0-based): the mean over the cells around it with mask = 1
__device__ inline double stagger_m1(int nx, const double *__restrict__ var, const int32_t *__restrict__ m, int i,
                                    int j, long long lead = 1, int k = 0) {
    const long long c00 = i + (long long)nx * j, c10 = c00 + 1, c01 = c00 + nx, c11 = c01 + 1;
    const double sv = (double)m[c01] * var[k + lead * c01] + (double)m[c11] * var[k + lead * c11] +
                      (double)m[c00] * var[k + lead * c00] + (double)m[c10] * var[k + lead * c10];
    const double sm = (double)(m[c01] + m[c11] + m[c00] + m[c10]);
    return sm > 0.0 ? sv / sm : 0.0;
}

// DIVA after the assembly (velo_higher:2742-2790): stag_omega over the ice-plus-land
// mask and beta_eff = beta / (1 + beta stag_omega), or 1/stag_omega under no-slip
__global__ __launch_bounds__(256) void k_beta_eff(GridDev g, const double *__restrict__ om,
                                                  const int32_t *__restrict__ ipl, const double *__restrict__ beta,
                                                  int noslip, double *__restrict__ som, double *__restrict__ beff) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)g.sx * g.sy) return;
    const int i = (int)(q % g.sx), j = (int)(q / g.sx);
    const double s = stagger_m1(g.nx, om, ipl, i, j);
    som[q] = s;
    if (noslip) beff[q] = s > 0.0 ? 1.0 / s : 0.0;
    else beff[q] = beta[q] / (1.0 + beta[q] * s);
}

// DIVA after each solve (velo_higher:3661-3700): basal traction beta_eff u on every
// vertex; on active vertices the bed velocity u - tau_b stag_omega and the levels above
// u_b + tau_b stag_omega_k(k) (compute_3d_velocity_diva, diva_level_index = 0)
__global__ __launch_bounds__(256) void k_diva_velocity(GridDev g, int nz3, const double *__restrict__ omk,
                                                       const int32_t *__restrict__ ipl,
                                                       const int32_t *__restrict__ av,
                                                       const double *__restrict__ beff,
                                                       const double *__restrict__ som,
                                                       const double *__restrict__ vel, double *__restrict__ btr,
                                                       double *__restrict__ u3, double *__restrict__ v3) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ncol = (long long)g.sx * g.sy;
    if (q >= ncol) return;
    const int i = (int)(q % g.sx), j = (int)(q / g.sx);
    const double u2 = vel[2 * q], v2 = vel[2 * q + 1];
    const double tx = beff[q] * u2, ty = beff[q] * v2;
    btr[q] = tx;
    btr[ncol + q] = ty;
    if (!av[q]) return;
    const double ub = u2 - tx * som[q], vb = v2 - ty * som[q];
    double *uc = u3 + (long long)nz3 * q, *vc = v3 + (long long)nz3 * q;
    uc[nz3 - 1] = ub;
    vc[nz3 - 1] = vb;
    for (int k = 0; k < nz3 - 1; ++k) {
        const double sk = stagger_m1(g.nx, omk, ipl, i, j, nz3, k);
        uc[k] = ub + tx * sk;
        vc[k] = vb + ty * sk;
    }
}

// staggered_parallel_halo of a (lead, nx-1, ny-1) array (parallel_slap.F90:2807-2925)
__global__ __launch_bounds__(256) void k_halo_lead(GridDev g, int lead, double *__restrict__ a) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ncol = (long long)g.sx * g.sy;
    if (t >= ncol * lead) return;
    const int k = (int)(t % lead);
    const long long q = t / lead;
    const int i = (int)(q % g.sx), j = (int)(q / g.sx);
    const int ii = halo_image(i, g.sx, g.periodic), jj = halo_image(j, g.sy, g.periodic);
    if (ii == i && jj == j) return;
    a[t] = (ii < 0 || jj < 0) ? 0.0 : a[k + (long long)lead * (ii + (long long)g.sx * jj)];
}

// velocity of owned inactive vertices is zero before the loop (velo_higher:1960-1972)
__global__ __launch_bounds__(256) void k_zero_inactive(GridDev g, int nz3, const int32_t *__restrict__ av,
                                                       double *__restrict__ vel, double *__restrict__ u3,
                                                       double *__restrict__ v3) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)g.sx * g.sy) return;
    const int i = (int)(q % g.sx), j = (int)(q / g.sx);
    if (!is_owned(g, i, j) || av[q]) return;
    vel[2 * q] = 0.0;
    vel[2 * q + 1] = 0.0;
    for (int k = 0; k < nz3; ++k) {
        u3[k + (long long)nz3 * q] = 0.0;
        v3[k + (long long)nz3 * q] = 0.0;
    }
}

// compute_3d_velocity_L1L2, the cell part (velo_higher:5990-6050): strain rates of the
// basal (2D) velocity at the center of each active cell of the assembly's range and the
// parallel effective strain rate; zero elsewhere.  cs: 5 planes ux uy vx vy eps
__global__ __launch_bounds__(256) void k_l1l2_cells(GridDev g, const int32_t *__restrict__ ac,
                                                    const double *__restrict__ xV, const double *__restrict__ yV,
                                                    const double *__restrict__ vel, double *__restrict__ cs,
                                                    DevScal *__restrict__ sc) {
    const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ncell = (long long)g.nx * g.ny;
    if (cell >= ncell) return;
    const int i = (int)(cell % g.nx) + 1, j = (int)(cell / g.nx) + 1;
    const bool ok = i >= CISM_NHALO + 1 && i <= g.nx - CISM_NHALO + 1 && j >= CISM_NHALO + 1 &&
                    j <= g.ny - CISM_NHALO + 1 && ac[cell] != 0;
    double ux = 0, uy = 0, vx = 0, vy = 0, eps = 0;
    if (ok) {
        double x[4], y[4], u[4], v[4];
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            const long long vtx = (i - 1 + c_di2[n]) + (long long)g.sx * (j - 1 + c_dj2[n]);
            x[n] = xV[vtx];
            y[n] = yV[vtx];
            u[n] = vel[2 * vtx];
            v[n] = vel[2 * vtx + 1];
        }
        const double dxr[4] = {-0.25, 0.25, 0.25, -0.25}, dyr[4] = {-0.25, -0.25, 0.25, 0.25};
        double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            J11 = J11 + dxr[n] * x[n];
            J12 = J12 + dxr[n] * y[n];
            J21 = J21 + dyr[n] * x[n];
            J22 = J22 + dyr[n] * y[n];
        }
        const double det = J11 * J22 - J12 * J21;
        if (fabs(det) > 0.0) {
            const double i11 = J22 / det, i12 = -J12 / det, i21 = -J21 / det, i22 = J11 / det;
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                const double dx = i11 * dxr[n] + i12 * dyr[n], dy = i21 * dxr[n] + i22 * dyr[n];
                ux = ux + dx * u[n];
                uy = uy + dy * u[n];
                vx = vx + dx * v[n];
                vy = vy + dy * v[n];
            }
        } else {
            atomicOr(&sc->singular_flag, 1);
        }
        eps = sqrt(ux * ux + vy * vy + ux * vy + 0.25 * ((vx + uy) * (vx + uy)));
    }
    cs[cell] = ux;
    cs[ncell + cell] = uy;
    cs[2 * ncell + cell] = vx;
    cs[3 * ncell + cell] = vy;
    cs[4 * ncell + cell] = eps;
}

// compute_3d_velocity_L1L2, the column part (velo_higher:6060-6350, method 1): one lane
// per vertex.  Owned active vertices: first, top down, the integral of efvs from each
// layer's base to the surface in the four cells around, the centered gradients of the
// depth-integrated membrane stresses (glissade_gradient over ice cells) and the vertical
// shear stresses tau_xz, tau_yz (kept in u3/v3 meanwhile); then, bottom up, u(k) =
// u(k+1) + 2 A tau_eff^2 dz tau_xz with A and tau_parallel^2 staggered over ice cells.
// Every other vertex gets zero above the bed.
__global__ __launch_bounds__(256) void k_l1l2_column(GridDev g, DepthDev d, const int32_t *__restrict__ ice,
                                                     const int32_t *__restrict__ ac,
                                                     const int32_t *__restrict__ av,
                                                     const int32_t *__restrict__ um,
                                                     const int32_t *__restrict__ vm,
                                                     const double *__restrict__ sthck,
                                                     const double *__restrict__ cs, double dx, double dy,
                                                     double *__restrict__ u3, double *__restrict__ v3) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ncol = (long long)g.sx * g.sy, ncell = (long long)g.nx * g.ny;
    if (q >= ncol) return;
    const int i = (int)(q % g.sx), j = (int)(q / g.sx);
    const int nz3 = d.nz3, L = nz3 - 1;
    double *uc = u3 + (long long)nz3 * q, *vc = v3 + (long long)nz3 * q;
    if (!is_owned(g, i, j) || !av[q]) {
        for (int k = 0; k < L; ++k) uc[k] = vc[k] = 0.0;
        return;
    }
    // cells SW, SE, NW, NE of the vertex (0-based (i,j), (i+1,j), (i,j+1), (i+1,j+1))
    long long cid[4];
    bool act[4];
    int m[4];
    double ux[4], uy[4], vx[4], vy[4], eps[4], H[4], eint[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int ic = i + (t & 1), jc = j + (t >> 1);
        cid[t] = ic + (long long)g.nx * jc;
        act[t] = ic >= CISM_NHALO && ic <= g.nx - CISM_NHALO && jc >= CISM_NHALO && jc <= g.ny - CISM_NHALO &&
                 ac[cid[t]] != 0;
        m[t] = ice[cid[t]];
        ux[t] = cs[cid[t]];
        uy[t] = cs[ncell + cid[t]];
        vx[t] = cs[2 * ncell + cid[t]];
        vy[t] = cs[3 * ncell + cid[t]];
        eps[t] = cs[4 * ncell + cid[t]];
        H[t] = d.thck[cid[t]];
        eint[t] = 0.0;
    }
    // edge masks of glissade_gradient (gradient_margin_in = 1)
    const bool exs = m[0] == 1 && m[1] == 1, exn = m[2] == 1 && m[3] == 1;
    const bool eyw = m[0] == 1 && m[2] == 1, eye = m[1] == 1 && m[3] == 1;
    const double sthk = sthck[q];
    for (int k = 0; k < L; ++k) {
        double w1[4], w2[4], w3[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            if (act[t]) eint[t] = eint[t] + d.efvs3[k + (long long)L * cid[t]] * (d.sigma[k + 1] - d.sigma[k]) * H[t];
            w1[t] = eint[t] * (2.0 * ux[t] + vy[t]);
            w2[t] = eint[t] * (uy[t] + vx[t]);
            w3[t] = eint[t] * (2.0 * vy[t] + ux[t]);
        }
        auto gx = [&](const double *w) {
            const double xn = (w[3] - w[2]) / dx, xs = (w[1] - w[0]) / dx;
            return (exs && exn) ? (xn + xs) / 2.0 : exs ? xs : exn ? xn : 0.0;
        };
        auto gy = [&](const double *w) {
            const double ye = (w[3] - w[1]) / dy, yw = (w[2] - w[0]) / dy;
            return (eyw && eye) ? (ye + yw) / 2.0 : eyw ? yw : eye ? ye : 0.0;
        };
        const double depth = 0.5 * (d.sigma[k] + d.sigma[k + 1]) * sthk;
        uc[k] = -(CISM_RHOI * CISM_GRAV * depth * d.dsdx[q]) + 2.0 * gx(w1) + gy(w2);
        vc[k] = -(CISM_RHOI * CISM_GRAV * depth * d.dsdy[q]) + gx(w2) + 2.0 * gy(w3);
    }
    const bool umk = um[q] == 1, vmk = vm[q] == 1;
    for (int k = L - 1; k >= 0; --k) {
        double tp[4], fl[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            tp[t] = act[t] ? 2.0 * d.efvs3[k + (long long)L * cid[t]] * eps[t] : 0.0;
            fl[t] = d.flwa[k + (long long)L * cid[t]];
        }
        // staggering order of glissade_stagger: (i,j+1), (i+1,j+1), (i,j), (i+1,j)
        const double sm = (double)(m[2] + m[3] + m[0] + m[1]);
        double st = 0.0, sf = 0.0;
        if (sm > 0.0) {
            st = ((double)m[2] * tp[2] + (double)m[3] * tp[3] + (double)m[0] * tp[0] + (double)m[1] * tp[1]) / sm;
            sf = ((double)m[2] * fl[2] + (double)m[3] * fl[3] + (double)m[0] * fl[0] + (double)m[1] * fl[1]) / sm;
        }
        st = st * st;
        const double txz = uc[k], tyz = vc[k];
        // tau_eff_sq**((gn-1)/2) with gn = 3 is tau_eff_sq itself
        const double tes = st + txz * txz + tyz * tyz;
        const double fact = 2.0 * sf * tes * (d.sigma[k + 1] - d.sigma[k]) * sthk;
        uc[k] = umk ? uc[L] : uc[k + 1] + fact * txz;
        vc[k] = vmk ? vc[L] : vc[k + 1] + fact * tyz;
    }
}

static DepthDev depth_dev(cism_ctx *c) {
    DepthDev d;
    d.nz3 = c->dep_nz3;
    d.sigma = c->ssa_sigma;
    d.flwa = c->dep_flwa;
    d.ff3 = c->ssa_flw3;
    d.thck = c->dep_thck;
    d.dsdx = c->dsdx;
    d.dsdy = c->dsdy;
    d.btr = c->dep_btr;
    d.efvs3 = c->dep_efvs3;
    d.eqp = c->dep_eqp;
    d.omk = c->dep_omk;
    d.om = c->dep_om;
    return d;
}

int ensure_depth(cism_ctx *c, int nz3) {
    CISM_TRY(ensure_ssa(c, nz3));
    if (c->dep_nz3 == nz3) return CISM_OK;
    double **bufs[] = {&c->dep_flwa, &c->dep_thck, &c->dep_efvs3, &c->dep_eqp, &c->dep_omk, &c->dep_om,
                       &c->dep_beff, &c->dep_som, &c->dep_btr, &c->dep_u3, &c->dep_v3, &c->dep_cs,
                       &c->dep_ub};
    for (double **b : bufs)
        if (*b) {
            HIP_TRY(hipFree(*b));
            *b = nullptr;
        }
    if (c->dep_ice) HIP_TRY(hipFree(c->dep_ice));
    if (c->dep_ipl) HIP_TRY(hipFree(c->dep_ipl));
    c->dep_ice = c->dep_ipl = nullptr;
    c->dep_nz3 = 0;
    const GridDev &g = c->g;
    const size_t ncell = (size_t)g.nx * g.ny, ncol = (size_t)g.sx * g.sy, L = nz3 - 1;
    const size_t sizes[] = {L * ncell, ncell, L * ncell, 4 * L * ncell, (size_t)nz3 * ncell, ncell,
                            ncol, ncol, 2 * ncol, (size_t)nz3 * ncol, (size_t)nz3 * ncol, 5 * ncell, 2 * ncol};
    for (int t = 0; t < 13; ++t) HIP_TRY(hipMalloc(bufs[t], sizeof(double) * sizes[t]));
    HIP_TRY(hipMalloc(&c->dep_ice, sizeof(int32_t) * ncell));
    HIP_TRY(hipMalloc(&c->dep_ipl, sizeof(int32_t) * ncell));
    HIP_TRY(hipMemsetAsync(c->dep_om, 0, sizeof(double) * ncell, c->stream));
    HIP_TRY(hipMemsetAsync(c->dep_omk, 0, sizeof(double) * nz3 * ncell, c->stream));
    c->dep_nz3 = nz3;
    return CISM_OK;
}

// DIVA's efvs_qp of every quadrature point from the caller's efvs (velo_higher:1168-1171)
__global__ __launch_bounds__(256) void k_eqp_init(long long ncell, int L, const double *__restrict__ efvs3,
                                                  double *__restrict__ eqp) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell * 4 * L) return;
    const int k = (int)(t % L);
    const long long cell = t / (4LL * L);
    eqp[t] = efvs3[k + (long long)L * cell];
}

int depth_prepare_device(cism_ctx *c, int approx) {
    const GridDev &g = c->g;
    const long long ncell = (long long)g.nx * g.ny;
    const int L = c->dep_nz3 - 1;
    if (approx == 4) {
        hipLaunchKernelGGL(k_eqp_init, dim3(nblocks_for(ncell * 4 * L)), dim3(256), 0, c->stream, ncell, L,
                           c->dep_efvs3, c->dep_eqp);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_zero_inactive, dim3(nblocks_for((long long)g.sx * g.sy)), dim3(256), 0, c->stream, g,
                       c->dep_nz3, c->av, c->x, c->dep_u3, c->dep_v3);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// the L1L2/DIVA assembly: element kernel, (DIVA) omega halo, stag_omega and beta_eff,
// then the node gather with the local basal friction when add_basal
int assemble_depth_device(cism_ctx *c, int approx, int whichefvs, double efvs_constant, int add_basal,
                          int noslip, double dx, double dy) {
    const GridDev &g = c->g;
    const long long ncell = (long long)g.nx * g.ny, ncol = (long long)g.sx * g.sy;
    const DepthDev d = depth_dev(c);
    if (approx == 3)
        hipLaunchKernelGGL(k_element_depth<3>, dim3(nblocks_for(ncell)), dim3(256), 0, c->stream, g, d, c->ac,
                           c->xV, c->yV, c->x, c->sthck, whichefvs, efvs_constant, c->ssa_Ke, c->sc);
    else
        hipLaunchKernelGGL(k_element_depth<4>, dim3(nblocks_for(ncell)), dim3(256), 0, c->stream, g, d, c->ac,
                           c->xV, c->yV, c->x, c->sthck, whichefvs, efvs_constant, c->ssa_Ke, c->sc);
    HIP_TRY(hipGetLastError());
    if (approx == 4) {
        hipLaunchKernelGGL(k_cell_halo, dim3(nblocks_for(ncell)), dim3(256), 0, c->stream, g.nx, g.ny, g.periodic,
                           c->dep_om);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_beta_eff, dim3(nblocks_for(ncol)), dim3(256), 0, c->stream, g, c->dep_om, c->dep_ipl,
                           c->beta, noslip, c->dep_som, c->dep_beff);
        HIP_TRY(hipGetLastError());
    }
    c->a_zero_ok = 0;
    c->a64s_valid = 0;
    c->a64_valid = 1;
    hipLaunchKernelGGL(k_gather_2d, dim3(nblocks_for(ncol)), dim3(256), 0, c->stream, g, c->ac, c->av, c->ssa_Ke,
                       approx == 4 ? c->dep_beff : c->beta, add_basal, dx * dy / CISM_VOL0, c->A);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

const double *depth_basal_beta(cism_ctx *c, int approx) { return approx == 4 ? c->dep_beff : c->beta; }

// DIVA after each solve: basal traction, the 3D velocity, its halo
int diva_post_solve_device(cism_ctx *c) {
    const GridDev &g = c->g;
    const long long ncol = (long long)g.sx * g.sy;
    const int nz3 = c->dep_nz3;
    hipLaunchKernelGGL(k_diva_velocity, dim3(nblocks_for(ncol)), dim3(256), 0, c->stream, g, nz3, c->dep_omk,
                       c->dep_ipl, c->av, c->dep_beff, c->dep_som, c->x, c->dep_btr, c->dep_u3, c->dep_v3);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_halo_lead, dim3(nblocks_for(ncol * nz3)), dim3(256), 0, c->stream, g, nz3, c->dep_u3);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_halo_lead, dim3(nblocks_for(ncol * nz3)), dim3(256), 0, c->stream, g, nz3, c->dep_v3);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// L1L2 after the loop: the bed level from the 2D velocity, the levels above by
// compute_3d_velocity_L1L2, the halo
__global__ __launch_bounds__(256) void k_bed_level(long long ncol, int nz3, const double *__restrict__ vel,
                                                   double *__restrict__ u3, double *__restrict__ v3) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ncol) return;
    u3[(nz3 - 1) + (long long)nz3 * q] = vel[2 * q];
    v3[(nz3 - 1) + (long long)nz3 * q] = vel[2 * q + 1];
}

int l1l2_velocity_device(cism_ctx *c, double dx, double dy) {
    const GridDev &g = c->g;
    const long long ncol = (long long)g.sx * g.sy, ncell = (long long)g.nx * g.ny;
    const int nz3 = c->dep_nz3;
    hipLaunchKernelGGL(k_bed_level, dim3(nblocks_for(ncol)), dim3(256), 0, c->stream, ncol, nz3, c->x,
                       c->dep_u3, c->dep_v3);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_l1l2_cells, dim3(nblocks_for(ncell)), dim3(256), 0, c->stream, g, c->ac, c->xV, c->yV,
                       c->x, c->dep_cs, c->sc);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_l1l2_column, dim3(nblocks_for(ncol)), dim3(256), 0, c->stream, g, depth_dev(c),
                       c->dep_ice, c->ac, c->av, c->umask, c->vmask, c->sthck, c->dep_cs, dx, dy, c->dep_u3,
                       c->dep_v3);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_halo_lead, dim3(nblocks_for(ncol * nz3)), dim3(256), 0, c->stream, g, nz3, c->dep_u3);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_halo_lead, dim3(nblocks_for(ncol * nz3)), dim3(256), 0, c->stream, g, nz3, c->dep_v3);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

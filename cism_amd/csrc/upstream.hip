// upstream.hip -- the velocity solve's inputs computed on the device from the
// ice geometry (glissade_velo_higher_solve's set-up, velo_higher:1449-1939):
// masks, staggered thickness and surface, the surface-elevation gradient,
// vertex geometry and active cells/vertices, and the flow factor.  The results
// stay in the context (the solve reads them from there) and are optionally
// copied to the caller.
#include "internal.h"

#include <math.h>

// glissade_get_masks without a subgrid calving front (glissade_masks.F90:130-210):
// bit 0 ice, 1 floating, 2 ocean, 3 land, 4 active, 5 ice-or-land
__global__ __launch_bounds__(256) void k_masks(GridDev g, const double *__restrict__ thck,
                                               const double *__restrict__ topg, double eus,
                                               double thklim, int32_t *__restrict__ bits) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.nx * g.ny) return;
    const double H = thck[t], b = topg[t];
    const int ice = H > thklim;
    const int ocean = b < eus && !ice;
    const int floating = (b - eus < (-CISM_RHOI / CISM_RHOO) * H) && ice;
    const int land = b >= eus;
    bits[t] = ice | floating << 1 | ocean << 2 | land << 3 | ice << 4 | (ice || land) << 5;
}

// glissade_stagger, stagger_margin_in = 1 (glissade_grid_operators.F90:118-146), for
// thickness and surface together; vertex (i, j) (0-based) averages the cells of
// ice_plus_land_mask among (i..i+1, j..j+1) in the order (i,j+1), (i+1,j+1), (i,j), (i+1,j)
__global__ __launch_bounds__(256) void k_stagger(GridDev g, const double *__restrict__ thck,
                                                 const double *__restrict__ usrf,
                                                 const int32_t *__restrict__ bits,
                                                 double *__restrict__ sthck, double *__restrict__ susrf) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (long long)g.sx * g.sy) return;
    const int i = (int)(v % g.sx), j = (int)(v / g.sx);
    const long long c[4] = {i + (long long)g.nx * (j + 1), (i + 1) + (long long)g.nx * (j + 1),
                            i + (long long)g.nx * j, (i + 1) + (long long)g.nx * j};
    double st = 0.0, ss = 0.0, sm = 0.0;
    for (int q = 0; q < 4; ++q) {
        const double m = (bits[c[q]] >> 5) & 1;
        st = st + m * thck[c[q]];
        ss = ss + m * usrf[c[q]];
    }
    for (int q = 0; q < 4; ++q) sm += (bits[c[q]] >> 5) & 1;
    sthck[v] = sm > 0 ? st / sm : 0.0;
    susrf[v] = sm > 0 ? ss / sm : 0.0;
}

// surface-elevation gradient across the edge between cells a (lower index) and b:
// HO_GRADIENT_CENTERED with HO_GRADIENT_MARGIN_HYBRID (glissade_grid_operators.F90:830-990):
// taken where the upper cell is active and the lower one active or land; a land upper
// cell at or below thklim contributes 0
__device__ inline double edge_grad(const double *usrf, const double *thck, const int32_t *bits,
                                   long long a, long long b, double h, double thklim) {
    const bool up_a = usrf[a] > usrf[b];
    const long long u = up_a ? a : b, l = up_a ? b : a;
    const int bu = bits[u], bl = bits[l];
    const bool ok = ((bu >> 4) & 1) && (((bl >> 4) & 1) || ((bl >> 3) & 1));
    if (!ok) return 0.0;
    const double ef = ((bu >> 3) & 1) ? (thck[u] > thklim ? 1.0 : 0.0) : 1.0;
    const double sign = up_a ? -1.0 : 1.0;
    return ef * sign * (usrf[u] - usrf[l]) / h;
}

// ds/dx, ds/dy at vertices with the max_slope limit; halo vertices take their periodic
// image's value (staggered_parallel_halo) or 0 on an outflow grid
__global__ __launch_bounds__(256) void k_surface_gradient(GridDev g, const double *__restrict__ usrf,
                                                          const double *__restrict__ thck,
                                                          const int32_t *__restrict__ bits, double dx,
                                                          double dy, double thklim, double max_slope,
                                                          double *__restrict__ dsdx,
                                                          double *__restrict__ dsdy) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (long long)g.sx * g.sy) return;
    const int i0 = (int)(v % g.sx), j0 = (int)(v / g.sx);
    const int i = halo_image(i0, g.sx, g.periodic), j = halo_image(j0, g.sy, g.periodic);
    double gx = 0.0, gy = 0.0;
    if (i >= 0 && j >= 0) {
        auto cell = [&](int ci, int cj) { return ci + (long long)g.nx * cj; };
        const double ex0 = edge_grad(usrf, thck, bits, cell(i, j), cell(i + 1, j), dx, thklim);
        const double ex1 = edge_grad(usrf, thck, bits, cell(i, j + 1), cell(i + 1, j + 1), dx, thklim);
        const double ey0 = edge_grad(usrf, thck, bits, cell(i, j), cell(i, j + 1), dy, thklim);
        const double ey1 = edge_grad(usrf, thck, bits, cell(i + 1, j), cell(i + 1, j + 1), dy, thklim);
        gx = 0.5 * (ex0 + ex1);
        gy = 0.5 * (ey0 + ey1);
        gx = gx > 0.0 ? fmin(gx, max_slope) : fmax(gx, -max_slope);
        gy = gy > 0.0 ? fmin(gy, max_slope) : fmax(gy, -max_slope);
    }
    dsdx[v] = gx;
    dsdy[v] = gy;
}

// get_vertex_geometry (velo_higher:4163-4313): coordinates; active cells are active
// ice in cells nhalo+1 .. n-nhalo+1 (1-based); active vertices are their corners
__global__ __launch_bounds__(256) void k_active_cells(GridDev g, const int32_t *__restrict__ bits,
                                                      int32_t *__restrict__ ac) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.nx * g.ny) return;
    const int i = (int)(t % g.nx) + 1, j = (int)(t / g.nx) + 1;
    const bool in = i >= CISM_NHALO + 1 && i <= g.nx - CISM_NHALO + 1 && j >= CISM_NHALO + 1 &&
                    j <= g.ny - CISM_NHALO + 1;
    ac[t] = in && ((bits[t] >> 4) & 1);
}

__global__ __launch_bounds__(256) void k_vertex_geometry(GridDev g, double dx, double dy,
                                                         const int32_t *__restrict__ ac,
                                                         double *__restrict__ xV, double *__restrict__ yV,
                                                         int32_t *__restrict__ av) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (long long)g.sx * g.sy) return;
    const int i = (int)(v % g.sx), j = (int)(v / g.sx);
    xV[v] = dx * (i + 1);
    yV[v] = dy * (j + 1);
    const long long c = i + (long long)g.nx * j;
    av[v] = (ac[c] | ac[c + 1] | ac[c + g.nx] | ac[c + g.nx + 1]) != 0;
}

// flwafact = 0.5 A^(-1/n), n = 3, on cells nhalo+1 .. n-nhalo+1 where A > 0
// (velo_higher:1930-1939); layer-fastest like the reference's (nz-1, nx, ny)
__global__ __launch_bounds__(256) void k_flwafact(GridDev g, const double *__restrict__ flwa,
                                                  double *__restrict__ ff) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.ne) return;
    const long long cell = e / (g.nz - 1);
    const int i = (int)(cell % g.nx) + 1, j = (int)(cell / g.nx) + 1;
    const bool in = i >= CISM_NHALO + 1 && i <= g.nx - CISM_NHALO + 1 && j >= CISM_NHALO + 1 &&
                    j <= g.ny - CISM_NHALO + 1;
    const double A = flwa[e];
    ff[e] = in && A > 0.0 ? 0.5 * pow(A, -1.0 / 3.0) : 0.0;
}

static int fetch(cism_ctx *c, void *host, const void *dev, size_t bytes) {
    if (host) HIP_TRY(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, c->stream));
    return CISM_OK;
}

__global__ void k_bit(long long n, const int32_t *__restrict__ bits, int b, int32_t *__restrict__ out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = (bits[t] >> b) & 1;
}

extern "C" int cism_glissade_velo_inputs(cism_ctx *c, double dx, double dy, double thklim, double eus,
                                         double max_slope, const double *thck, const double *usrf,
                                         const double *topg, const double *flwa, int32_t *ice_mask,
                                         int32_t *floating_mask, int32_t *ocean_mask, int32_t *land_mask,
                                         int32_t *active_cell, int32_t *active_vertex, double *xVertex,
                                         double *yVertex, double *stagusrf, double *stagthck,
                                         double *dusrf_dx, double *dusrf_dy, double *flwafact) {
    if (!c || !thck || !usrf || !topg || !flwa || !(dx > 0) || !(dy > 0) || !(max_slope > 0)) {
        cism_set_error("glissade_velo_inputs: invalid arguments");
        return CISM_ERR_ARG;
    }
    HIP_TRY(hipSetDevice(c->device));
    const GridDev &g = c->g;
    if (g.nz < 2) {
        cism_set_error("glissade_velo_inputs needs a 3D context (nz >= 2)");
        return CISM_ERR_ARG;
    }
    const long long ncell = (long long)g.nx * g.ny, ncol = (long long)g.sx * g.sy;
    // device staging: cell fields in the element scratch (allocated by the first assembly
    // or here), bits in the node-mask buffer of u (nn >= ncell ints)
    CISM_TRY(ensure_ke(c));
    double *d_thck = c->Ke, *d_usrf = c->Ke + ncell, *d_topg = c->Ke + 2 * ncell, *d_flwa = c->Ke + 3 * ncell;
    int32_t *bits = c->umask;   // overwritten by the next solve's Dirichlet masks
    HIP_TRY(hipMemcpyAsync(d_thck, thck, sizeof(double) * ncell, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(d_usrf, usrf, sizeof(double) * ncell, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(d_topg, topg, sizeof(double) * ncell, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(d_flwa, flwa, sizeof(double) * g.ne, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_masks, dim3(nblocks_for(ncell)), dim3(256), 0, c->stream, g, d_thck, d_topg, eus,
                       thklim, bits);
    hipLaunchKernelGGL(k_stagger, dim3(nblocks_for(ncol)), dim3(256), 0, c->stream, g, d_thck, d_usrf, bits,
                       c->sthck, c->susrf);
    hipLaunchKernelGGL(k_surface_gradient, dim3(nblocks_for(ncol)), dim3(256), 0, c->stream, g, d_usrf,
                       d_thck, bits, dx, dy, thklim, max_slope, c->dsdx, c->dsdy);
    hipLaunchKernelGGL(k_active_cells, dim3(nblocks_for(ncell)), dim3(256), 0, c->stream, g, bits, c->ac);
    hipLaunchKernelGGL(k_vertex_geometry, dim3(nblocks_for(ncol)), dim3(256), 0, c->stream, g, dx, dy, c->ac,
                       c->xV, c->yV, c->av);
    hipLaunchKernelGGL(k_flwafact, dim3(nblocks_for(g.ne)), dim3(256), 0, c->stream, g, d_flwa, c->flwafact);
    HIP_TRY(hipGetLastError());
    int32_t *masks[4] = {ice_mask, floating_mask, ocean_mask, land_mask};
    int32_t *mbuf = c->vmask;   // scratch for one mask at a time
    for (int b = 0; b < 4; ++b) {
        if (!masks[b]) continue;
        hipLaunchKernelGGL(k_bit, dim3(nblocks_for(ncell)), dim3(256), 0, c->stream, ncell, bits, b, mbuf);
        HIP_TRY(hipGetLastError());
        CISM_TRY(fetch(c, masks[b], mbuf, sizeof(int32_t) * ncell));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    CISM_TRY(fetch(c, active_cell, c->ac, sizeof(int32_t) * ncell));
    CISM_TRY(fetch(c, active_vertex, c->av, sizeof(int32_t) * ncol));
    CISM_TRY(fetch(c, xVertex, c->xV, sizeof(double) * ncol));
    CISM_TRY(fetch(c, yVertex, c->yV, sizeof(double) * ncol));
    CISM_TRY(fetch(c, stagusrf, c->susrf, sizeof(double) * ncol));
    CISM_TRY(fetch(c, stagthck, c->sthck, sizeof(double) * ncol));
    CISM_TRY(fetch(c, dusrf_dx, c->dsdx, sizeof(double) * ncol));
    CISM_TRY(fetch(c, dusrf_dy, c->dsdy, sizeof(double) * ncol));
    CISM_TRY(fetch(c, flwafact, c->flwafact, sizeof(double) * g.ne));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CISM_OK;
}

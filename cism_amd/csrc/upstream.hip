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

// The inputs of the solve from the cell fields thck, usrf, topg (nx, ny) and flwa
// (nz-1, nx, ny), all host arrays: results in the context (sthck, susrf, dsdx, dsdy, ac,
// xV, yV, av, flwafact); the cell masks as bits (k_masks) in c->umask, until the next
// writer of the Dirichlet masks.
int velo_inputs_device(cism_ctx *c, double dx, double dy, double thklim, double eus, double max_slope,
                       const double *thck, const double *usrf, const double *topg, const double *flwa) {
    CISM_TRY(velo_inputs_upload(c, thck, usrf, topg, flwa));
    return velo_inputs_compute(c, dx, dy, thklim, eus, max_slope);
}

// the cell fields into device staging: the element scratch (allocated by the first
// assembly or here) holds thck, usrf, topg and flwa until the next assembly
int velo_inputs_upload(cism_ctx *c, const double *thck, const double *usrf, const double *topg,
                       const double *flwa) {
    const GridDev &g = c->g;
    if (!thck || !usrf || !topg || !flwa) {
        cism_set_error("glissade_velo_inputs: null cell field");
        return CISM_ERR_ARG;
    }
    if (g.nz < 2) {
        cism_set_error("glissade_velo_inputs needs a 3D context (nz >= 2)");
        return CISM_ERR_ARG;
    }
    const long long ncell = (long long)g.nx * g.ny;
    CISM_TRY(ensure_ke(c));
    HIP_TRY(hipMemcpyAsync(c->Ke, thck, sizeof(double) * ncell, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->Ke + ncell, usrf, sizeof(double) * ncell, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->Ke + 2 * ncell, topg, sizeof(double) * ncell, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->Ke + 3 * ncell, flwa, sizeof(double) * g.ne, hipMemcpyHostToDevice, c->stream));
    return CISM_OK;
}

// the inputs from the staged cell fields; the mask bits go to the node-mask buffer of u
// (nn >= ncell ints), overwritten by the next solve's Dirichlet masks
int velo_inputs_compute(cism_ctx *c, double dx, double dy, double thklim, double eus, double max_slope) {
    if (!(dx > 0) || !(dy > 0) || !(max_slope > 0)) {
        cism_set_error("glissade_velo_inputs: invalid arguments");
        return CISM_ERR_ARG;
    }
    const GridDev &g = c->g;
    const long long ncell = (long long)g.nx * g.ny, ncol = (long long)g.sx * g.sy;
    double *d_thck = c->Ke, *d_usrf = c->Ke + ncell, *d_topg = c->Ke + 2 * ncell, *d_flwa = c->Ke + 3 * ncell;
    int32_t *bits = c->umask;
    c->dirichlet_any = 1;   // the masks are scratch now: no longer known empty
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
    return CISM_OK;
}

extern "C" int cism_glissade_velo_inputs(cism_ctx *c, double dx, double dy, double thklim, double eus,
                                         double max_slope, const double *thck, const double *usrf,
                                         const double *topg, const double *flwa, int32_t *ice_mask,
                                         int32_t *floating_mask, int32_t *ocean_mask, int32_t *land_mask,
                                         int32_t *active_cell, int32_t *active_vertex, double *xVertex,
                                         double *yVertex, double *stagusrf, double *stagthck,
                                         double *dusrf_dx, double *dusrf_dy, double *flwafact) {
    if (!c) {
        cism_set_error("glissade_velo_inputs: null context");
        return CISM_ERR_ARG;
    }
    HIP_TRY(hipSetDevice(c->device));
    CISM_TRY(velo_inputs_device(c, dx, dy, thklim, eus, max_slope, thck, usrf, topg, flwa));
    const GridDev &g = c->g;
    const long long ncell = (long long)g.nx * g.ny, ncol = (long long)g.sx * g.sy;
    const int32_t *bits = c->umask;
    int32_t *masks[4] = {ice_mask, floating_mask, ocean_mask, land_mask};
    int32_t *mbuf = c->vmask;   // scratch for one mask at a time
    c->dirichlet_any = 1;   // the masks are scratch now: no longer known empty
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

// ---------------------------------------------------------------------------
// grounded fraction and basal traction
// ---------------------------------------------------------------------------
// glissade_grounded_fraction (glissade_grounding_line.F90:55-178), whichground 0
// (NO_GLP) or 2 (ALL).  The vertex mask tests land(i+1, j+1) twice and never
// land(i+1, j), as the reference's condition does.
__global__ __launch_bounds__(256) void k_grounded_fraction(GridDev g, const int32_t *__restrict__ ice,
                                                           const int32_t *__restrict__ fl,
                                                           const int32_t *__restrict__ land, int whichground,
                                                           double *__restrict__ fg) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (long long)g.sx * g.sy) return;
    const int i = (int)(v % g.sx), j = (int)(v / g.sx);
    const long long sw = i + (long long)g.nx * j, se = sw + 1, nw = sw + g.nx, ne = nw + 1;
    const bool vm = ice[nw] == 1 || ice[ne] == 1 || ice[sw] == 1 || ice[se] == 1 || land[nw] == 1 ||
                    land[ne] == 1 || land[sw] == 1 || land[ne] == 1;
    double f = 0.0;
    if (vm) {
        if (whichground == 2) {
            f = 1.0;
        } else {
            auto cg = [&](long long c) { return (ice[c] == 1 && fl[c] == 0) || land[c] == 1; };
            f = (cg(nw) || cg(ne) || cg(sw) || cg(se)) ? 1.0 : 0.0;
        }
    }
    fg[v] = f;
}

// calcbeta (glissade_basal_traction.F90:67-598) on every vertex: the basal speed
// sqrt(u^2 + v^2 + 0.01^2) (capped by beta_powerlaw_umax when > 0), the law of
// whichbabc, then beta * f_ground and the beta_grounded_min floor where f_ground > 0.
// A negative or NaN beta raises *bad.  ub, vb: basal velocity with stride `st`
// between vertices (st = 1 for plain 2D arrays, 2 nz for the bed level of the
// interleaved device velocity).
__global__ __launch_bounds__(256) void k_calcbeta(GridDev g, cism_basal_physics bp, const double *__restrict__ ub,
                                                  const double *__restrict__ vb, long long st,
                                                  const double *__restrict__ bext,
                                                  const double *__restrict__ N, const double *__restrict__ fg,
                                                  double *__restrict__ beta, int *__restrict__ bad) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (long long)g.sx * g.sy) return;
    const double u = ub[v * st], w = vb[v * st];
    double speed = sqrt(u * u + w * w + 1.0e-2 * 1.0e-2);
    if (bp.beta_powerlaw_umax > 0.0) speed = fmin(speed, bp.beta_powerlaw_umax);
    double b = 0.0;
    switch (bp.whichbabc) {
    case 0: b = bp.ho_beta_const; break;
    case 4: b = bp.ho_beta_large; break;
    case 5: b = bext[v]; break;
    case 9: b = bp.powerlaw_c * pow(speed, 1.0 / bp.powerlaw_m - 1.0); break;
    case 11: {
        const double m = bp.powerlaw_m, n = N[v];
        const double num = bp.powerlaw_c * bp.coulomb_c * n;
        const double den = pow(pow(bp.powerlaw_c, m) * speed + pow(bp.coulomb_c * n, m), 1.0 / m);
        b = (num / den) * pow(speed, 1.0 / m - 1.0);
        if (b > 1.0e8) b = 1.0e8;
        break;
    }
    case 12: {
        const double tp = bp.powerlaw_c * pow(speed, 1.0 / bp.powerlaw_m);
        const double tc = bp.coulomb_c * N[v];
        b = tc <= tp ? tc / speed : tp / speed;
        break;
    }
    case 14: {
        const int j1 = (int)(v / g.sx) + 1;   // 1-based row, strip ns = 5 .. nsn-5
        b = (j1 >= 5 && j1 <= g.ny - 5) ? 100.0 : 1.0e4;
        break;
    }
    default: break;
    }
    b = b * fg[v];
    if (fg[v] > 0.0 && b < bp.beta_grounded_min) b = bp.beta_grounded_min;
    if (!(b >= 0.0)) bad[0] = 1;
    beta[v] = b;
}

// staggered_parallel_halo of one vertex field: a halo vertex takes its periodic image
// (always an owned vertex, so no entry is read after it is written) or 0 (outflow)
__global__ __launch_bounds__(256) void k_vertex_halo(GridDev g, double *__restrict__ a) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (long long)g.sx * g.sy) return;
    const int i = (int)(v % g.sx), j = (int)(v / g.sx);
    const int ii = halo_image(i, g.sx, g.periodic), jj = halo_image(j, g.sy, g.periodic);
    if (ii == i && jj == j) return;
    a[v] = (ii < 0 || jj < 0) ? 0.0 : a[ii + (long long)g.sx * jj];
}

static bool babc_supported(int w) {
    return w == 0 || w == 4 || w == 5 || w == 9 || w == 11 || w == 12 || w == 14;
}

static int launch_calcbeta(cism_ctx *c, const cism_basal_physics &bp, const double *ub, const double *vb,
                           long long st, const double *bext, const double *N, const double *fg, double *beta,
                           int *bad) {
    const GridDev &g = c->g;
    const long long ncol = (long long)g.sx * g.sy;
    hipLaunchKernelGGL(k_calcbeta, dim3(nblocks_for(ncol)), dim3(256), 0, c->stream, g, bp, ub, vb, st, bext, N,
                       fg, beta, bad);
    hipLaunchKernelGGL(k_vertex_halo, dim3(nblocks_for(ncol)), dim3(256), 0, c->stream, g, beta);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// the Picard loop's beta (c->beta) from the current basal velocity, the bed level of
// c->x; a bad value raises sc->beta_bad
int calcbeta_device(cism_ctx *c) {
    const GridDev &g = c->g;
    const double *ub = c->x + 2 * (long long)(g.nz - 1);
    return launch_calcbeta(c, c->bp, ub, ub + 1, 2LL * g.nz, c->bp_bext, c->bp_N, c->bp_fg, c->beta,
                           &c->sc->beta_bad);
}

int calcbeta_from_device(cism_ctx *c, const double *ub, const double *vb, long long stride) {
    return launch_calcbeta(c, c->bp, ub, vb, stride, c->bp_bext, c->bp_N, c->bp_fg, c->beta, &c->sc->beta_bad);
}

static int check_basal_fields(cism_ctx *c, const cism_basal_physics *bp, const double *beta_external,
                              const double *effecpress_stag, const double *f_ground) {
    const GridDev &g = c->g;
    const size_t ncol = (size_t)g.sx * g.sy;
    if (!babc_supported(bp->whichbabc)) {
        cism_set_error("calcbeta: which_ho_babc %d is not supported (0, 4, 5, 9, 11, 12, 14)", bp->whichbabc);
        return CISM_ERR_UNSUPPORTED;
    }
    if (!f_ground || (bp->whichbabc == 5 && !beta_external) ||
        ((bp->whichbabc == 11 || bp->whichbabc == 12) && !effecpress_stag) ||
        ((bp->whichbabc == 9 || bp->whichbabc == 11 || bp->whichbabc == 12) && !(bp->powerlaw_m > 0.0))) {
        cism_set_error("calcbeta: missing input field or powerlaw_m <= 0");
        return CISM_ERR_ARG;
    }
    if (bp->whichbabc == 5) {   // the reference's check that beta was read (maxval >= 0)
        double mx = -1.0e300;
        for (size_t t = 0; t < ncol; ++t) mx = fmax(mx, beta_external[t]);
        if (mx < 0.0) {
            cism_set_error("calcbeta: all external beta values are < 0");
            return CISM_ERR_ARG;
        }
    }
    return CISM_OK;
}

// f_ground, beta_external, N to the device (unused fields are zero-filled)
static int upload_basal_fields(cism_ctx *c, const double *beta_external, const double *effecpress_stag,
                               const double *f_ground, double *d_bext, double *d_N, double *d_fg) {
    const size_t ncol = (size_t)c->g.sx * c->g.sy;
    HIP_TRY(hipMemcpyAsync(d_fg, f_ground, sizeof(double) * ncol, hipMemcpyHostToDevice, c->stream));
    if (beta_external)
        HIP_TRY(hipMemcpyAsync(d_bext, beta_external, sizeof(double) * ncol, hipMemcpyHostToDevice, c->stream));
    else
        HIP_TRY(hipMemsetAsync(d_bext, 0, sizeof(double) * ncol, c->stream));
    if (effecpress_stag)
        HIP_TRY(hipMemcpyAsync(d_N, effecpress_stag, sizeof(double) * ncol, hipMemcpyHostToDevice, c->stream));
    else
        HIP_TRY(hipMemsetAsync(d_N, 0, sizeof(double) * ncol, c->stream));
    return CISM_OK;
}

extern "C" int cism_glissade_grounded_fraction(cism_ctx *c, const int32_t *ice_mask,
                                               const int32_t *floating_mask, const int32_t *land_mask,
                                               int whichground, double *f_ground) {
    if (!c || !ice_mask || !floating_mask || !land_mask || !f_ground) {
        cism_set_error("grounded_fraction: invalid arguments");
        return CISM_ERR_ARG;
    }
    if (whichground != 0 && whichground != 2) {
        cism_set_error("grounded_fraction: which_ho_ground %d not supported (0 or 2; the reference "
                       "stops on the GLP option 1)", whichground);
        return CISM_ERR_UNSUPPORTED;
    }
    HIP_TRY(hipSetDevice(c->device));
    const GridDev &g = c->g;
    const long long ncell = (long long)g.nx * g.ny, ncol = (long long)g.sx * g.sy;
    // masks staged in the node-mask buffers (nn >= ncell ints each) and the load vector
    int32_t *d_ice = c->umask, *d_fl = c->vmask, *d_land = c->umask + ncell;
    c->dirichlet_any = 1;   // the masks are scratch now: no longer known empty
    if (3 * ncell > 2 * g.nn) {
        cism_set_error("grounded_fraction: grid too small for the staging buffers");
        return CISM_ERR_ARG;
    }
    HIP_TRY(hipMemcpyAsync(d_ice, ice_mask, sizeof(int32_t) * ncell, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(d_fl, floating_mask, sizeof(int32_t) * ncell, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(d_land, land_mask, sizeof(int32_t) * ncell, hipMemcpyHostToDevice, c->stream));
    double *d_fg = c->vstage;
    hipLaunchKernelGGL(k_grounded_fraction, dim3(nblocks_for(ncol)), dim3(256), 0, c->stream, g, d_ice, d_fl,
                       d_land, whichground, d_fg);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(f_ground, d_fg, sizeof(double) * ncol, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CISM_OK;
}

extern "C" int cism_calcbeta(cism_ctx *c, const cism_basal_physics *bp, const double *ubas,
                             const double *vbas, const double *beta_external, const double *effecpress_stag,
                             const double *f_ground, double *beta) {
    if (!c || !bp || !ubas || !vbas || !beta) {
        cism_set_error("calcbeta: invalid arguments");
        return CISM_ERR_ARG;
    }
    HIP_TRY(hipSetDevice(c->device));
    const GridDev &g = c->g;
    const size_t ncol = (size_t)g.sx * g.sy;
    CISM_TRY(check_basal_fields(c, bp, beta_external, effecpress_stag, f_ground));
    // six vertex fields of temporary device memory; the context's own fields (the
    // Picard loop's beta and basal physics) are left as they are
    double *tmp = nullptr;
    HIP_TRY(hipMalloc(&tmp, sizeof(double) * 6 * ncol));
    double *d_u = tmp, *d_v = tmp + ncol, *d_bext = tmp + 2 * ncol, *d_N = tmp + 3 * ncol,
           *d_fg = tmp + 4 * ncol, *d_beta = tmp + 5 * ncol;
    int st = upload_basal_fields(c, beta_external, effecpress_stag, f_ground, d_bext, d_N, d_fg);
    if (st == CISM_OK) {
        hipError_t e = hipMemcpyAsync(d_u, ubas, sizeof(double) * ncol, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(d_v, vbas, sizeof(double) * ncol, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(&c->sc->beta_bad, 0, sizeof(int), c->stream);
        if (e != hipSuccess) {
            cism_set_error("calcbeta: %s", hipGetErrorString(e));
            st = CISM_ERR_HIP;
        }
    }
    if (st == CISM_OK) st = launch_calcbeta(c, *bp, d_u, d_v, 1, d_bext, d_N, d_fg, d_beta, &c->sc->beta_bad);
    if (st == CISM_OK && hipMemcpyAsync(beta, d_beta, sizeof(double) * ncol, hipMemcpyDeviceToHost, c->stream) != hipSuccess) {
        cism_set_error("calcbeta: copy of beta failed");
        st = CISM_ERR_HIP;
    }
    if (st == CISM_OK) st = sync_scalars(c);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (st != CISM_OK) return st;
    if (c->sc_host->beta_bad) {
        cism_set_error("Invalid beta value in calcbeta (negative or NaN)");
        return CISM_ERR_NAN;
    }
    return CISM_OK;
}

extern "C" int cism_set_basal_physics(cism_ctx *c, const cism_basal_physics *bp, const double *beta_external,
                                      const double *effecpress_stag, const double *f_ground) {
    if (!c) {
        cism_set_error("set_basal_physics: invalid arguments");
        return CISM_ERR_ARG;
    }
    HIP_TRY(hipSetDevice(c->device));
    if (!bp) {
        c->bp_on = 0;
        return CISM_OK;
    }
    CISM_TRY(check_basal_fields(c, bp, beta_external, effecpress_stag, f_ground));
    const size_t ncol = (size_t)c->g.sx * c->g.sy;
    if (!c->bp_fg) {
        HIP_TRY(hipMalloc(&c->bp_fg, sizeof(double) * ncol));
        HIP_TRY(hipMalloc(&c->bp_bext, sizeof(double) * ncol));
        HIP_TRY(hipMalloc(&c->bp_N, sizeof(double) * ncol));
    }
    CISM_TRY(upload_basal_fields(c, beta_external, effecpress_stag, f_ground, c->bp_bext, c->bp_N, c->bp_fg));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->bp = *bp;
    c->bp_on = 1;
    return CISM_OK;
}

// ---------------------------------------------------------------------------
// cism_glissade_velo_higher_solve's basal set-up from the resident geometry
// ---------------------------------------------------------------------------
// the mask bits of k_masks as int32 planes: ice, floating, ocean, land, and a zero
// calving-front plane (no subgrid calving front)
__global__ __launch_bounds__(256) void k_mask_planes(long long ncell, const int32_t *__restrict__ bits,
                                                     int32_t *__restrict__ planes) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    const int b = bits[t];
    planes[t] = b & 1;
    planes[ncell + t] = (b >> 1) & 1;
    planes[2 * ncell + t] = (b >> 2) & 1;
    planes[3 * ncell + t] = (b >> 3) & 1;
    planes[4 * ncell + t] = 0;
}

int mask_planes_device(cism_ctx *c) {
    const long long ncell = (long long)c->g.nx * c->g.ny;
    if (!c->planes) HIP_TRY(hipMalloc(&c->planes, sizeof(int32_t) * 5 * ncell));
    hipLaunchKernelGGL(k_mask_planes, dim3(nblocks_for(ncell)), dim3(256), 0, c->stream, ncell, c->umask,
                       c->planes);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

int basal_fields_upload(cism_ctx *c, const cism_basal_physics *bp, const double *beta_external,
                        const double *effecpress_stag) {
    const long long ncol = (long long)c->g.sx * c->g.sy;
    // the same checks as calcbeta's; f_ground comes from the device here
    double fg_dummy = 0.0;
    CISM_TRY(check_basal_fields(c, bp, beta_external, effecpress_stag, &fg_dummy));
    if (!c->bp_fg) {
        HIP_TRY(hipMalloc(&c->bp_fg, sizeof(double) * ncol));
        HIP_TRY(hipMalloc(&c->bp_bext, sizeof(double) * ncol));
        HIP_TRY(hipMalloc(&c->bp_N, sizeof(double) * ncol));
    }
    if (beta_external)
        HIP_TRY(hipMemcpyAsync(c->bp_bext, beta_external, sizeof(double) * ncol, hipMemcpyHostToDevice, c->stream));
    else
        HIP_TRY(hipMemsetAsync(c->bp_bext, 0, sizeof(double) * ncol, c->stream));
    if (effecpress_stag)
        HIP_TRY(hipMemcpyAsync(c->bp_N, effecpress_stag, sizeof(double) * ncol, hipMemcpyHostToDevice, c->stream));
    else
        HIP_TRY(hipMemsetAsync(c->bp_N, 0, sizeof(double) * ncol, c->stream));
    return CISM_OK;
}

// after basal_fields_upload: f_ground from the mask planes; bp becomes the Picard loop's
int basal_setup_device(cism_ctx *c, const int32_t *planes, int whichground, const cism_basal_physics *bp) {
    const GridDev &g = c->g;
    const long long ncell = (long long)g.nx * g.ny, ncol = (long long)g.sx * g.sy;
    if (whichground != 0 && whichground != 2) {
        cism_set_error("which_ho_ground %d not supported (0 or 2; the reference stops on the GLP option 1)",
                       whichground);
        return CISM_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL(k_grounded_fraction, dim3(nblocks_for(ncol)), dim3(256), 0, c->stream, g, planes,
                       planes + ncell, planes + 3 * ncell, whichground, c->bp_fg);
    HIP_TRY(hipGetLastError());
    c->bp = *bp;
    c->bp_on = 1;
    return CISM_OK;
}

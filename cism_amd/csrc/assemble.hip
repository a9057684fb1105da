// assemble.hip -- finite-element assembly of the Blatter-Pattyn blocks and the
// operations applied to the assembled system before the linear solve
// (glissade_velo_higher.F90).
//
// Assembly is split in two atomic-free kernels:
//  1. k_element: one workgroup = 16 hexahedral elements x 8 lanes.  Lane p
//     evaluates quadrature point p (Jacobian, inverse, basis derivatives,
//     Glen-law viscosity) into LDS; then lane a sums over the points the rows of
//     node a of the five point-weighted products Pxx = sum_p f x_a x_b, Pyy,
//     Pzz, Pxy = sum_p f x_a y_b and Pyx = sum_p f y_a x_b, reading the other
//     nodes' derivatives as broadcasts, and forms from them its row of Kuu and
//     Kvv (the upper part: b >= a) and of Kuv.  Kuu = 4 Pxx + Pyy + Pzz,
//     Kvv = Pxx + 4 Pyy + Pzz and Kuv = 2 Pxy + Pyx are the reference's sums
//     over points of f (4 x_a x_b + y_a y_b + z_a z_b) etc. regrouped, so the
//     entries agree with it to rounding.  Kvu = Kuv^T is not stored.  The 136
//     entries go element-fastest to a scratch array.
//  2. k_gather: one lane per node sums, slot by slot, the entries of the (up to
//     8) elements that share the node, in the reference's element loop order
//     (j, i, k ascending), so every coefficient is the reference's sum in the
//     reference's order; the local basal friction term is added last.
#include "internal.h"

#include <math.h>

// node n of a hexahedron: horizontal offsets (ix, jy) in {0,1} from the SW
// vertex, layer kl = 0 upper (level k) / 1 lower (level k+1)   (velo_higher:453-486):
// ix = {0,1,1,0,0,1,1,0}, jy = {0,0,1,1,0,0,1,1}, kl = {1,1,1,1,0,0,0,0}.  Written as
// functions of n so that in unrolled loops the basis-function signs fold at compile time.
__host__ __device__ constexpr int c_node_ix(int n) { return ((n + 1) >> 1) & 1; }
__host__ __device__ constexpr int c_node_jy(int n) { return (n >> 1) & 1; }
__host__ __device__ constexpr int c_node_kl(int n) { return n < 4 ? 1 : 0; }
static_assert(c_node_ix(0) == 0 && c_node_ix(1) == 1 && c_node_ix(2) == 1 && c_node_ix(3) == 0 &&
              c_node_ix(4) == 0 && c_node_ix(5) == 1 && c_node_ix(6) == 1 && c_node_ix(7) == 0, "ix");
static_assert(c_node_jy(0) == 0 && c_node_jy(1) == 0 && c_node_jy(2) == 1 && c_node_jy(3) == 1 &&
              c_node_jy(4) == 0 && c_node_jy(5) == 0 && c_node_jy(6) == 1 && c_node_jy(7) == 1, "jy");

__host__ __device__ constexpr int node_id(int ix, int jy, int kl) {
    return kl == 1 ? (jy == 0 ? (ix == 0 ? 0 : 1) : (ix == 0 ? 3 : 2))
                   : (jy == 0 ? (ix == 0 ? 4 : 5) : (ix == 0 ? 7 : 6));
}
// index of (a, b), a <= b, in the packed upper triangle of an 8x8 matrix
__host__ __device__ constexpr int tri_index(int a, int b) {
    return a <= b ? (a * (15 - a)) / 2 + b : (b * (15 - b)) / 2 + a;
}

#define EL_PER_WG 32     // elements per workgroup of the stress kernel
#define EL_ELEM 16       // elements per workgroup of k_element (8 lanes each): measured
                         // best of 8, 16, 24 and 32 at the 1200 x 1400 x 11 size

// element scratch: tiles of 64 elements, entry-major inside a tile, so a wave's
// reads and writes of one entry are one contiguous 512-byte run.  k_gather writes the
// matrix with non-temporal stores (whole 16-byte pairs, 1 KiB per wave instruction): the
// matrix is not read again before the pass has run, and the gather's scratch reads keep
// the L2 (4.83 -> 4.26 ms at 1200 x 1400 x 11).  The element kernel's scratch stores, one
// double per lane at a 512-byte stride, stay ordinary: as non-temporal stores they took
// 6.20 ms against 5.66.
__host__ __device__ inline long long keidx(int ent, long long e) {
    return (((e >> 6) * CISM_KE_ENTRIES + ent) << 6) + (e & 63);
}

// LDS record of quadrature point p of an element: the basis derivatives x[8], y[8],
// z[8] of its nodes (24 doubles, so each record and each node pair in it is 16-byte
// aligned); the weight f = efvs det / vol0 and the point's efvs sit in arrays of their own.
// An element's eight records are padded to 194 doubles: the eight elements a wave reads
// side by side then start in eight disjoint groups of four banks.  The nodes' (x, y, z, u,
// v) share the records' memory (a barrier separates their last read from the first
// record store), which keeps the workgroup under 26.7 KB: six workgroups (three waves per
// SIMD) fit a CU.
#define QP_REC 24
#define EL_REC (8 * QP_REC + 2)
struct QPLds {
    union {
        alignas(16) double rec[EL_ELEM][EL_REC];
        double nodes[EL_ELEM][8 * 5 + 1];                                              // x y z u v
    };
    double f[EL_ELEM][8];
    double ev[EL_ELEM][8];
};
static_assert(sizeof(QPLds) + EL_ELEM * sizeof(int) <= 163840 / 6, "six element workgroups per CU");

// One quadrature point p of a hexahedral element whose 8 nodes are given as
// (x, y, z, u, v) records: basis-function derivatives (get_basis_function_
// derivatives_3d, velo_higher:6360-6579), the velocity gradient (ux uy uz vx vy vz)
// and the Glen-law viscosity (compute_effective_viscosity, velo_higher:7199-7378).
// A zero Jacobian determinant gives zero derivatives and sets `singular`.
__device__ __forceinline__ void eval_qp(int p, const double *__restrict__ nodes, int whichapprox,
                                        int whichefvs, double efvs_constant, double ff,
                                        double *dxo, double *dyo, double *dzo, double &det,
                                        double &ev, double (&strain)[6], bool &singular) {
    const double r = 1.0 / sqrt(3.0);
    const double xq = (p == 0 || p == 3 || p == 4 || p == 7) ? -r : r;
    const double yq = (p < 2 || p == 4 || p == 5) ? -r : r;
    const double zq = p < 4 ? -r : r;
    double a[8], b[8], cz[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const double sx = c_node_ix(n) ? 1.0 : -1.0, sy = c_node_jy(n) ? 1.0 : -1.0;
        const double sz = c_node_kl(n) ? -1.0 : 1.0;
        const double ax = sx > 0 ? 1.0 + xq : 1.0 - xq;
        const double ay = sy > 0 ? 1.0 + yq : 1.0 - yq;
        const double az = sz > 0 ? 1.0 + zq : 1.0 - zq;
        a[n] = sx * (ay * az / 8.0);
        b[n] = sy * (ax * az / 8.0);
        cz[n] = sz * (ax * ay / 8.0);
    }
    // sums of products with fused multiply-adds (one rounding per term)
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const double X = nodes[n * 5 + 0], Y = nodes[n * 5 + 1], Z = nodes[n * 5 + 2];
        J[0][0] = fma(a[n], X, J[0][0]);  J[0][1] = fma(a[n], Y, J[0][1]);  J[0][2] = fma(a[n], Z, J[0][2]);
        J[1][0] = fma(b[n], X, J[1][0]);  J[1][1] = fma(b[n], Y, J[1][1]);  J[1][2] = fma(b[n], Z, J[1][2]);
        J[2][0] = fma(cz[n], X, J[2][0]); J[2][1] = fma(cz[n], Y, J[2][1]); J[2][2] = fma(cz[n], Z, J[2][2]);
    }
    double cof[3][3];
    cof[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    cof[0][1] = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]);
    cof[0][2] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    cof[1][0] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]);
    cof[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
    cof[1][2] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]);
    cof[2][0] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    cof[2][1] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]);
    cof[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    det = J[0][0] * cof[0][0] + J[0][1] * cof[0][1] + J[0][2] * cof[0][2];
    double Ji[3][3];
    singular = !(fabs(det) > 0.0);
    if (!singular) {
        const double rdet = 1.0 / det;
#pragma unroll
        for (int cc = 0; cc < 3; ++cc)
#pragma unroll
            for (int rr = 0; rr < 3; ++rr) Ji[rr][cc] = cof[cc][rr] * rdet;
    } else {
#pragma unroll
        for (int cc = 0; cc < 3; ++cc)
#pragma unroll
            for (int rr = 0; rr < 3; ++rr) Ji[rr][cc] = 0.0;
    }
    double ux = 0, uy = 0, uz = 0, vx = 0, vy = 0, vz = 0;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const double ddx = fma(Ji[0][2], cz[n], fma(Ji[0][1], b[n], Ji[0][0] * a[n]));
        const double ddy = fma(Ji[1][2], cz[n], fma(Ji[1][1], b[n], Ji[1][0] * a[n]));
        const double ddz = fma(Ji[2][2], cz[n], fma(Ji[2][1], b[n], Ji[2][0] * a[n]));
        if (dxo) {
            dxo[n] = ddx;
            dyo[n] = ddy;
            dzo[n] = ddz;
        }
        const double u = nodes[n * 5 + 3], v = nodes[n * 5 + 4];
        ux = fma(ddx, u, ux); uy = fma(ddy, u, uy); uz = fma(ddz, u, uz);
        vx = fma(ddx, v, vx); vy = fma(ddy, v, vy); vz = fma(ddz, v, vz);
    }
    strain[0] = ux; strain[1] = uy; strain[2] = uz; strain[3] = vx; strain[4] = vy; strain[5] = vz;
    if (whichefvs == 0) {
        ev = efvs_constant;
    } else if (whichefvs == 1) {
        const double effstrain = (500.0 / CISM_SCYR) / 200.0e3 * CISM_SCYR;
        ev = ff * pow(effstrain, (1.0 - 3.0) / 3.0);
    } else {
        const double emin = 1.0e-8;
        double sq;
        if (whichapprox == 0)
            sq = emin * emin + 0.25 * (uz * uz + vz * vz);
        else if (whichapprox == 1)
            sq = emin * emin + (ux * ux + vy * vy + ux * vy + 0.25 * ((vx + uy) * (vx + uy)));
        else
            sq = emin * emin + (ux * ux + vy * vy + ux * vy + 0.25 * ((vx + uy) * (vx + uy)))
               + 0.25 * (uz * uz + vz * vz);
        ev = ff * pow(sq, ((1.0 - 3.0) / 3.0) / 2);
    }
}

template <int APPROX>
__global__ __launch_bounds__(8 * EL_ELEM) void k_element(GridDev g, const double *__restrict__ sigma,
                                                 const int32_t *__restrict__ ac,
                                                 const double *__restrict__ xV,
                                                 const double *__restrict__ yV,
                                                 const double *__restrict__ uvel,
                                                 const double *__restrict__ susrf,
                                                 const double *__restrict__ sthck,
                                                 const double *__restrict__ flwafact,
                                                 int whichefvs, double efvs_constant,
                                                 double *__restrict__ Ke, double *__restrict__ efvs,
                                                 DevScal *__restrict__ sc) {
    __shared__ QPLds L;
    __shared__ int valid[EL_ELEM];
    const int el = threadIdx.x >> 3, sub = threadIdx.x & 7;
    // element and node indices fit in 32 bits (the context rejects larger grids)
    const int e = blockIdx.x * EL_ELEM + el;
    const int nzm = g.nz - 1;
    int k = 0, ic = 0, jc = 0;
    bool ok = e < (int)g.ne;
    if (ok) {
        const int cell = e / nzm;
        k = e - cell * nzm + 1;                 // 1-based layer
        jc = cell / g.nx;
        ic = cell - jc * g.nx + 1;              // 1-based cell
        jc += 1;
        ok = ic >= CISM_NHALO + 1 && ic <= g.nx - CISM_NHALO + 1 && jc >= CISM_NHALO + 1 &&
             jc <= g.ny - CISM_NHALO + 1 && ac[cell] != 0;
    }
    if (sub == 0) valid[el] = ok;
    // a group with no active element only clears its efvs
    if (!__syncthreads_or(ok)) {
        if (threadIdx.x < EL_ELEM && blockIdx.x * EL_ELEM + (int)threadIdx.x < (int)g.ne)
            efvs[blockIdx.x * EL_ELEM + threadIdx.x] = 0.0;
        return;
    }
    // phase 0: lane (el, n) gathers node n (velo_higher:5100-5127)
    if (ok) {
        const int n = sub;
        const int in = ic - 1 + c_node_ix(n), jn = jc - 1 + c_node_jy(n), kn = k + c_node_kl(n);
        const int v = (in - 1) + g.sx * (jn - 1);
        const int qn = (kn - 1) + g.nz * v;
        L.nodes[el][n * 5 + 0] = xV[v];
        L.nodes[el][n * 5 + 1] = yV[v];
        L.nodes[el][n * 5 + 2] = susrf[v] - sigma[kn - 1] * sthck[v];
        L.nodes[el][n * 5 + 3] = uvel[2 * qn];
        L.nodes[el][n * 5 + 4] = uvel[2 * qn + 1];
    }
    __syncthreads();
    // phase 1: lane (el, p) evaluates quadrature point p (velo_higher:6360-6579, 7199-7378)
    double dx[8], dy[8], dz[8], evp = 0.0, fw = 0.0;
    if (ok) {
        const int p = sub;
        double det, ev, st[6];
        bool singular;
        eval_qp(p, L.nodes[el], APPROX, whichefvs, efvs_constant, flwafact[e], dx, dy, dz, det, ev, st,
                singular);
        if (singular) atomicOr(&sc->singular_flag, 1);
        evp = ev;
        fw = ev * 1.0 * det / CISM_VOL0;
    }
    __syncthreads();   // every lane has read the nodes: the records may overwrite them
    if (ok) {
        const int p = sub;
        double *r = &L.rec[el][p * QP_REC];
#pragma unroll
        for (int n = 0; n < 8; n += 2) {
            *reinterpret_cast<double2 *>(r + n) = make_double2(dx[n], dx[n + 1]);
            *reinterpret_cast<double2 *>(r + 8 + n) = make_double2(dy[n], dy[n + 1]);
            *reinterpret_cast<double2 *>(r + 16 + n) = make_double2(dz[n], dz[n + 1]);
        }
        L.f[el][p] = fw;
        L.ev[el][p] = evp;
    }
    __syncthreads();
    // element efvs = mean over quadrature points (velo_higher:5167-5172)
    if (threadIdx.x < EL_ELEM) {
        const int ee = blockIdx.x * EL_ELEM + threadIdx.x;
        if (ee < (int)g.ne) {
            double s = 0.0;
            if (valid[threadIdx.x])
                for (int p = 0; p < 8; ++p) s = s + L.ev[threadIdx.x][p];
            efvs[ee] = valid[threadIdx.x] ? s / 8 : 0.0;
        }
    }
    if (!ok) return;
    // phase 2: lane (el, a) sums row a of the point-weighted products over the
    // points, in point order (compute_element_matrix, velo_higher:7960-7982)
    const int a = sub;
    double Pxx[8], Pyy[8], Pzz[8], Pxy[8], Pyx[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) Pxx[b] = Pyy[b] = Pzz[b] = Pxy[b] = Pyx[b] = 0.0;
#pragma unroll 1
    for (int p = 0; p < 8; ++p) {
        const double *r = &L.rec[el][p * QP_REC];
        const double f = L.f[el][p];
        const double fx = f * r[a], fy = f * r[8 + a], fz = f * r[16 + a];
#pragma unroll
        for (int b = 0; b < 8; b += 2) {
            if (APPROX != 0) {
                const double2 X = *reinterpret_cast<const double2 *>(r + b);
                const double2 Y = *reinterpret_cast<const double2 *>(r + 8 + b);
                Pxx[b] = fma(fx, X.x, Pxx[b]);  Pxx[b + 1] = fma(fx, X.y, Pxx[b + 1]);
                Pyy[b] = fma(fy, Y.x, Pyy[b]);  Pyy[b + 1] = fma(fy, Y.y, Pyy[b + 1]);
                Pxy[b] = fma(fx, Y.x, Pxy[b]);  Pxy[b + 1] = fma(fx, Y.y, Pxy[b + 1]);
                Pyx[b] = fma(fy, X.x, Pyx[b]);  Pyx[b + 1] = fma(fy, X.y, Pyx[b + 1]);
            }
            if (APPROX != 1) {
                const double2 Z = *reinterpret_cast<const double2 *>(r + 16 + b);
                Pzz[b] = fma(fz, Z.x, Pzz[b]);  Pzz[b + 1] = fma(fz, Z.y, Pzz[b + 1]);
            }
        }
    }
    double *kcol = Ke + keidx(0, e);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        if (b >= a) {
            double ku, kv;
            if (APPROX == 0) {
                ku = kv = Pzz[b];
            } else if (APPROX == 1) {
                ku = 4.0 * Pxx[b] + Pyy[b];
                kv = Pxx[b] + 4.0 * Pyy[b];
            } else {
                ku = 4.0 * Pxx[b] + Pyy[b] + Pzz[b];
                kv = Pxx[b] + 4.0 * Pyy[b] + Pzz[b];
            }
            const int t = tri_index(a, b);
            kcol[t << 6] = ku;
            kcol[(36 + t) << 6] = kv;
        }
        kcol[(72 + a + 8 * b) << 6] = APPROX == 0 ? 0.0 : 2.0 * Pxy[b] + Pyx[b];
    }
}

// ---------------------------------------------------------------------------
// compute_internal_stress (velo_higher:6968-7195): the element mean over its 8
// quadrature points of 2 efvs times the strain-rate components, and the
// effective stress.  Same element ranges and quadrature as the assembly; SIA
// uses only the vertical shear, SSA only the horizontal gradients.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_stress(GridDev g, const double *__restrict__ sigma,
                                                const int32_t *__restrict__ ac,
                                                const double *__restrict__ xV,
                                                const double *__restrict__ yV,
                                                const double *__restrict__ uvel,
                                                const double *__restrict__ susrf,
                                                const double *__restrict__ sthck,
                                                const double *__restrict__ flwafact, int whichapprox,
                                                int whichefvs, double efvs_constant,
                                                double *__restrict__ tau, DevScal *__restrict__ sc) {
    __shared__ double nodes[EL_PER_WG][8 * 5 + 1];
    __shared__ double prod[EL_PER_WG][8][5];
    const int el = threadIdx.x >> 3, sub = threadIdx.x & 7;
    const long long e = (long long)blockIdx.x * EL_PER_WG + el;
    const int nzm = g.nz - 1;
    int k = 0, ic = 0, jc = 0;
    bool ok = e < g.ne;
    if (ok) {
        k = (int)(e % nzm) + 1;
        const long long cell = e / nzm;
        ic = (int)(cell % g.nx) + 1;
        jc = (int)(cell / g.nx) + 1;
        ok = ic >= CISM_NHALO + 1 && ic <= g.nx - CISM_NHALO + 1 && jc >= CISM_NHALO + 1 &&
             jc <= g.ny - CISM_NHALO + 1 && ac[(ic - 1) + (long long)g.nx * (jc - 1)] != 0;
    }
    if (ok) {
        const int n = sub;
        const int in = ic - 1 + c_node_ix(n), jn = jc - 1 + c_node_jy(n), kn = k + c_node_kl(n);
        const long long v = (in - 1) + (long long)g.sx * (jn - 1);
        const long long qn = (kn - 1) + (long long)g.nz * v;
        nodes[el][n * 5 + 0] = xV[v];
        nodes[el][n * 5 + 1] = yV[v];
        nodes[el][n * 5 + 2] = susrf[v] - sigma[kn - 1] * sthck[v];
        nodes[el][n * 5 + 3] = uvel[2 * qn];
        nodes[el][n * 5 + 4] = uvel[2 * qn + 1];
    }
    __syncthreads();
    if (ok) {
        const int p = sub;
        double det, ev, st[6];
        bool singular;
        eval_qp(p, nodes[el], whichapprox, whichefvs, efvs_constant, flwafact[e], nullptr, nullptr,
                nullptr, det, ev, st, singular);
        if (singular) atomicOr(&sc->singular_flag, 1);
        double ux = st[0], uy = st[1], uz = st[2], vx = st[3], vy = st[4], vz = st[5];
        if (whichapprox == 0) ux = uy = vx = vy = 0.0;
        else if (whichapprox == 1) uz = vz = 0.0;
        prod[el][p][0] = ev * uz;
        prod[el][p][1] = ev * vz;
        prod[el][p][2] = 2.0 * ev * ux;
        prod[el][p][3] = 2.0 * ev * vy;
        prod[el][p][4] = ev * (vx + uy);
    }
    __syncthreads();
    if (sub == 0 && e < g.ne) {
        double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (ok) {
            for (int p = 0; p < 8; ++p)
                for (int c = 0; c < 5; ++c) t[c] = t[c] + prod[el][p][c];
            for (int c = 0; c < 5; ++c) t[c] = t[c] / 8;
        }
        const double xz = t[0], yz = t[1], xx = t[2], yy = t[3], xy = t[4];
        const long long ne = g.ne;
        tau[e] = xz;
        tau[ne + e] = yz;
        tau[2 * ne + e] = xx;
        tau[3 * ne + e] = yy;
        tau[4 * ne + e] = xy;
        tau[5 * ne + e] = ok ? sqrt(xx * xx + yy * yy + xx * yy + xy * xy + xz * xz + yz * yz) : 0.0;
    }
}

int internal_stress_device(cism_ctx *c, int whichapprox, int whichefvs, double efvs_constant, double *tau) {
    const GridDev &g = c->g;
    const int nb = (int)((g.ne + EL_PER_WG - 1) / EL_PER_WG);
    hipLaunchKernelGGL(k_stress, dim3(nb), dim3(256), 0, c->stream, g, c->sigma, c->ac, c->xV, c->yV, c->x,
                       c->susrf, c->sthck, c->flwafact, whichapprox, whichefvs, efvs_constant, tau, c->sc);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// cell (1-based i) -> the owned cell whose value a periodic halo cell takes
// (parallel_halo, parallel_slap.F90:1794-1860), or -1 for an outflow halo cell
__device__ inline int cell_image(int i, int n, int periodic) {
    if (i > CISM_NHALO && i <= n - CISM_NHALO) return i;
    if (!periodic) return -1;
    return i <= CISM_NHALO ? i + (n - 2 * CISM_NHALO) : i - (n - 2 * CISM_NHALO);
}

// compute_basal_friction_heatflx (velo_higher:6755-6964) on the basal layer of the
// velocity: one lane per cell, halo cells by their periodic image
// xb: basal velocity of vertex column col at xb[cs*col] (u) and xb[cs*col + 1] (v)
__global__ __launch_bounds__(256) void k_bfric(GridDev g, const int32_t *__restrict__ ac,
                                               const int32_t *__restrict__ av,
                                               const double *__restrict__ xb, long long cs,
                                               const double *__restrict__ beta, int standard,
                                               double *__restrict__ out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.nx * g.ny) return;
    const int i0 = (int)(t % g.nx) + 1, j0 = (int)(t / g.nx) + 1;
    const int i = cell_image(i0, g.nx, g.periodic), j = cell_image(j0, g.ny, g.periodic);
    double val = 0.0;
    if (i > 0 && j > 0) {
        // basal velocity and beta at vertex (iv, jv), 1-based
        auto ub = [&](int iv, int jv) { return xb[cs * ((iv - 1) + (long long)g.sx * (jv - 1))]; };
        auto vb = [&](int iv, int jv) { return xb[cs * ((iv - 1) + (long long)g.sx * (jv - 1)) + 1]; };
        auto bt = [&](int iv, int jv) { return beta[(iv - 1) + (long long)g.sx * (jv - 1)]; };
        if (standard) {
            if (ac[(i - 1) + (long long)g.nx * (j - 1)]) {
                const double r = 1.0 / sqrt(3.0);
                const double xq[4] = {-r, r, r, -r}, yq[4] = {-r, -r, r, r};
                const int di[4] = {-1, 0, 0, -1}, dj[4] = {-1, -1, 0, 0};
                double u[4], v[4], b[4];
                for (int n = 0; n < 4; ++n) {
                    u[n] = ub(i + di[n], j + dj[n]);
                    v[n] = vb(i + di[n], j + dj[n]);
                    b[n] = bt(i + di[n], j + dj[n]);
                }
                double acc = 0.0, sw = 0.0;
                for (int p = 0; p < 4; ++p) {
                    double phi[4];
                    phi[0] = (1.0 - xq[p]) * (1.0 - yq[p]) / 4.0;
                    phi[1] = (1.0 + xq[p]) * (1.0 - yq[p]) / 4.0;
                    phi[2] = (1.0 + xq[p]) * (1.0 + yq[p]) / 4.0;
                    phi[3] = (1.0 - xq[p]) * (1.0 + yq[p]) / 4.0;
                    double uq = 0.0, vq = 0.0, bq = 0.0;
                    for (int n = 0; n < 4; ++n) {
                        uq = uq + phi[n] * u[n];
                        vq = vq + phi[n] * v[n];
                        bq = bq + phi[n] * b[n];
                    }
                    acc = acc + 1.0 * bq * (uq * uq + vq * vq);
                    sw = sw + 1.0;
                }
                val = acc / (sw * CISM_SCYR);
            }
        } else {
            auto stv = [&](int iv, int jv) {
                if (!av[(iv - 1) + (long long)g.sx * (jv - 1)]) return 0.0;
                const double q = bt(iv, jv) * (ub(iv, jv) * ub(iv, jv) + vb(iv, jv) * vb(iv, jv));
                return q / CISM_SCYR;
            };
            val = (stv(i, j) + stv(i - 1, j) + stv(i, j - 1) + stv(i - 1, j - 1)) / 4.0;
        }
    }
    out[t] = val;
}

// basal traction (velo_higher:3893-3894): beta times the basal velocity, every vertex
__global__ __launch_bounds__(256) void k_btract(GridDev g, const double *__restrict__ x,
                                                const double *__restrict__ beta,
                                                double *__restrict__ out) {
    const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ncol = (long long)g.sx * g.sy;
    if (col >= ncol) return;
    const long long q = (g.nz - 1) + (long long)g.nz * col;
    out[col] = beta[col] * x[2 * q];
    out[ncol + col] = beta[col] * x[2 * q + 1];
}

int basal_diagnostics_device(cism_ctx *c, int whichassemble_bfric, double *bfric, double *btract,
                             const double *xb, long long cs) {
    const GridDev &g = c->g;
    if (!xb) {
        xb = c->x + 2 * (g.nz - 1);
        cs = 2LL * g.nz;
    }
    if (bfric) {
        hipLaunchKernelGGL(k_bfric, dim3(nblocks_for((long long)g.nx * g.ny)), dim3(256), 0, c->stream, g,
                           c->ac, c->av, xb, cs, c->beta, whichassemble_bfric == 0, bfric);
        HIP_TRY(hipGetLastError());
    }
    if (btract) {
        hipLaunchKernelGGL(k_btract, dim3(nblocks_for((long long)g.sx * g.sy)), dim3(256), 0, c->stream, g,
                           c->x, c->beta, btract);
        HIP_TRY(hipGetLastError());
    }
    return CISM_OK;
}

// The element scratch ends with one tile of zeros (ensure_ke): the gather reads its first
// column in place of an element that is not there, so every one of its loads is
// unconditional, and adding +0.0 leaves every sum unchanged.  Taking it from Ke itself
// keeps every load of the gather on the restrict-qualified scratch, so the compiler may
// move the next slot's loads above this slot's stores to A and As.
__host__ __device__ inline long long ke_zero_offset(long long nepad) { return keidx(0, nepad); }

// entry ent of the element whose scratch column starts at base (base = Ke + keidx(0, e))
__device__ __forceinline__ double ke_at(const double *__restrict__ base, int ent) {
    return base[ent << 6];
}

// value of K_block(nr, nc) of an element from its scratch column
__device__ __forceinline__ double ke_val(const double *__restrict__ base, int blk, int nr, int nc) {
    if (blk == 0) return ke_at(base, tri_index(nr, nc));
    if (blk == 3) return ke_at(base, 36 + tri_index(nr, nc));
    if (blk == 1) return ke_at(base, 72 + nr + 8 * nc);   // Kuv(nr, nc)
    return ke_at(base, 72 + nc + 8 * nr);                 // Kvu(nr, nc) = Kuv(nc, nr)
}

// One lane per node (k, i, j): A(m, node) = sum over elements in (j, i, k) order
// (element_to_global_matrix_3d, velo_higher:7990-8054), then the local basal
// friction term (basal_sliding_bc LOCAL, velo_higher:8216-8232).  A node with no
// active element around it gets zero rows.  An owned node with an inactive vertex
// and no active cell around it, now and at the previous assembly (av_zero, ac_zero,
// when given), is not rewritten: the previous assembly left its rows zero, and
// since then the Dirichlet update and the basal term touch active vertices only and
// the halo update non-owned nodes.  Node and element indices fit in 32 bits (the
// context rejects larger grids).
__global__ __launch_bounds__(256) void k_gather(GridDev g, const int32_t *__restrict__ ac,
                                                const int32_t *__restrict__ av,
                                                const int32_t *__restrict__ av_zero,
                                                const int32_t *__restrict__ ac_zero,
                                                const double *__restrict__ Ke,
                                                const double *__restrict__ beta, int add_basal,
                                                double basal_scale, double *__restrict__ A, int store64,
                                                float *__restrict__ As, const int32_t *__restrict__ as_av,
                                                const int32_t *__restrict__ as_ac,
                                                const double *__restrict__ x, const double *__restrict__ bl,
                                                double *__restrict__ tv, double *__restrict__ rsv,
                                                double *__restrict__ partials) {
    // As (the mixed-precision path without Dirichlet rows or a separate basal pass): the
    // lane also writes its node's FP32 symmetric storage (slots CISM_DIAG .. 26) and, as
    // k_resvec<true> with b = bl (the load vector, no Dirichlet row changes it), the
    // residual of its row from the coefficients it has just formed: tv = A x - b on owned
    // active rows (0 elsewhere), rsv = b - A x on owned rows (the halo rows follow from their
    // images by a vector halo update); partials per block: |tv|^2 and |b|^2 over owned
    // active rows, |rsv|^2 and |b|^2 over owned rows, the nonzeros of owned active rows
    // (every lane reaches the reduction).  x must carry its halo.  store64 = 0: the FP64
    // matrix is not written (the mixed path that confirms no linear solve in FP64 reads it
    // at most once per Picard iteration: its residual comes from a pass of this kernel with
    // only rsv, from the element matrices).
    double red[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const bool res = rsv != nullptr;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < (int)g.nn) {
    const int k = q % g.nz + 1;                // 1-based
    const int col = q / g.nz;
    const int i = col % g.sx + 1, j = col / g.sx + 1;
    const int nzm = g.nz - 1;
    // candidate elements: cell (i+dI, j+dJ), layer k-1+dK
    const double *base[2][2][2];
    bool any = false, was = false;
#pragma unroll
    for (int dJ = 0; dJ < 2; ++dJ)
#pragma unroll
        for (int dI = 0; dI < 2; ++dI) {
            const int icl = i + dI, jcl = j + dJ;
            const bool okc = icl >= CISM_NHALO + 1 && icl <= g.nx - CISM_NHALO + 1 &&
                             jcl >= CISM_NHALO + 1 && jcl <= g.ny - CISM_NHALO + 1;
            const int cell = (icl - 1) + g.nx * (jcl - 1);
            const bool act = okc && ac[cell] != 0;
            if (ac_zero) was = was || (okc && ac_zero[cell] != 0);
#pragma unroll
            for (int dK = 0; dK < 2; ++dK) {
                const int ke = k - 1 + dK;
                const bool ok = act && ke >= 1 && ke <= nzm;
                any = any || ok;
                base[dJ][dI][dK] = Ke + (ok ? keidx(0, (ke - 1) + (long long)nzm * cell) : ke_zero_offset(g.nepad));
            }
        }
    const bool basal = add_basal && k == g.nz && av[col];
    if (!any && !basal) {
        const bool owned = is_owned(g, i - 1, j - 1);
        if (store64 && !(av_zero && !was && !av[col] && !av_zero[col] && owned)) {
            const double2 z = make_double2(0.0, 0.0);
#pragma unroll
            for (int m = 0; m < 27; ++m) {
                *reinterpret_cast<double2 *>(A + aidx(0, m, q)) = z;
                *reinterpret_cast<double2 *>(A + aidx(2, m, q)) = z;
            }
        }
        // the FP32 row is zero already when the node had no active cell at its last write
        if (As && !(as_av && !av[col] && !as_av[col] && owned && !touches_active_cell(g, as_ac, i - 1, j - 1))) {
#pragma unroll
            for (int s = 0; s < CISM_SYM_SLOTS; ++s)
                *reinterpret_cast<float4 *>(As + sidx(0, s, q)) = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        if (res) {   // no active row here (an owned active vertex touches an active cell)
            if (tv) reinterpret_cast<double2 *>(tv)[q] = make_double2(0.0, 0.0);
            if (owned) {
                const double2 bb = reinterpret_cast<const double2 *>(bl)[q];
                reinterpret_cast<double2 *>(rsv)[q] = bb;
                red[2] = bb.x * bb.x + bb.y * bb.y;
                red[3] = bb.x * bb.x + bb.y * bb.y;
            }
        }
    } else {
    const bool owned = is_owned(g, i - 1, j - 1);
    const bool count = res && owned && av[col];
    double nnz = 0.0, su = 0.0, sv = 0.0;
    // slot by slot, the four blocks together: each slot's coefficients leave as two
    // 16-byte stores (the tile keeps a node's four blocks of one slot adjacent)
#pragma unroll
    for (int m = 0; m < 27; ++m) {
        const int iA = m % 3 - 1, jA = (m / 3) % 3 - 1, kA = m / 9 - 1;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int blk = 0; blk < 4; ++blk) {
#pragma unroll
            for (int dJ = 0; dJ < 2; ++dJ)
#pragma unroll
                for (int dI = 0; dI < 2; ++dI)
#pragma unroll
                    for (int dK = 0; dK < 2; ++dK) {
                        const int ix = 1 - dI, jy = 1 - dJ, kl = 1 - dK;
                        const int cx = ix + iA, cy = jy + jA, cl = kl + kA;
                        if (cx < 0 || cx > 1 || cy < 0 || cy > 1 || cl < 0 || cl > 1) continue;
                        acc[blk] = acc[blk] + ke_val(base[dJ][dI][dK], blk, node_id(ix, jy, kl),
                                                     node_id(cx, cy, cl));
                    }
            if (basal && m == CISM_DIAG && (blk == 0 || blk == 3))
                acc[blk] = acc[blk] + basal_scale * beta[col];
        }
        if (store64) store_blocks_nt(A, q, m, acc[0], acc[1], acc[2], acc[3]);
        if (As && m >= CISM_DIAG) {
            typedef float fvec4_nt __attribute__((ext_vector_type(4)));
            const fvec4_nt v = {(float)acc[0], (float)acc[1], (float)acc[2], (float)acc[3]};
            __builtin_nontemporal_store(v, reinterpret_cast<fvec4_nt *>(As + sidx(0, m - CISM_DIAG, q)));
        }
        if (count) {
            nnz += (double)((acc[0] != 0.0) + (acc[1] != 0.0) + (acc[2] != 0.0) + (acc[3] != 0.0));
            // the row product in k_resvec's order (slots in indxA order, the same operations)
            if (k - 1 + kA >= 0 && k - 1 + kA < g.nz) {
                const long long qn = q + kA + (long long)g.nz * (iA + (long long)g.sx * jA);
                const double2 xx = reinterpret_cast<const double2 *>(x)[qn];
                su = su + acc[0] * xx.x + acc[1] * xx.y;
                sv = sv + acc[2] * xx.x + acc[3] * xx.y;
            }
        }
    }
    if (res) {
        const double2 bb = reinterpret_cast<const double2 *>(bl)[q];
        if (count) {
            const double ru = su - bb.x, rv = sv - bb.y;
            if (tv) reinterpret_cast<double2 *>(tv)[q] = make_double2(ru, rv);
            reinterpret_cast<double2 *>(rsv)[q] = make_double2(-ru, -rv);
            red[0] = ru * ru + rv * rv;
            red[1] = bb.x * bb.x + bb.y * bb.y;
            red[2] = ru * ru + rv * rv;
            red[3] = bb.x * bb.x + bb.y * bb.y;
        } else {
            if (tv) reinterpret_cast<double2 *>(tv)[q] = make_double2(0.0, 0.0);
            if (owned) {
                reinterpret_cast<double2 *>(rsv)[q] = bb;
                red[2] = bb.x * bb.x + bb.y * bb.y;
                red[3] = bb.x * bb.x + bb.y * bb.y;
            }
        }
    }
    red[4] = nnz;
    }
    }
    if (partials) {
        if (res && !tv) {
            // the residual-only pass: |r|^2 alone, in k_residual's one-value record, so that
            // its reduction adds in the same order
            double r1[1] = {red[2]};
            block_reduce_store<1>(r1, partials);
        } else {
            block_reduce_store<5>(red, partials);
        }
    }
}

// basal_sliding_bc with HO_ASSEMBLE_BETA_STANDARD (velo_higher:8236-8363) in gather
// form, after the assembly: one lane per basal-layer node and horizontal slot (iA, jA, 0)
// adds, onto the assembled Auu and Avv coefficients, the terms
//   K = beta_qp w_p detJ / vol0 phi_r(p) phi_c(p)
// of every active cell that holds the node (corner r) and its neighbour (corner c): cells
// in the reference's loop order (j, i ascending), quadrature points in order, one term at
// a time as the reference accumulates them.  detJ comes from the 2D Jacobian of the cell's
// corners (get_basis_function_derivatives_2d, velo_higher:6583-6751), beta_qp is the
// bilinear interpolation of the corners' beta; Kvv = Kuu.
__global__ __launch_bounds__(256) void k_basal_standard(GridDev g, const int32_t *__restrict__ ac,
                                                        const double *__restrict__ xV,
                                                        const double *__restrict__ yV,
                                                        const double *__restrict__ beta,
                                                        double *__restrict__ A) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ncol = (long long)g.sx * g.sy;
    if (t >= 9 * ncol) return;
    const int s = (int)(t / ncol);
    const long long col = t - s * ncol;
    const int iA = s % 3 - 1, jA = s / 3 - 1;
    const int ii = (int)(col % g.sx) + 1, jj = (int)(col / g.sx) + 1;   // 1-based vertex
    // corners SW, SE, NE, NW of a cell relative to its NE vertex (ishift(3, n), jshift(3, n))
    const int di[4] = {-1, 0, 0, -1}, dj[4] = {-1, -1, 0, 0};
    const double r3 = 1.0 / sqrt(3.0);
    const double xq[4] = {-r3, r3, r3, -r3}, yq[4] = {-r3, -r3, r3, r3};
    const long long q = (long long)(g.nz - 1) + (long long)g.nz * col;   // basal-layer node
    const int m = slot_of(iA, jA, 0);
    double au = 0.0, av = 0.0;
    bool any = false;
    // the cells with this vertex as a corner, in (j, i) order; the vertex is their NE, NW,
    // SE and SW corner respectively
#pragma unroll 1
    for (int cc = 0; cc < 4; ++cc) {
        const int ci = ii + (cc & 1), cj = jj + (cc >> 1);
        const int nr = cc == 0 ? 2 : cc == 1 ? 3 : cc == 2 ? 1 : 0;
        if (ci < CISM_NHALO + 1 || ci > g.nx - CISM_NHALO + 1 || cj < CISM_NHALO + 1 ||
            cj > g.ny - CISM_NHALO + 1 || !ac[(ci - 1) + (long long)g.nx * (cj - 1)])
            continue;
        int nc = -1;   // the neighbour (ii + iA, jj + jA) as a corner of this cell
        for (int n = 0; n < 4; ++n)
            if (ci + di[n] == ii + iA && cj + dj[n] == jj + jA) nc = n;
        if (nc < 0) continue;
        double x[4], y[4], b[4];
        for (int n = 0; n < 4; ++n) {
            const long long v = (ci + di[n] - 1) + (long long)g.sx * (cj + dj[n] - 1);
            x[n] = xV[v];
            y[n] = yV[v];
            b[n] = beta[v];
        }
        if (!any) {
            au = A[aidx(0, m, q)];
            av = A[aidx(3, m, q)];
            any = true;
        }
        for (int p = 0; p < 4; ++p) {
            double phi[4], dxr[4], dyr[4];
            phi[0] = (1.0 - xq[p]) * (1.0 - yq[p]) / 4.0;
            phi[1] = (1.0 + xq[p]) * (1.0 - yq[p]) / 4.0;
            phi[2] = (1.0 + xq[p]) * (1.0 + yq[p]) / 4.0;
            phi[3] = (1.0 - xq[p]) * (1.0 + yq[p]) / 4.0;
            dxr[0] = -(1.0 - yq[p]) / 4.0; dxr[1] = (1.0 - yq[p]) / 4.0;
            dxr[2] = (1.0 + yq[p]) / 4.0;  dxr[3] = -(1.0 + yq[p]) / 4.0;
            dyr[0] = -(1.0 - xq[p]) / 4.0; dyr[1] = -(1.0 + xq[p]) / 4.0;
            dyr[2] = (1.0 + xq[p]) / 4.0;  dyr[3] = (1.0 - xq[p]) / 4.0;
            double J11 = 0.0, J12 = 0.0, J21 = 0.0, J22 = 0.0;
            for (int n = 0; n < 4; ++n) {
                J11 = J11 + dxr[n] * x[n];
                J12 = J12 + dxr[n] * y[n];
                J21 = J21 + dyr[n] * x[n];
                J22 = J22 + dyr[n] * y[n];
            }
            const double detJ = J11 * J22 - J12 * J21;
            double bq = 0.0;
            for (int n = 0; n < 4; ++n) bq = bq + phi[n] * b[n];
            const double K = 0.0 + bq * 1.0 * detJ / CISM_VOL0 * phi[nr] * phi[nc];
            au = au + K;
            av = av + K;
        }
    }
    if (any) {
        A[aidx(0, m, q)] = au;
        A[aidx(3, m, q)] = av;
    }
}

int basal_standard_device(cism_ctx *c, const double *beta) {
    const GridDev &g = c->g;
    const long long n = 9LL * g.sx * g.sy;
    c->a64s_valid = 0;
    hipLaunchKernelGGL(k_basal_standard, dim3(nblocks_for(n)), dim3(256), 0, c->stream, g, c->ac, c->xV,
                       c->yV, beta ? beta : c->beta, c->A);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

int assemble_device(cism_ctx *c, int whichapprox, int whichefvs, double efvs_constant,
                    int add_basal, double dx, double dy, float *As, int store64) {
    CISM_TRY(ensure_ke(c));
    c->a64s_valid = 0;
    const GridDev &g = c->g;
    const int nb = (int)((g.ne + EL_ELEM - 1) / EL_ELEM);
    auto kern = whichapprox == 0 ? k_element<0> : whichapprox == 1 ? k_element<1> : k_element<2>;
    hipLaunchKernelGGL(kern, dim3(nb), dim3(8 * EL_ELEM), 0, c->stream, g, c->sigma, c->ac, c->xV, c->yV, c->x,
                       c->susrf, c->sthck, c->flwafact, whichefvs, efvs_constant, c->Ke, c->efvs, c->sc);
    HIP_TRY(hipGetLastError());
    const double basal_scale = dx * dy / CISM_VOL0;
    const int nbg = nblocks_for(g.nn);
    if (!As) store64 = 1;
    hipLaunchKernelGGL(k_gather, dim3(nbg), dim3(256), 0, c->stream, g, c->ac, c->av,
                       c->a_zero_ok ? c->av_zero : nullptr, c->a_zero_ok ? c->ac_zero : nullptr, c->Ke, c->beta,
                       add_basal, basal_scale, c->A, store64, As, As && c->as_ok ? c->as_av_prev : nullptr,
                       As && c->as_ok ? c->as_ac_prev : nullptr, c->x, c->load, As ? c->t : nullptr,
                       As ? c->r : nullptr, As ? c->partials : nullptr);
    HIP_TRY(hipGetLastError());
    if (As) {
        // the FP32 matrix of this assembly (its halo rows follow the FP64 halo update,
        // launch_a32_halo); the residual of residual_vector_device(c, c->t, c->r) in c->t,
        // c->r (halo rows: launch_halo_vec(c->r)) and sums[0..3]; sums[4] = nonzeros of owned
        // active rows
        coarse_invalidate(c);
        CISM_TRY(launch_reduce(c, nbg, 5, RED_SUMS));
        HIP_TRY(hipMemcpyAsync(c->as_av_prev, c->av, sizeof(int32_t) * (size_t)g.sx * g.sy,
                               hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(c->as_ac_prev, c->ac, sizeof(int32_t) * (size_t)g.nx * g.ny,
                               hipMemcpyDeviceToDevice, c->stream));
        c->as_ok = 1;
    }
    HIP_TRY(hipMemcpyAsync(c->av_zero, c->av, sizeof(int32_t) * (size_t)g.sx * g.sy, hipMemcpyDeviceToDevice,
                           c->stream));
    HIP_TRY(hipMemcpyAsync(c->ac_zero, c->ac, sizeof(int32_t) * (size_t)g.nx * g.ny, hipMemcpyDeviceToDevice,
                           c->stream));
    // without the FP64 rows (store64 = 0) neither the matrix nor its zero rows are current
    c->a_zero_ok = store64;
    c->a64_valid = store64;
    c->asm_add_basal = add_basal;
    c->asm_basal_scale = basal_scale;
    return CISM_OK;
}

// ---------------------------------------------------------------------------
// dirichlet_boundary_conditions_3d (velo_higher:8385-8578) in gather form: one
// lane per row N applies (a) N's own row replacement when N is a prescribed
// node, and (b) the column eliminations of every prescribed neighbour D, in the
// reference's order of D (j, i, k ascending, u before v), so the right-hand
// side accumulates its corrections in the reference's order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_dirichlet(GridDev g, const int32_t *__restrict__ av,
                                                   const int32_t *__restrict__ um,
                                                   const int32_t *__restrict__ vm,
                                                   const double *__restrict__ vel,
                                                   double *__restrict__ A, double *__restrict__ b) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nn) return;
    const int k = (int)(q % g.nz) + 1;
    const long long col = q / g.nz;
    const int i = (int)(col % g.sx) + 1, j = (int)(col / g.sx) + 1;
    // prescribed-node loop range: i, j in nhalo .. n-nhalo+1 and active (velo_higher:8441-8443)
    auto in_range = [&](int ii, int jj) {
        return ii >= CISM_NHALO && ii <= g.nx - CISM_NHALO + 1 && jj >= CISM_NHALO &&
               jj <= g.ny - CISM_NHALO + 1 && av[(ii - 1) + (long long)g.sx * (jj - 1)] != 0;
    };
    const bool Nu = um[q] == 1, Nv = vm[q] == 1;
    const bool Nrange = in_range(i, j);
    // (a) own row
    if (Nrange && (Nu || Nv)) {
        for (int m = 0; m < 27; ++m) {
            const int iA = m % 3 - 1, jA = (m / 3) % 3 - 1, kA = m / 9 - 1;
            if (k + kA < 1 || k + kA > g.nz || i + iA < 1 || i + iA > g.sx || j + jA < 1 || j + jA > g.sy)
                continue;
            if (m == CISM_DIAG) continue;
            if (Nu) { A[aidx(0, m, q)] = 0.0; A[aidx(1, m, q)] = 0.0; }
            if (Nv) { A[aidx(2, m, q)] = 0.0; A[aidx(3, m, q)] = 0.0; }
        }
        if (Nu) {
            b[2 * q] = vel[2 * q];
            A[aidx(0, CISM_DIAG, q)] = 1.0;
            A[aidx(1, CISM_DIAG, q)] = 0.0;
            A[aidx(2, CISM_DIAG, q)] = 0.0;
        }
        if (Nv) {
            b[2 * q + 1] = vel[2 * q + 1];
            A[aidx(1, CISM_DIAG, q)] = 0.0;
            A[aidx(2, CISM_DIAG, q)] = 0.0;
            A[aidx(3, CISM_DIAG, q)] = 1.0;
        }
    }
    if (Nu && Nv) return;   // no column elimination touches a doubly prescribed row
    // (b) column eliminations from prescribed neighbours D = N + o
    double bu = b[2 * q], bv = b[2 * q + 1];
    bool touched = false;
    for (int oj = -1; oj <= 1; ++oj)
        for (int oi = -1; oi <= 1; ++oi)
            for (int ok = -1; ok <= 1; ++ok) {
                if (oi == 0 && oj == 0 && ok == 0) continue;
                const int di = i + oi, dj = j + oj, dk = k + ok;
                if (dk < 1 || dk > g.nz || di < 1 || di > g.sx || dj < 1 || dj > g.sy) continue;
                const long long qd = (dk - 1) + (long long)g.nz * ((di - 1) + (long long)g.sx * (dj - 1));
                const bool Du = um[qd] == 1, Dv = vm[qd] == 1;
                if (!(Du || Dv) || !in_range(di, dj)) continue;
                const int m = slot_of(oi, oj, ok);
                if (Du) {
                    const double ud = vel[2 * qd];
                    if (!Nu) { bu = bu - A[aidx(0, m, q)] * ud; A[aidx(0, m, q)] = 0.0; }
                    if (!Nv) { bv = bv - A[aidx(2, m, q)] * ud; A[aidx(2, m, q)] = 0.0; }
                }
                if (Dv) {
                    const double vd = vel[2 * qd + 1];
                    if (!Nu) { bu = bu - A[aidx(1, m, q)] * vd; A[aidx(1, m, q)] = 0.0; }
                    if (!Nv) { bv = bv - A[aidx(3, m, q)] * vd; A[aidx(3, m, q)] = 0.0; }
                }
                touched = true;
            }
    if (touched) {
        if (!Nu) b[2 * q] = bu;
        if (!Nv) b[2 * q + 1] = bv;
    }
}

int dirichlet_device(cism_ctx *c) {
    if (!c->dirichlet_any) return CISM_OK;   // no prescribed node: nothing to replace or eliminate
    c->a64s_valid = 0;
    hipLaunchKernelGGL(k_dirichlet, dim3(nblocks_for(c->g.nn)), dim3(256), 0, c->stream, c->g, c->av,
                       c->umask, c->vmask, c->x, c->A, c->b);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// ---------------------------------------------------------------------------
// check_symmetry_assembled_matrix_3d (velo_higher:9533-9741).  Each pair of
// coefficients (A(m, row), A'(26-m, col)) is disjoint from every other pair.  The
// reference visits a pair from each of its two rows that is owned and active, and
// replaces both by their mean as soon as one visit finds |v1 - v2| < 1e-8 |diag|
// (the visiting row's Auu diagonal for the Auu and Auv tests, its Avv diagonal for
// the Avv and Avu tests); a repaired pair compares equal at the other visit.  So a
// pair is repaired iff it qualifies from at least one visiting side.
//
// Here one lane per node owns the pairs of its slots CISM_DIAG .. 26 (the diagonal
// slot's Auv/Avu pair and the 13 couplings after it), tests both sides on the
// unmodified values and writes both coefficients of a repaired pair.  No other lane
// reads or writes those coefficients (a lane reads only its own slots from CISM_DIAG
// on and, of other nodes, the lower partner slots of its own pairs and the diagonal
// Auu/Avv, which no repair touches), so the pass is in place, race-free and
// independent of the launch.  partials, per block: [0] the visits that find an
// asymmetry above the threshold (the reference's warnings, on the unmodified matrix),
// [1] the nonzero coefficients of owned active rows after the repair
// (count_nonzeros_3d, velo_higher:9355-9410; a row's coefficients before its diagonal
// are counted by the lane that owns their pair).  As: when not null, the lane also
// writes its node's FP32 symmetric storage (slots CISM_DIAG .. 26, sidx) from the
// repaired values, which saves the mixed-precision path a separate conversion pass.
// ---------------------------------------------------------------------------
//
// skip (ac, ac_prev, av_prev; when A is the assembly's and As was written by such a
// pass before): an owned node with an inactive vertex and no active cell around it,
// now and at that pass, is left out.  Its rows of A are zero (no element touches
// it, and the Dirichlet and basal updates act on active vertices), so are its
// couplings in the neighbours' rows, and its rows of As are zero from that pass.
__global__ __launch_bounds__(256) void k_sym_repair(GridDev g, const int32_t *__restrict__ av,
                                                    double *__restrict__ A, float *__restrict__ As,
                                                    double *__restrict__ As64,
                                                    const int32_t *__restrict__ ac,
                                                    const int32_t *__restrict__ ac_prev,
                                                    const int32_t *__restrict__ av_prev,
                                                    double *__restrict__ partials) {
    double acc[2] = {0.0, 0.0};
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    bool skip = q >= g.nn;
    int k = 0, i = 0, j = 0;
    long long col = 0;
    if (!skip) {
        k = (int)(q % g.nz);
        col = q / g.nz;
        i = (int)(col % g.sx);
        j = (int)(col / g.sx);
        skip = ac && !av[col] && !av_prev[col] && is_owned(g, i, j) && !touches_active_cell(g, ac, i, j) &&
               !touches_active_cell(g, ac_prev, i, j);
    }
    if (!skip) {
        const bool vis_q = is_owned(g, i, j) && av[col];
        const double dgu = A[aidx(0, CISM_DIAG, q)], dgv = A[aidx(3, CISM_DIAG, q)];
        double warn = 0.0, nnz = 0.0;
        auto count4 = [&](long long e0, long long e2) {   // four blocks: (b0, b1) at e0, (b2, b3) at e2
            return (double)((A[e0] != 0.0) + (A[e0 + 1] != 0.0) + (A[e2] != 0.0) + (A[e2 + 1] != 0.0));
        };
        // slot m of row q is final once its pairs are repaired (only this lane writes a
        // row's slots from the diagonal on): its FP32 and FP64 symmetric copies
        auto store32 = [&](int m) {
            if (!As && !As64) return;
            const double2 a = *reinterpret_cast<const double2 *>(A + aidx(0, m, q));
            const double2 b = *reinterpret_cast<const double2 *>(A + aidx(2, m, q));
            if (As) {
                float4 o;
                o.x = (float)a.x; o.y = (float)a.y; o.z = (float)b.x; o.w = (float)b.y;
                *reinterpret_cast<float4 *>(As + sidx(0, m - CISM_DIAG, q)) = o;
            }
            if (As64) {
                const dvec2_nt a2 = {a.x, a.y}, b2 = {b.x, b.y};
                __builtin_nontemporal_store(a2, reinterpret_cast<dvec2_nt *>(As64 + s64idx(0, m - CISM_DIAG, q)));
                __builtin_nontemporal_store(b2, reinterpret_cast<dvec2_nt *>(As64 + s64idx(2, m - CISM_DIAG, q)));
            }
        };
        // one coefficient pair (e1, e2) with values v1, v2; qualification from each side
        auto pair = [&](long long e1, long long e2, double v1, double v2, bool va, double da, bool vb,
                        double db) {
            if (v1 == v2) return;
            const double d = fabs(v2 - v1);
            const bool qa = va && d < 1.0e-8 * fabs(da), qb = vb && d < 1.0e-8 * fabs(db);
            warn += (va && !qa ? 1.0 : 0.0) + (vb && !qb ? 1.0 : 0.0);
            if (qa || qb) {
                const double avg = 0.5 * (v1 + v2);
                A[e1] = avg;
                A[e2] = avg;
            }
        };
        // the diagonal slot: Auv(q) ~ Avu(q), tested against both diagonals
        {
            const long long e1 = aidx(1, CISM_DIAG, q), e2 = aidx(2, CISM_DIAG, q);
            const double v1 = A[e1], v2 = A[e2];
            if (vis_q && v1 != v2) {
                const double d = fabs(v2 - v1);
                const bool qa = d < 1.0e-8 * fabs(dgu), qb = d < 1.0e-8 * fabs(dgv);
                warn += (qa ? 0.0 : 1.0) + (qb ? 0.0 : 1.0);
                if (qa || qb) {
                    const double avg = 0.5 * (v1 + v2);
                    A[e1] = avg;
                    A[e2] = avg;
                }
            }
            if (vis_q) nnz += count4(aidx(0, CISM_DIAG, q), aidx(2, CISM_DIAG, q));
            store32(CISM_DIAG);
        }
        for (int m = CISM_DIAG + 1; m < 27; ++m) {
            const int iA = m % 3 - 1, jA = (m / 3) % 3 - 1, kA = m / 9 - 1;
            if (k + kA < 0 || k + kA >= g.nz || i + iA < 0 || i + iA >= g.sx || j + jA < 0 || j + jA >= g.sy) {
                if (vis_q) nnz += count4(aidx(0, m, q), aidx(2, m, q));
                store32(m);
                continue;
            }
            const int mm = 26 - m;
            const long long qc = q + kA + (long long)g.nz * (iA + (long long)g.sx * jA);
            const long long cc = col + iA + (long long)g.sx * jA;
            const bool vis_c = is_owned(g, i + iA, j + jA) && av[cc];
            if (!vis_q && !vis_c) {
                store32(m);
                continue;
            }
            const double cgu = A[aidx(0, CISM_DIAG, qc)], cgv = A[aidx(3, CISM_DIAG, qc)];
            double a0, a1, a2, a3, b0, b1, b2, b3;
            load_blocks(A, q, m, a0, a1, a2, a3);
            load_blocks(A, qc, mm, b0, b1, b2, b3);
            pair(aidx(0, m, q), aidx(0, mm, qc), a0, b0, vis_q, dgu, vis_c, cgu);   // Auu ~ Auu
            pair(aidx(1, m, q), aidx(2, mm, qc), a1, b2, vis_q, dgu, vis_c, cgv);   // Auv ~ Avu
            pair(aidx(2, m, q), aidx(1, mm, qc), a2, b1, vis_q, dgv, vis_c, cgu);   // Avu ~ Auv
            pair(aidx(3, m, q), aidx(3, mm, qc), a3, b3, vis_q, dgv, vis_c, cgv);   // Avv ~ Avv
            if (vis_q) nnz += count4(aidx(0, m, q), aidx(2, m, q));
            if (vis_c) nnz += count4(aidx(0, mm, qc), aidx(2, mm, qc));
            store32(m);
        }
        acc[0] = warn;
        acc[1] = nnz;
    }
    block_reduce_store<2>(acc, partials);
}

// symmetry repair; sums[0] = warnings, sums[1] = nonzeros of owned active rows.  As: the
// FP32 symmetric storage to write as well, or null; As64 (c->A64s): the FP64 one, valid
// after sym64_commit when the pass counted no warning.
int symmetry_device(cism_ctx *c, float *As, double *As64) {
    const GridDev &g = c->g;
    const int nb = nblocks_for(g.nn);
    // rows outside the ice may be left out only when A is the assembly's (a_zero_ok)
    // and As was last written from such an A (never with As64, which covers every row)
    const bool skip = As && !As64 && c->a_zero_ok && c->as_ok;
    c->a64s_valid = 0;
    c->a64s_pending = As64 != nullptr;
    hipLaunchKernelGGL(k_sym_repair, dim3(nb), dim3(256), 0, c->stream, g, c->av, c->A, As, As64,
                       skip ? c->ac : nullptr, c->as_ac_prev, c->as_av_prev, c->partials);
    HIP_TRY(hipGetLastError());
    if (As) {
        coarse_invalidate(c);
        c->as_ok = c->a_zero_ok;
        if (c->as_ok) {
            HIP_TRY(hipMemcpyAsync(c->as_av_prev, c->av, sizeof(int32_t) * (size_t)g.sx * g.sy,
                                   hipMemcpyDeviceToDevice, c->stream));
            HIP_TRY(hipMemcpyAsync(c->as_ac_prev, c->ac, sizeof(int32_t) * (size_t)g.nx * g.ny,
                                   hipMemcpyDeviceToDevice, c->stream));
        }
    }
    CISM_TRY(launch_reduce(c, nb, 2, RED_SUMS));
    return CISM_OK;
}

void sym64_commit(cism_ctx *c) {
    c->a64s_valid = c->a64s_pending && c->sc_host->sums[0] == 0.0;
    c->a64s_pending = 0;
}

// r = b - A x (c->r; halo rows from their images) of the last assembly's matrix, formed
// from its element matrices by a pass of the gather that stores nothing else: the FP64
// residual of the mixed path after an assembly without the FP64 rows, the same values as
// k_residual's from the stored matrix (and the same norm, reduced in the same order).
// sums[0] = |r|^2 over owned nodes.  x must carry its halo.
int residual_from_elements(cism_ctx *c) {
    const GridDev &g = c->g;
    const int nbg = nblocks_for(g.nn);
    hipLaunchKernelGGL(k_gather, dim3(nbg), dim3(256), 0, c->stream, g, c->ac, c->av, (const int32_t *)nullptr,
                       (const int32_t *)nullptr, c->Ke, c->beta, c->asm_add_basal, c->asm_basal_scale, c->A, 0,
                       (float *)nullptr, (const int32_t *)nullptr, (const int32_t *)nullptr, c->x, c->b,
                       (double *)nullptr, c->r, c->partials);
    HIP_TRY(hipGetLastError());
    CISM_TRY(launch_reduce(c, nbg, 1, RED_SUMS));
    return launch_halo_vec(c, c->r);
}

// count_nonzeros_3d (velo_higher:9355-9410) -> sums[0]
__global__ __launch_bounds__(256) void k_nnz(GridDev g, const int32_t *__restrict__ av,
                                             const double *__restrict__ A, double *__restrict__ partials) {
    double acc[1] = {0.0};
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long col = q / g.nz;
    const int i = (int)(col % g.sx), j = (int)(col / g.sx);
    if (q < g.nn && is_owned(g, i, j) && av[col]) {
        int n = 0;
        for (int m = 0; m < 27; ++m) {
            double a0, a1, a2, a3;
            load_blocks_nt(A, q, m, a0, a1, a2, a3);
            n += (a0 != 0.0) + (a1 != 0.0) + (a2 != 0.0) + (a3 != 0.0);
        }
        acc[0] = n;
    }
    block_reduce_store<1>(acc, partials);
}

int count_nonzeros_device(cism_ctx *c) {
    const int nb = nblocks_for(c->g.nn);
    hipLaunchKernelGGL(k_nnz, dim3(nb), dim3(256), 0, c->stream, c->g, c->av, c->A, c->partials);
    HIP_TRY(hipGetLastError());
    return launch_reduce(c, nb, 1, RED_SUMS);
}

// compute_residual_vector_3d (velo_higher:8768-8927): r = A x - b on owned
// active rows (0 elsewhere); sums[0] = |r|^2, sums[1] = |b|^2 over the same rows.
// With rs (the first residual of the iterative refinement, which needs the same
// product): rs = b - A x on the whole grid, the negation of r on owned active
// rows, b on owned inactive rows, and on halo rows the value of their periodic
// image (0 without one); sums[2] = |rs|^2 and sums[3] = |b|^2 over owned rows.
template <bool RS>
__global__ __launch_bounds__(256) void k_resvec(GridDev g, const int32_t *__restrict__ av,
                                                const double *__restrict__ A, const double *__restrict__ x,
                                                const double *__restrict__ b, double *__restrict__ r,
                                                double *__restrict__ rs, double *__restrict__ partials) {
    constexpr int NV = RS ? 4 : 2;
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    const long long cst = g.nz, rst = (long long)g.nz * g.sx;
    // one lane per node
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < g.nn) {
        const int k = (int)(q % g.nz);
        const long long col = q / g.nz;
        int i = (int)(col % g.sx), j = (int)(col / g.sx);
        const bool own = is_owned(g, i, j);
        long long qs = q;     // row whose product this lane forms
        bool none = false;    // a halo row without an image
        if (RS && !own) {
            const int ii = halo_image(i, g.sx, g.periodic), jj = halo_image(j, g.sy, g.periodic);
            if (ii < 0 || jj < 0) none = true;
            else { i = ii; j = jj; qs = k + (long long)g.nz * (i + (long long)g.sx * j); }
        } else if (!own) {
            none = true;
        }
        double ru = 0.0, rv = 0.0;
        bool act = false;
        if (!none && av[qs / g.nz]) {
            act = true;
            double su = 0.0, sv = 0.0;
#pragma unroll
            for (int m = 0; m < 27; ++m) {
                const int iA = m % 3 - 1, jA = (m / 3) % 3 - 1, kA = m / 9 - 1;
                if (k + kA < 0 || k + kA >= g.nz) continue;
                const long long qn = qs + kA + iA * cst + jA * rst;
                const double2 xx = *reinterpret_cast<const double2 *>(x + 2 * qn);
                double uu, uv, vu, vv;
                load_blocks_nt(A, qs, m, uu, uv, vu, vv);
                su = su + uu * xx.x + uv * xx.y;
                sv = sv + vu * xx.x + vv * xx.y;
            }
            ru = su - b[2 * qs];
            rv = sv - b[2 * qs + 1];
        }
        if (own) {
            r[2 * q] = ru;
            r[2 * q + 1] = rv;
            const double bu = b[2 * q], bv = b[2 * q + 1];
            if (act) {
                acc[0] += ru * ru + rv * rv;
                acc[1] += bu * bu + bv * bv;
            }
            if (RS) {
                const double su = act ? -ru : bu, sv = act ? -rv : bv;
                rs[2 * q] = su;
                rs[2 * q + 1] = sv;
                acc[NV - 2] += su * su + sv * sv;
                acc[NV - 1] += bu * bu + bv * bv;
            }
        } else {
            r[2 * q] = 0.0;
            r[2 * q + 1] = 0.0;
            if (RS) {
                rs[2 * q] = none ? 0.0 : act ? -ru : b[2 * qs];
                rs[2 * q + 1] = none ? 0.0 : act ? -rv : b[2 * qs + 1];
            }
        }
    }
    block_reduce_store<NV>(acc, partials);
}

int residual_vector_device(cism_ctx *c, double *ru, double *rs) {
    const int nb = nblocks_for(c->g.nn);
    if (rs)
        hipLaunchKernelGGL(k_resvec<true>, dim3(nb), dim3(256), 0, c->stream, c->g, c->av, c->A, c->x, c->b,
                           ru, rs, c->partials);
    else
        hipLaunchKernelGGL(k_resvec<false>, dim3(nb), dim3(256), 0, c->stream, c->g, c->av, c->A, c->x,
                           c->b, ru, nullptr, c->partials);
    HIP_TRY(hipGetLastError());
    return launch_reduce(c, nb, rs ? 4 : 2, RED_SUMS);
}

// compute_residual_velocity_3d (velo_higher:9111-9234) -> sums[0] (and the
// count in sums[1] for the mean variant)
__global__ __launch_bounds__(256) void k_resvel(GridDev g, int whichresid, const double *__restrict__ x,
                                                const double *__restrict__ xs,
                                                double *__restrict__ partials) {
    double acc[2] = {0.0, 0.0};
    const long long stride = (long long)gridDim.x * blockDim.x;
    const int kmax = (whichresid == 1 || whichresid == 2) ? g.nz - 1 : g.nz;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < g.nn; q += stride) {
        const int k = (int)(q % g.nz);
        const long long col = q / g.nz;
        const int i = (int)(col % g.sx), j = (int)(col / g.sx);
        if (!is_owned(g, i, j) || k >= kmax) continue;
        const double u = x[2 * q], v = x[2 * q + 1];
        const double speed = sqrt(u * u + v * v);
        if (speed != 0.0) {
            const double uo = xs[2 * q], vo = xs[2 * q + 1];
            const double old = sqrt(uo * uo + vo * vo);
            const double dd = fabs((old - speed) / speed);
            if (whichresid == 2) { acc[0] += dd; acc[1] += 1.0; }
            else acc[0] = fmax(acc[0], dd);
        }
    }
    if (whichresid == 2) block_reduce_store<2>(acc, partials);
    else block_max_store<2>(acc, partials);
}

int residual_velocity_device(cism_ctx *c, int whichresid) {
    const int nb = red_blocks_for(c->g.nn);
    hipLaunchKernelGGL(k_resvel, dim3(nb), dim3(256), 0, c->stream, c->g, whichresid, c->x, c->usav,
                       c->partials);
    HIP_TRY(hipGetLastError());
    return launch_reduce(c, nb, 2, whichresid == 2 ? RED_SUMS : RED_MAX);
}

// load_vector_gravity (velo_higher:4317-4532).  LOCAL: one lane per node.
// STANDARD: one lane per node gathers, element by element in (j, i, k) order and
// quadrature point by quadrature point, the terms the reference scatters.
__global__ __launch_bounds__(256) void k_load_local(GridDev g, const double *__restrict__ wsig,
                                                    double dx, double dy, const int32_t *__restrict__ av,
                                                    const double *__restrict__ sthck,
                                                    const double *__restrict__ dsdx,
                                                    const double *__restrict__ dsdy,
                                                    double *__restrict__ load) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nn) return;
    const int k = (int)(q % g.nz) + 1;
    const long long col = q / g.nz;
    if (!av[col]) return;
    const double dz = sthck[col] * (wsig[k] - wsig[k - 1]);
    const double f = CISM_RHOI * CISM_GRAV * dx * dy * dz / CISM_VOL0;
    load[2 * q] = load[2 * q] - f * dsdx[col];
    load[2 * q + 1] = load[2 * q + 1] - f * dsdy[col];
}

__global__ __launch_bounds__(256) void k_load_standard(GridDev g, const double *__restrict__ sigma,
                                                       const int32_t *__restrict__ ac,
                                                       const double *__restrict__ xV,
                                                       const double *__restrict__ yV,
                                                       const double *__restrict__ susrf,
                                                       const double *__restrict__ sthck,
                                                       const double *__restrict__ dsdx,
                                                       const double *__restrict__ dsdy,
                                                       double *__restrict__ load) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nn) return;
    const int k = (int)(q % g.nz) + 1;
    const long long col = q / g.nz;
    const int i = (int)(col % g.sx) + 1, j = (int)(col / g.sx) + 1;
    double lu = load[2 * q], lv = load[2 * q + 1];
    const double r = 1.0 / sqrt(3.0);
    for (int dJ = 0; dJ < 2; ++dJ)
        for (int dI = 0; dI < 2; ++dI)
            for (int dK = 0; dK < 2; ++dK) {
                const int icl = i + dI, jcl = j + dJ, ke = k - 1 + dK;
                if (icl < CISM_NHALO + 1 || icl > g.nx - CISM_NHALO + 1 || jcl < CISM_NHALO + 1 ||
                    jcl > g.ny - CISM_NHALO + 1 || ke < 1 || ke > g.nz - 1)
                    continue;
                if (!ac[(icl - 1) + (long long)g.nx * (jcl - 1)]) continue;
                const int me = node_id(1 - dI, 1 - dJ, 1 - dK);
                double X[8], Y[8], Z[8], SX[8], SY[8];
                for (int n = 0; n < 8; ++n) {
                    const int in = icl - 1 + c_node_ix(n), jn = jcl - 1 + c_node_jy(n), kn = ke + c_node_kl(n);
                    const long long v = (in - 1) + (long long)g.sx * (jn - 1);
                    X[n] = xV[v]; Y[n] = yV[v];
                    Z[n] = susrf[v] - sigma[kn - 1] * sthck[v];
                    SX[n] = dsdx[v]; SY[n] = dsdy[v];
                }
                for (int p = 0; p < 8; ++p) {
                    const double xq = (p == 0 || p == 3 || p == 4 || p == 7) ? -r : r;
                    const double yq = (p < 2 || p == 4 || p == 5) ? -r : r;
                    const double zq = p < 4 ? -r : r;
                    double a[8], b[8], cz[8], ph[8];
                    for (int n = 0; n < 8; ++n) {
                        const double sx = c_node_ix(n) ? 1.0 : -1.0, sy = c_node_jy(n) ? 1.0 : -1.0;
                        const double sz = c_node_kl(n) ? -1.0 : 1.0;
                        const double ax = sx > 0 ? 1.0 + xq : 1.0 - xq;
                        const double ay = sy > 0 ? 1.0 + yq : 1.0 - yq;
                        const double az = sz > 0 ? 1.0 + zq : 1.0 - zq;
                        ph[n] = ax * ay * az / 8.0;
                        a[n] = sx * (ay * az / 8.0);
                        b[n] = sy * (ax * az / 8.0);
                        cz[n] = sz * (ax * ay / 8.0);
                    }
                    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                    for (int n = 0; n < 8; ++n) {
                        J[0][0] = J[0][0] + a[n] * X[n];  J[0][1] = J[0][1] + a[n] * Y[n];  J[0][2] = J[0][2] + a[n] * Z[n];
                        J[1][0] = J[1][0] + b[n] * X[n];  J[1][1] = J[1][1] + b[n] * Y[n];  J[1][2] = J[1][2] + b[n] * Z[n];
                        J[2][0] = J[2][0] + cz[n] * X[n]; J[2][1] = J[2][1] + cz[n] * Y[n]; J[2][2] = J[2][2] + cz[n] * Z[n];
                    }
                    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
                    const double c01 = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]);
                    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
                    const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
                    double qx = 0.0, qy = 0.0;
                    for (int n = 0; n < 8; ++n) { qx = qx + ph[n] * SX[n]; qy = qy + ph[n] * SY[n]; }
                    lu = lu - CISM_RHOI * CISM_GRAV * 1.0 * det / CISM_VOL0 * qx * ph[me];
                    lv = lv - CISM_RHOI * CISM_GRAV * 1.0 * det / CISM_VOL0 * qy * ph[me];
                }
            }
    load[2 * q] = lu;
    load[2 * q + 1] = lv;
}

int load_gravity_device(cism_ctx *c, int whichassemble_taud, double dx, double dy) {
    const GridDev &g = c->g;
    if (whichassemble_taud == 1)
        hipLaunchKernelGGL(k_load_local, dim3(nblocks_for(g.nn)), dim3(256), 0, c->stream, g, c->wsig,
                           dx, dy, c->av, c->sthck, c->dsdx, c->dsdy, c->load);
    else
        hipLaunchKernelGGL(k_load_standard, dim3(nblocks_for(g.nn)), dim3(256), 0, c->stream, g,
                           c->sigma, c->ac, c->xV, c->yV, c->susrf, c->sthck, c->dsdx, c->dsdy,
                           c->load);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// ---------------------------------------------------------------------------
// load_vector_lateral_bc + lateral_shelf_bc (velo_higher:4536-4950): pressure of
// ice against ocean on vertical faces of active cells that border an ocean cell
// (or an inactive calving-front cell).  One lane per node gathers, in the
// reference's order (cells j, i; faces W, E, S, N; layers; Gauss points), the
// terms the reference scatters onto the node.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_load_lateral(GridDev g, const double *__restrict__ sigma,
                                                      const int32_t *__restrict__ ocean,
                                                      const int32_t *__restrict__ cfront,
                                                      const int32_t *__restrict__ ac,
                                                      const double *__restrict__ xV,
                                                      const double *__restrict__ yV,
                                                      const double *__restrict__ susrf,
                                                      const double *__restrict__ sthck,
                                                      double *__restrict__ load) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nn) return;
    const int k = (int)(q % g.nz) + 1;
    const long long col = q / g.nz;
    const int i = (int)(col % g.sx) + 1, j = (int)(col / g.sx) + 1;
    auto C2i = [&](const int32_t *a, int ii, int jj) { return a[(ii - 1) + (long long)g.nx * (jj - 1)]; };
    auto V2 = [&](const double *a, int ii, int jj) { return a[(ii - 1) + (long long)g.sx * (jj - 1)]; };
    double lu = load[2 * q], lv = load[2 * q + 1];
    bool any = false;
    const double r = 1.0 / sqrt(3.0);
    for (int dJ = 0; dJ < 2; ++dJ)
        for (int dI = 0; dI < 2; ++dI) {
            const int ic = i + dI, jc = j + dJ;
            if (ic < CISM_NHALO + 1 || ic > g.nx - CISM_NHALO + 1 || jc < CISM_NHALO + 1 ||
                jc > g.ny - CISM_NHALO + 1 || !C2i(ac, ic, jc))
                continue;
            const int ni[4] = {ic - 1, ic + 1, ic, ic}, nj[4] = {jc, jc, jc - 1, jc + 1};
            for (int f = 0; f < 4; ++f) {
                if (!(C2i(ocean, ni[f], nj[f]) == 1 ||
                      (C2i(cfront, ni[f], nj[f]) == 1 && !C2i(ac, ni[f], nj[f]))))
                    continue;
                int in0, jn0, in1, jn1;
                if (f == 0) { in0 = ic - 1; jn0 = jc; in1 = ic - 1; jn1 = jc - 1; }
                else if (f == 1) { in0 = ic; jn0 = jc - 1; in1 = ic; jn1 = jc; }
                else if (f == 2) { in0 = ic - 1; jn0 = jc - 1; in1 = ic; jn1 = jc - 1; }
                else { in0 = ic; jn0 = jc; in1 = ic - 1; jn1 = jc; }
                int n0;
                if (in0 == i && jn0 == j) n0 = 0;
                else if (in1 == i && jn1 == j) n0 = 1;
                else continue;
                double x[4], s[4], h[4];
                x[0] = f <= 1 ? V2(yV, in0, jn0) : V2(xV, in0, jn0);
                x[1] = f <= 1 ? V2(yV, in1, jn1) : V2(xV, in1, jn1);
                x[2] = x[1]; x[3] = x[0];
                s[0] = V2(susrf, in0, jn0); s[1] = V2(susrf, in1, jn1); s[2] = s[1]; s[3] = s[0];
                h[0] = V2(sthck, in0, jn0); h[1] = V2(sthck, in1, jn1); h[2] = h[1]; h[3] = h[0];
                for (int kf = k - 1; kf <= k; ++kf) {
                    if (kf < 1 || kf > g.nz - 1) continue;
                    const int n = kf == k - 1 ? n0 : 3 - n0;   // lower pair 0,1; upper 3 (above 0), 2 (above 1)
                    double y[4];
                    y[0] = s[0] - sigma[kf] * h[0];
                    y[1] = s[1] - sigma[kf] * h[1];
                    y[2] = s[2] - sigma[kf - 1] * h[2];
                    y[3] = s[3] - sigma[kf - 1] * h[3];
                    for (int p = 0; p < 4; ++p) {
                        const double xq = (p == 0 || p == 3) ? -r : r, yq = p < 2 ? -r : r;
                        double phi[4], dxr[4], dyr[4];
                        for (int m = 0; m < 4; ++m) {
                            const double sx = (m == 0 || m == 3) ? -1.0 : 1.0, sy = m < 2 ? -1.0 : 1.0;
                            const double ax = sx > 0 ? 1.0 + xq : 1.0 - xq;
                            const double ay = sy > 0 ? 1.0 + yq : 1.0 - yq;
                            phi[m] = ax * ay / 4.0;
                            dxr[m] = sx * (ay / 4.0);
                            dyr[m] = sy * (ax / 4.0);
                        }
                        double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
                        for (int m = 0; m < 4; ++m) {
                            J11 = J11 + dxr[m] * x[m];
                            J12 = J12 + dxr[m] * y[m];
                            J21 = J21 + dyr[m] * x[m];
                            J22 = J22 + dyr[m] * y[m];
                        }
                        const double detJ = fabs(J11 * J22 - J12 * J21);
                        double hq = 0.0, sq = 0.0;
                        for (int m = 0; m < 4; ++m) { hq = hq + phi[m] * h[m]; sq = sq + phi[m] * s[m]; }
                        const double ratio = sq / hq < 1.0 ? sq / hq : 1.0;
                        const double p_av = 0.5 * CISM_RHOI * CISM_GRAV * hq -
                                            0.5 * CISM_RHOO * CISM_GRAV * hq * ((1.0 - ratio) * (1.0 - ratio));
                        const double t = p_av * 1.0 * detJ / CISM_VOL0 * phi[n];
                        if (f == 0) lu = lu - t;
                        else if (f == 1) lu = lu + t;
                        else if (f == 2) lv = lv - t;
                        else lv = lv + t;
                        any = true;
                    }
                }
            }
        }
    if (any) {
        load[2 * q] = lu;
        load[2 * q + 1] = lv;
    }
}

int load_lateral_device(cism_ctx *c, const int32_t *ocean, const int32_t *cfront) {
    hipLaunchKernelGGL(k_load_lateral, dim3(nblocks_for(c->g.nn)), dim3(256), 0, c->stream, c->g, c->sigma,
                       ocean, cfront, c->ac, c->xV, c->yV, c->susrf, c->sthck, c->load);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

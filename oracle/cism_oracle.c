/*
 * cism_oracle.c -- CPU oracle (test infrastructure only; see cism_oracle.h).
 *
 * A plain-C restatement of the reference's 3D Blatter-Pattyn path.  Each
 * function cites the reference routine it follows (paths relative to the
 * CISM source tree).  Loop orders, index ranges and the order of floating
 * point operations follow the reference so that the oracle reproduces its
 * rounding, not only its mathematics.
 *
 * What pins each function (tests/test_oracle_*.py):
 *   orc_tables                 literal ishift/jshift/kshift/indxA tables of the
 *                              reference; partition of unity; sum of derivatives 0
 *   orc_basis_derivatives      box element: J = diag(dx/2,dy/2,dz/2), detJ closed form;
 *                              sheared column: analytic J; affine field reproduced exactly
 *   orc_efvs_qp                linear velocity field -> analytic strain rates -> closed form
 *   orc_element_matrix         box element, constant efvs: Kronecker sums of 1D linear
 *                              stiffness / mass matrices; Kvu == Kuv^T bitwise
 *   orc_assemble_3d            brute-force scipy CSR assembly from element matrices;
 *                              rigid translation u=const in the interior; symmetry
 *   orc_dirichlet_3d           explicit column-elimination formula on a dense matrix
 *   orc_halo                   explicit periodic index maps
 *   orc_load_lateral_bc        closed form: uniform face -> p_av * H * face length
 *   orc_matvec_3d              brute-force CSR product
 *   orc_tridiag                scipy.linalg.solve_banded
 *   orc_pcg_*                  scipy.sparse.linalg.spsolve; standard vs Chron-Gear
 *                              iterates agree; niters equal
 *   orc_picard_solve           slab analytic first-order solution (tests/slab in the
 *                              reference); converged-residual invariant
 */
#include "cism_oracle.h"

/* ORC_OMP (liborc_omp.so only): the hot loops run on all host cores with OpenMP, for the
 * all-core CPU baseline.  That build sums in another order and is never the checker; the
 * serial build (liborc.so) compiles the pragmas away and keeps the reference's order. */
#ifdef ORC_OMP
#define ORC_PAR_FOR _Pragma("omp parallel for schedule(static)")
#define ORC_PAR_SUM_S _Pragma("omp parallel for schedule(static) reduction(+:s)")
#else
#define ORC_PAR_FOR
#define ORC_PAR_SUM_S
#endif

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* glimmer_physcon.F90:64-75 (non-CESM values) */
static const double RHOI = 910.0, RHOO = 1028.0, GRAV = 9.81, SCYR = 31536000.0;
static const int GN = 3;
/* glissade_velo_higher.F90:1111 */
static const double VOL0 = 1.0e9;
/* glimmer_paramets.F90:146-147 */
static const double LEN_SCALE = 200.0e3;
#define VEL_SCALE (500.0 / SCYR)

#define IDX_A(iA, jA, kA) (1 + ((iA) + 1) + 3 * ((jA) + 1) + 9 * ((kA) + 1))

/* Fortran-layout accessors, 1-based logical indices, local nx, ny, nz in scope */
#define V3(a, k, i, j) (a)[((k)-1) + (size_t)nz * (((i)-1) + (size_t)(nx - 1) * ((j)-1))]
#define A4(a, m, k, i, j) \
    (a)[((m)-1) + (size_t)27 * (((k)-1) + (size_t)nz * (((i)-1) + (size_t)(nx - 1) * ((j)-1)))]
#define S2(a, i, j) (a)[((i)-1) + (size_t)(nx - 1) * ((j)-1)]
#define C2(a, i, j) (a)[((i)-1) + (size_t)nx * ((j)-1)]
#define F3(a, k, i, j) (a)[((k)-1) + (size_t)(nz - 1) * (((i)-1) + (size_t)nx * ((j)-1))]

/* ------------------------------------------------------------------------ */
/* tables: glissade_velo_higher_init (velo_higher:270-650)                  */
/* ------------------------------------------------------------------------ */

/* node n (0-based) position within the element: x/y offset 0|1, layer 0 = upper (k), 1 = lower (k+1) */
static const int NODE_IX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
static const int NODE_JY[8] = {0, 0, 1, 1, 0, 0, 1, 1};
static const int NODE_KL[8] = {1, 1, 1, 1, 0, 0, 0, 0};

static double T_phi[8][8], T_dxr[8][8], T_dyr[8][8], T_dzr[8][8]; /* [p][n] */
static int T_ready = 0;

static void build_tables(void)
{
    if (T_ready) return;
    const double r = 1.0 / sqrt(3.0);
    /* quadrature point p has reference coordinates (sx*r, sy*r, sz*r), velo_higher:306-328 */
    static const int QX[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    static const int QY[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    static const int QZ[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    for (int p = 0; p < 8; ++p) {
        double xq = QX[p] * r, yq = QY[p] * r, zq = QZ[p] * r;
        for (int n = 0; n < 8; ++n) {
            /* reference-element node signs: x,y from the horizontal position,
               z = -1 on the lower face (nodes 1-4), +1 on the upper face */
            double sx = NODE_IX[n] ? 1.0 : -1.0;
            double sy = NODE_JY[n] ? 1.0 : -1.0;
            double sz = NODE_KL[n] ? -1.0 : 1.0;
            double ax = sx > 0 ? 1.0 + xq : 1.0 - xq;
            double ay = sy > 0 ? 1.0 + yq : 1.0 - yq;
            double az = sz > 0 ? 1.0 + zq : 1.0 - zq;
            T_phi[p][n] = ax * ay * az / 8.0;
            T_dxr[p][n] = sx * (ay * az / 8.0);
            T_dyr[p][n] = sy * (ax * az / 8.0);
            T_dzr[p][n] = sz * (ax * ay / 8.0);
        }
    }
    T_ready = 1;
}

void orc_tables(double *phi, double *dxr, double *dyr, double *dzr,
                int *ishift, int *jshift, int *kshift, int *indxA)
{
    build_tables();
    /* Fortran layout (n, p) for the basis tables, (nr, nc) for the shifts */
    for (int p = 0; p < 8; ++p)
        for (int n = 0; n < 8; ++n) {
            if (phi) phi[n + 8 * p] = T_phi[p][n];
            if (dxr) dxr[n + 8 * p] = T_dxr[p][n];
            if (dyr) dyr[n + 8 * p] = T_dyr[p][n];
            if (dzr) dzr[n + 8 * p] = T_dzr[p][n];
        }
    for (int nc = 0; nc < 8; ++nc)
        for (int nr = 0; nr < 8; ++nr) {
            if (ishift) ishift[nr + 8 * nc] = NODE_IX[nc] - NODE_IX[nr];
            if (jshift) jshift[nr + 8 * nc] = NODE_JY[nc] - NODE_JY[nr];
            if (kshift) kshift[nr + 8 * nc] = NODE_KL[nc] - NODE_KL[nr];
        }
    if (indxA)
        for (int kA = -1; kA <= 1; ++kA)
            for (int jA = -1; jA <= 1; ++jA)
                for (int iA = -1; iA <= 1; ++iA)
                    indxA[(iA + 1) + 3 * (jA + 1) + 9 * (kA + 1)] = IDX_A(iA, jA, kA);
}

/* parallel_slap.F90:563-686 with nhalo = 2: staggered_lhalo = 2, uhalo = 1 */
void orc_grid_init(orc_grid *g, int nx, int ny, int nz, int periodic)
{
    g->nx = nx; g->ny = ny; g->nz = nz; g->nhalo = 2;
    g->periodic = periodic;
    g->ilo = periodic ? 3 : 2;
    g->jlo = periodic ? 3 : 2;
    g->ihi = nx - 1 - 1;
    g->jhi = ny - 1 - 1;
}

/* ------------------------------------------------------------------------ */
/* get_basis_function_derivatives_3d (velo_higher:6360-6579)               */
/* ------------------------------------------------------------------------ */
int orc_basis_derivatives(const double *x, const double *y, const double *z, int p,
                          double *dx, double *dy, double *dz, double *detJ)
{
    build_tables();
    const double *a = T_dxr[p], *b = T_dyr[p], *c = T_dzr[p];
    double J[3][3] = {{0}};
    for (int n = 0; n < 8; ++n) {
        J[0][0] = J[0][0] + a[n] * x[n];
        J[0][1] = J[0][1] + a[n] * y[n];
        J[0][2] = J[0][2] + a[n] * z[n];
        J[1][0] = J[1][0] + b[n] * x[n];
        J[1][1] = J[1][1] + b[n] * y[n];
        J[1][2] = J[1][2] + b[n] * z[n];
        J[2][0] = J[2][0] + c[n] * x[n];
        J[2][1] = J[2][1] + c[n] * y[n];
        J[2][2] = J[2][2] + c[n] * z[n];
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
    double det = J[0][0] * cof[0][0] + J[0][1] * cof[0][1] + J[0][2] * cof[0][2];
    *detJ = det;
    if (!(fabs(det) > 0.0)) return ORC_ERR_SINGULAR;
    double Ji[3][3];
    for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 3; ++row) Ji[row][col] = cof[col][row] / det;
    for (int n = 0; n < 8; ++n) {
        dx[n] = Ji[0][0] * a[n] + Ji[0][1] * b[n] + Ji[0][2] * c[n];
        dy[n] = Ji[1][0] * a[n] + Ji[1][1] * b[n] + Ji[1][2] * c[n];
        dz[n] = Ji[2][0] * a[n] + Ji[2][1] * b[n] + Ji[2][2] * c[n];
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* compute_effective_viscosity (velo_higher:7199-7378)                      */
/* ------------------------------------------------------------------------ */
double orc_efvs_qp(int whichefvs, int whichapprox, double efvs_constant,
                   const double *dphx, const double *dphy, const double *dphz,
                   const double *u, const double *v, double flwafact)
{
    const double p_effstr = (1.0 - (double)GN) / (double)GN;
    const double p2_effstr = p_effstr / 2;
    const double emin = 1.0e-8;
    if (whichefvs == 0) return efvs_constant;
    if (whichefvs == 1) {
        double effstrain = VEL_SCALE / LEN_SCALE * SCYR;
        return flwafact * pow(effstrain, p_effstr);
    }
    double ux = 0, uy = 0, uz = 0, vx = 0, vy = 0, vz = 0, sq;
    if (whichapprox == 0) {            /* SIA */
        for (int n = 0; n < 8; ++n) { uz = uz + dphz[n] * u[n]; vz = vz + dphz[n] * v[n]; }
        sq = emin * emin + 0.25 * (uz * uz + vz * vz);
    } else if (whichapprox == 1) {     /* SSA */
        for (int n = 0; n < 8; ++n) {
            ux = ux + dphx[n] * u[n]; uy = uy + dphy[n] * u[n];
            vx = vx + dphx[n] * v[n]; vy = vy + dphy[n] * v[n];
        }
        sq = emin * emin + (ux * ux + vy * vy + ux * vy + 0.25 * ((vx + uy) * (vx + uy)));
    } else {                           /* Blatter-Pattyn */
        for (int n = 0; n < 8; ++n) {
            ux = ux + dphx[n] * u[n]; uy = uy + dphy[n] * u[n]; uz = uz + dphz[n] * u[n];
            vx = vx + dphx[n] * v[n]; vy = vy + dphy[n] * v[n]; vz = vz + dphz[n] * v[n];
        }
        sq = emin * emin + (ux * ux + vy * vy + ux * vy + 0.25 * ((vx + uy) * (vx + uy)))
           + 0.25 * (uz * uz + vz * vz);
    }
    return flwafact * pow(sq, p2_effstr);
}

/* ------------------------------------------------------------------------ */
/* One element: node gather (velo_higher:5100-5127), quadrature loop with   */
/* compute_element_matrix (velo_higher:7868-7986), element efvs mean        */
/* (velo_higher:5167-5172).  K is Fortran (nr, nc) layout, 8x8.             */
/* ------------------------------------------------------------------------ */
int orc_element_matrix(const orc_grid *g, int i, int j, int k,
                       const double *sigma, const double *xV, const double *yV,
                       const double *uvel, const double *vvel,
                       const double *stagusrf, const double *stagthck,
                       const double *flwafact, int whichapprox, int whichefvs,
                       double efvs_constant,
                       double *Kuu, double *Kuv, double *Kvu, double *Kvv,
                       double *efvs_elem)
{
    const int nx = g->nx, nz = g->nz;
    double x[8], y[8], z[8], u[8], v[8];
    for (int n = 0; n < 8; ++n) {
        int in = i - 1 + NODE_IX[n], jn = j - 1 + NODE_JY[n], kn = k + NODE_KL[n];
        x[n] = S2(xV, in, jn);
        y[n] = S2(yV, in, jn);
        z[n] = S2(stagusrf, in, jn) - sigma[kn - 1] * S2(stagthck, in, jn);
        u[n] = V3(uvel, kn, in, jn);
        v[n] = V3(vvel, kn, in, jn);
    }
    for (int e = 0; e < 64; ++e) Kuu[e] = Kuv[e] = Kvu[e] = Kvv[e] = 0.0;
    double efvs_qp[8];
    const double ff = F3(flwafact, k, i, j);
    for (int p = 0; p < 8; ++p) {
        double dx[8], dy[8], dz[8], detJ;
        int st = orc_basis_derivatives(x, y, z, p, dx, dy, dz, &detJ);
        if (st) return st;
        double efvs = orc_efvs_qp(whichefvs, whichapprox, efvs_constant, dx, dy, dz, u, v, ff);
        efvs_qp[p] = efvs;
        double f = efvs * 1.0 * detJ / VOL0;   /* wqp = 1 */
        for (int nc = 0; nc < 8; ++nc)
            for (int nr = 0; nr < 8; ++nr) {
                int e = nr + 8 * nc;
                if (whichapprox == 0) {
                    Kuu[e] = Kuu[e] + f * (dz[nr] * dz[nc]);
                    Kvv[e] = Kvv[e] + f * (dz[nr] * dz[nc]);
                } else if (whichapprox == 1) {
                    Kuu[e] = Kuu[e] + f * (4.0 * dx[nr] * dx[nc] + dy[nr] * dy[nc]);
                    Kuv[e] = Kuv[e] + f * (2.0 * dx[nr] * dy[nc] + dy[nr] * dx[nc]);
                    Kvu[e] = Kvu[e] + f * (2.0 * dy[nr] * dx[nc] + dx[nr] * dy[nc]);
                    Kvv[e] = Kvv[e] + f * (4.0 * dy[nr] * dy[nc] + dx[nr] * dx[nc]);
                } else {
                    Kuu[e] = Kuu[e] + f * (4.0 * dx[nr] * dx[nc] + dy[nr] * dy[nc] + dz[nr] * dz[nc]);
                    Kuv[e] = Kuv[e] + f * (2.0 * dx[nr] * dy[nc] + dy[nr] * dx[nc]);
                    Kvu[e] = Kvu[e] + f * (2.0 * dy[nr] * dx[nc] + dx[nr] * dy[nc]);
                    Kvv[e] = Kvv[e] + f * (4.0 * dy[nr] * dy[nc] + dx[nr] * dx[nc] + dz[nr] * dz[nc]);
                }
            }
    }
    double s = 0.0;
    for (int p = 0; p < 8; ++p) s = s + efvs_qp[p];
    *efvs_elem = s / 8;
    /* check_symmetry_element_matrix (velo_higher:9468-9529) */
    const double eps10 = 1.0e-10;
    for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) {
            if (fabs(Kuu[a + 8 * b] - Kuu[b + 8 * a]) > eps10) return ORC_ERR_ASYMMETRIC;
            if (fabs(Kvv[a + 8 * b] - Kvv[b + 8 * a]) > eps10) return ORC_ERR_ASYMMETRIC;
            if (fabs(Kuv[a + 8 * b] - Kvu[b + 8 * a]) > eps10) return ORC_ERR_ASYMMETRIC;
        }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* assemble_stiffness_matrix_3d (velo_higher:4954-5204) with                */
/* element_to_global_matrix_3d (velo_higher:7990-8054)                      */
/* ------------------------------------------------------------------------ */
int orc_assemble_3d(const orc_grid *g, const double *sigma, const int32_t *active_cell,
                    const double *xV, const double *yV,
                    const double *uvel, const double *vvel,
                    const double *stagusrf, const double *stagthck,
                    const double *flwafact, int whichapprox,
                    double *efvs, int whichefvs, double efvs_constant,
                    double *Auu, double *Auv, double *Avu, double *Avv)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz, nh = g->nhalo;
    size_t na = (size_t)27 * nz * (nx - 1) * (ny - 1);
    memset(Auu, 0, na * sizeof(double));
    memset(Auv, 0, na * sizeof(double));
    memset(Avu, 0, na * sizeof(double));
    memset(Avv, 0, na * sizeof(double));
    memset(efvs, 0, (size_t)(nz - 1) * nx * ny * sizeof(double));
#ifdef ORC_OMP
    /* cell rows of one parity scatter to disjoint vertex rows (j-1, j) */
    int status = ORC_OK;
    for (int color = 0; color < 2; ++color) {
        ORC_PAR_FOR
        for (int j = nh + 1 + color; j <= ny - nh + 1; j += 2)
            for (int i = nh + 1; i <= nx - nh + 1; ++i) {
                if (!C2(active_cell, i, j)) continue;
                double Kuu[64], Kuv[64], Kvu[64], Kvv[64];
                for (int k = 1; k <= nz - 1; ++k) {
                    double ev;
                    int st = orc_element_matrix(g, i, j, k, sigma, xV, yV, uvel, vvel, stagusrf,
                                                stagthck, flwafact, whichapprox, whichefvs,
                                                efvs_constant, Kuu, Kuv, Kvu, Kvv, &ev);
                    if (st) { status = st; continue; }
                    F3(efvs, k, i, j) = ev;
                    for (int nr = 0; nr < 8; ++nr) {
                        int ir = i - 1 + NODE_IX[nr], jr = j - 1 + NODE_JY[nr], kr = k + NODE_KL[nr];
                        for (int nc = 0; nc < 8; ++nc) {
                            int m = IDX_A(NODE_IX[nc] - NODE_IX[nr], NODE_JY[nc] - NODE_JY[nr],
                                          NODE_KL[nc] - NODE_KL[nr]);
                            int e = nr + 8 * nc;
                            A4(Auu, m, kr, ir, jr) += Kuu[e];
                            A4(Auv, m, kr, ir, jr) += Kuv[e];
                            A4(Avu, m, kr, ir, jr) += Kvu[e];
                            A4(Avv, m, kr, ir, jr) += Kvv[e];
                        }
                    }
                }
            }
    }
    return status;
#endif
    double Kuu[64], Kuv[64], Kvu[64], Kvv[64];
    for (int j = nh + 1; j <= ny - nh + 1; ++j)
        for (int i = nh + 1; i <= nx - nh + 1; ++i) {
            if (!C2(active_cell, i, j)) continue;
            for (int k = 1; k <= nz - 1; ++k) {
                double ev;
                int st = orc_element_matrix(g, i, j, k, sigma, xV, yV, uvel, vvel, stagusrf,
                                            stagthck, flwafact, whichapprox, whichefvs,
                                            efvs_constant, Kuu, Kuv, Kvu, Kvv, &ev);
                if (st) return st;
                F3(efvs, k, i, j) = ev;
                for (int nr = 0; nr < 8; ++nr) {
                    int ir = i - 1 + NODE_IX[nr], jr = j - 1 + NODE_JY[nr], kr = k + NODE_KL[nr];
                    for (int nc = 0; nc < 8; ++nc) {
                        int m = IDX_A(NODE_IX[nc] - NODE_IX[nr], NODE_JY[nc] - NODE_JY[nr],
                                      NODE_KL[nc] - NODE_KL[nr]);
                        int e = nr + 8 * nc;
                        A4(Auu, m, kr, ir, jr) += Kuu[e];
                        A4(Auv, m, kr, ir, jr) += Kuv[e];
                        A4(Avu, m, kr, ir, jr) += Kvu[e];
                        A4(Avv, m, kr, ir, jr) += Kvv[e];
                    }
                }
            }
        }
    return ORC_OK;
}

/* basal_sliding_bc, HO_ASSEMBLE_BETA_LOCAL (velo_higher:8216-8232), basal layer k = nz */
void orc_basal_sliding_local(const orc_grid *g, double dx, double dy,
                             const int32_t *active_vertex, const double *beta,
                             double *Auu, double *Avv)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const int m = IDX_A(0, 0, 0);
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i)
            if (S2(active_vertex, i, j)) {
                A4(Auu, m, nz, i, j) = A4(Auu, m, nz, i, j) + dx * dy / VOL0 * S2(beta, i, j);
                A4(Avv, m, nz, i, j) = A4(Avv, m, nz, i, j) + dx * dy / VOL0 * S2(beta, i, j);
            }
}

/* basal_sliding_bc, HO_ASSEMBLE_BETA_STANDARD (velo_higher:8236-8363), on the basal layer:
 * for each active cell (j, i ascending) and each quadrature point p, the 2D Jacobian of the
 * cell's corners (get_basis_function_derivatives_2d, velo_higher:6583-6751), beta at p
 * interpolated from the corners, and Kuu(nr, nc) = beta_qp w_p detJ / vol0 phi_nr phi_nc
 * added into slot indxA(iA, jA, 0) of node nr (Kvv = Kuu). */
void orc_basal_sliding_standard(const orc_grid *g, const int32_t *active_cell,
                                const double *xV, const double *yV, const double *beta,
                                double *Auu, double *Avv)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz, h = g->nhalo;
    const double r = 1.0 / sqrt(3.0);
    const double xq[4] = {-r, r, r, -r}, yq[4] = {-r, -r, r, r}, wq[4] = {1.0, 1.0, 1.0, 1.0};
    /* corners SW, SE, NE, NW relative to the cell's NE vertex (ishift(3, n), jshift(3, n)) */
    const int di[4] = {-1, 0, 0, -1}, dj[4] = {-1, -1, 0, 0};
    for (int j = h + 1; j <= ny - h + 1; ++j)
        for (int i = h + 1; i <= nx - h + 1; ++i) {
            if (!C2(active_cell, i, j)) continue;
            double x[4], y[4], b[4];
            for (int n = 0; n < 4; ++n) {
                x[n] = S2(xV, i + di[n], j + dj[n]);
                y[n] = S2(yV, i + di[n], j + dj[n]);
                b[n] = S2(beta, i + di[n], j + dj[n]);
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
                double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
                for (int n = 0; n < 4; ++n) {
                    J11 = J11 + dxr[n] * x[n];
                    J12 = J12 + dxr[n] * y[n];
                    J21 = J21 + dyr[n] * x[n];
                    J22 = J22 + dyr[n] * y[n];
                }
                const double detJ = J11 * J22 - J12 * J21;
                double beta_qp = 0.0;
                for (int n = 0; n < 4; ++n) beta_qp = beta_qp + phi[n] * b[n];
                for (int nr = 0; nr < 4; ++nr) {
                    const int ii = i + di[nr], jj = j + dj[nr];
                    for (int nc = 0; nc < 4; ++nc) {
                        const int m = IDX_A(di[nc] - di[nr], dj[nc] - dj[nr], 0);
                        const double K = 0.0 + beta_qp * wq[p] * detJ / VOL0 * phi[nr] * phi[nc];
                        A4(Auu, m, nz, ii, jj) = A4(Auu, m, nz, ii, jj) + K;
                        A4(Avv, m, nz, ii, jj) = A4(Avv, m, nz, ii, jj) + K;
                    }
                }
            }
        }
}

/* dirichlet_boundary_conditions_3d (velo_higher:8385-8578) */
void orc_dirichlet_3d(const orc_grid *g, const int32_t *active_vertex,
                      const int32_t *umask, const int32_t *vmask,
                      const double *uvel, const double *vvel,
                      double *Auu, double *Auv, double *Avu, double *Avv,
                      double *bu, double *bv)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz, nh = g->nhalo;
    const int m0 = IDX_A(0, 0, 0);
    for (int j = nh; j <= ny - nh + 1; ++j)
        for (int i = nh; i <= nx - nh + 1; ++i) {
            if (!S2(active_vertex, i, j)) continue;
            for (int k = 1; k <= nz; ++k) {
                if (V3(umask, k, i, j) == 1) {
                    V3(bu, k, i, j) = V3(uvel, k, i, j);
                    for (int kA = -1; kA <= 1; ++kA)
                        for (int jA = -1; jA <= 1; ++jA)
                            for (int iA = -1; iA <= 1; ++iA) {
                                int kk = k + kA, ii = i + iA, jj = j + jA;
                                if (kk < 1 || kk > nz || ii < 1 || ii > nx - 1 || jj < 1 || jj > ny - 1)
                                    continue;
                                if (iA == 0 && jA == 0 && kA == 0) {
                                    A4(Auu, m0, k, i, j) = 1.0;
                                    A4(Auv, m0, k, i, j) = 0.0;
                                    A4(Avu, m0, k, i, j) = 0.0;
                                } else {
                                    int m = IDX_A(iA, jA, kA);
                                    A4(Auu, m, k, i, j) = 0.0;
                                    A4(Auv, m, k, i, j) = 0.0;
                                    int mm = IDX_A(-iA, -jA, -kA);
                                    if (V3(umask, kk, ii, jj) != 1) {
                                        V3(bu, kk, ii, jj) = V3(bu, kk, ii, jj) - A4(Auu, mm, kk, ii, jj) * V3(uvel, k, i, j);
                                        A4(Auu, mm, kk, ii, jj) = 0.0;
                                    }
                                    if (V3(vmask, kk, ii, jj) != 1) {
                                        V3(bv, kk, ii, jj) = V3(bv, kk, ii, jj) - A4(Avu, mm, kk, ii, jj) * V3(uvel, k, i, j);
                                        A4(Avu, mm, kk, ii, jj) = 0.0;
                                    }
                                }
                            }
                }
                if (V3(vmask, k, i, j) == 1) {
                    V3(bv, k, i, j) = V3(vvel, k, i, j);
                    for (int kA = -1; kA <= 1; ++kA)
                        for (int jA = -1; jA <= 1; ++jA)
                            for (int iA = -1; iA <= 1; ++iA) {
                                int kk = k + kA, ii = i + iA, jj = j + jA;
                                if (kk < 1 || kk > nz || ii < 1 || ii > nx - 1 || jj < 1 || jj > ny - 1)
                                    continue;
                                if (iA == 0 && jA == 0 && kA == 0) {
                                    A4(Auv, m0, k, i, j) = 0.0;
                                    A4(Avu, m0, k, i, j) = 0.0;
                                    A4(Avv, m0, k, i, j) = 1.0;
                                } else {
                                    int m = IDX_A(iA, jA, kA);
                                    A4(Avu, m, k, i, j) = 0.0;
                                    A4(Avv, m, k, i, j) = 0.0;
                                    int mm = IDX_A(-iA, -jA, -kA);
                                    if (V3(umask, kk, ii, jj) != 1) {
                                        V3(bu, kk, ii, jj) = V3(bu, kk, ii, jj) - A4(Auv, mm, kk, ii, jj) * V3(vvel, k, i, j);
                                        A4(Auv, mm, kk, ii, jj) = 0.0;
                                    }
                                    if (V3(vmask, kk, ii, jj) != 1) {
                                        V3(bv, kk, ii, jj) = V3(bv, kk, ii, jj) - A4(Avv, mm, kk, ii, jj) * V3(vvel, k, i, j);
                                        A4(Avv, mm, kk, ii, jj) = 0.0;
                                    }
                                }
                            }
                }
            }
        }
}

/* ------------------------------------------------------------------------ */
/* staggered_parallel_halo_real8_3d/_4d (parallel_slap.F90:2807-2925).      */
/* The array has shape (lead, nx-1, ny-1); lead = nz or 27*nz.              */
/* ------------------------------------------------------------------------ */
void orc_halo(const orc_grid *g, int lead, double *a)
{
    const int sx = g->nx - 1, sy = g->ny - 1;   /* staggered extents */
    const int lh = 2, uh = 1;
#define H3(i, j) (a + (size_t)lead * (((i)-1) + (size_t)sx * ((j)-1)))
    if (g->periodic) {
        /* west/east strips over interior rows j = lh+1 .. sy-uh (copies taken before writes) */
        for (int j = lh + 1; j <= sy - uh; ++j) {
            double tmpw[1], *dummy = tmpw; (void)dummy;
            /* east halo (i = sx) <- owned column lh+1; west halo (1..lh) <- sx-uh-lh+1 .. sx-uh */
            for (int q = 0; q < uh; ++q)
                memcpy(H3(sx - uh + 1 + q, j), H3(lh + 1 + q, j), sizeof(double) * lead);
            for (int q = 0; q < lh; ++q)
                memcpy(H3(1 + q, j), H3(sx - uh - lh + 1 + q, j), sizeof(double) * lead);
        }
        /* north/south full rows */
        for (int q = 0; q < uh; ++q)
            memcpy(H3(1, sy - uh + 1 + q), H3(1, lh + 1 + q), sizeof(double) * lead * sx);
        for (int q = 0; q < lh; ++q)
            memcpy(H3(1, 1 + q), H3(1, sy - uh - lh + 1 + q), sizeof(double) * lead * sx);
    } else {
        for (int j = 1; j <= sy; ++j) {
            for (int i = 1; i <= lh - 1; ++i) memset(H3(i, j), 0, sizeof(double) * lead);
            for (int i = sx - uh + 1; i <= sx; ++i) memset(H3(i, j), 0, sizeof(double) * lead);
        }
        for (int j = 1; j <= lh - 1; ++j) memset(H3(1, j), 0, sizeof(double) * lead * sx);
        for (int j = sy - uh + 1; j <= sy; ++j) memset(H3(1, j), 0, sizeof(double) * lead * sx);
    }
#undef H3
}

/* check_symmetry_assembled_matrix_3d (velo_higher:9533-9741) */
int orc_check_symmetry_3d(const orc_grid *g, const int32_t *active_vertex,
                          double *Auu, double *Auv, double *Avu, double *Avv,
                          double *maxdiff_out)
{
    const int nx = g->nx, nz = g->nz;
    const double eps08 = 1.0e-8;
    const int m0 = IDX_A(0, 0, 0);
    double maxdiff = 0.0;
    int nwarn = 0;
    for (int j = g->jlo; j <= g->jhi; ++j)
        for (int i = g->ilo; i <= g->ihi; ++i) {
            if (!S2(active_vertex, i, j)) continue;
            for (int k = 1; k <= nz; ++k) {
                for (int pass = 0; pass < 2; ++pass) {
                    /* pass 0: Auu/Auv rows (diag from Auu); pass 1: Avv/Avu rows (diag from Avv) */
                    double *Ad = pass == 0 ? Auu : Avv;
                    double *Ao = pass == 0 ? Auv : Avu;   /* off-diagonal block of this row      */
                    double *At = pass == 0 ? Avu : Auv;   /* its transpose partner               */
                    double diag = A4(Ad, m0, k, i, j);
                    for (int jA = -1; jA <= 1; ++jA)
                        for (int iA = -1; iA <= 1; ++iA)
                            for (int kA = -1; kA <= 1; ++kA) {
                                if (k + kA < 1 || k + kA > nz) continue;
                                int m = IDX_A(iA, jA, kA), mm = IDX_A(-iA, -jA, -kA);
                                double v1 = A4(Ad, m, k, i, j), v2 = A4(Ad, mm, k + kA, i + iA, j + jA);
                                if (v2 != v1) {
                                    if (fabs(v2 - v1) > maxdiff) maxdiff = fabs(v2 - v1);
                                    if (fabs(v2 - v1) < eps08 * fabs(diag)) {
                                        double avg = 0.5 * (v1 + v2);
                                        A4(Ad, m, k, i, j) = avg;
                                        A4(Ad, mm, k + kA, i + iA, j + jA) = avg;
                                    } else nwarn++;
                                }
                                v1 = A4(Ao, m, k, i, j); v2 = A4(At, mm, k + kA, i + iA, j + jA);
                                if (pass == 1 && fabs(v2 - v1) > maxdiff) maxdiff = fabs(v2 - v1);
                                if (v2 != v1) {
                                    if (pass == 0 && fabs(v2 - v1) > maxdiff) maxdiff = fabs(v2 - v1);
                                    if (fabs(v2 - v1) < eps08 * fabs(diag)) {
                                        double avg = 0.5 * (v1 + v2);
                                        A4(Ao, m, k, i, j) = avg;
                                        A4(At, mm, k + kA, i + iA, j + jA) = avg;
                                    } else nwarn++;
                                }
                            }
                }
            }
        }
    if (maxdiff_out) *maxdiff_out = maxdiff;
    return nwarn;
}

/* count_nonzeros_3d (velo_higher:9355-9410) */
long orc_count_nonzeros_3d(const orc_grid *g, const int32_t *active_vertex,
                           const double *Auu, const double *Auv,
                           const double *Avu, const double *Avv)
{
    const int nx = g->nx, nz = g->nz;
    long n = 0;
    for (int j = g->jlo; j <= g->jhi; ++j)
        for (int i = g->ilo; i <= g->ihi; ++i) {
            if (!S2(active_vertex, i, j)) continue;
            for (int k = 1; k <= nz; ++k)
                for (int m = 1; m <= 27; ++m) {
                    n += A4(Auu, m, k, i, j) != 0.0;
                    n += A4(Auv, m, k, i, j) != 0.0;
                    n += A4(Avu, m, k, i, j) != 0.0;
                    n += A4(Avv, m, k, i, j) != 0.0;
                }
        }
    return n;
}

/* the 27-point row product shared by matvec and the residual (pcg:2607-2645) */
static inline void row_product(int nx, int ny, int nz, int k, int i, int j,
                               const double *Auu, const double *Auv,
                               const double *Avu, const double *Avv,
                               const double *xu, const double *xv, double *yu, double *yv)
{
    double su = *yu, sv = *yv;
    for (int kA = -1; kA <= 1; ++kA)
        for (int jA = -1; jA <= 1; ++jA)
            for (int iA = -1; iA <= 1; ++iA) {
                int kk = k + kA, ii = i + iA, jj = j + jA;
                if (kk < 1 || kk > nz || ii < 1 || ii > nx - 1 || jj < 1 || jj > ny - 1) continue;
                int m = IDX_A(iA, jA, kA);
                su = su + A4(Auu, m, k, i, j) * V3(xu, kk, ii, jj) + A4(Auv, m, k, i, j) * V3(xv, kk, ii, jj);
                sv = sv + A4(Avu, m, k, i, j) * V3(xu, kk, ii, jj) + A4(Avv, m, k, i, j) * V3(xv, kk, ii, jj);
            }
    *yu = su; *yv = sv;
}

/* compute_residual_vector_3d (velo_higher:8768-8927) */
void orc_residual_vector_3d(const orc_grid *g, const int32_t *active_vertex,
                            const double *Auu, const double *Auv,
                            const double *Avu, const double *Avv,
                            const double *bu, const double *bv,
                            const double *uvel, const double *vvel,
                            double *ru, double *rv,
                            double *L2_norm, double *L2_norm_relative)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    size_t nv = (size_t)nz * (nx - 1) * (ny - 1);
    memset(ru, 0, nv * sizeof(double));
    memset(rv, 0, nv * sizeof(double));
    for (int j = g->jlo; j <= g->jhi; ++j)
        for (int i = g->ilo; i <= g->ihi; ++i) {
            if (!S2(active_vertex, i, j)) continue;
            for (int k = 1; k <= nz; ++k)
                row_product(nx, ny, nz, k, i, j, Auu, Auv, Avu, Avv, uvel, vvel,
                            &V3(ru, k, i, j), &V3(rv, k, i, j));
        }
    double L2 = 0.0;
    for (int j = g->jlo; j <= g->jhi; ++j)
        for (int i = g->ilo; i <= g->ihi; ++i) {
            if (!S2(active_vertex, i, j)) continue;
            for (int k = 1; k <= nz; ++k) {
                V3(ru, k, i, j) = V3(ru, k, i, j) - V3(bu, k, i, j);
                V3(rv, k, i, j) = V3(rv, k, i, j) - V3(bv, k, i, j);
                L2 = L2 + V3(ru, k, i, j) * V3(ru, k, i, j) + V3(rv, k, i, j) * V3(rv, k, i, j);
            }
        }
    L2 = sqrt(L2);
    *L2_norm = L2;
    if (L2_norm_relative) {
        double rhs = 0.0;
        for (int j = g->jlo; j <= g->jhi; ++j)
            for (int i = g->ilo; i <= g->ihi; ++i) {
                if (!S2(active_vertex, i, j)) continue;
                for (int k = 1; k <= nz; ++k)
                    rhs = rhs + V3(bu, k, i, j) * V3(bu, k, i, j) + V3(bv, k, i, j) * V3(bv, k, i, j);
            }
        rhs = sqrt(rhs);
        *L2_norm_relative = rhs > 0.0 ? L2 / rhs : 0.0;
    }
}

/* compute_residual_velocity_3d (velo_higher:9111-9234); MEANU (2) is not parallel-safe
 * in the reference and is restated the same way */
double orc_residual_velocity_3d(const orc_grid *g, int whichresid,
                                const double *uvel, const double *vvel,
                                const double *usav, const double *vsav)
{
    const int nx = g->nx, nz = g->nz;
    double r = 0.0;
    long count = 0;
    int kmax = (whichresid == 1 || whichresid == 2) ? nz - 1 : nz;
    for (int j = g->jlo; j <= g->jhi; ++j)
        for (int i = g->ilo; i <= g->ihi; ++i)
            for (int k = 1; k <= kmax; ++k) {
                double u = V3(uvel, k, i, j), v = V3(vvel, k, i, j);
                double speed = sqrt(u * u + v * v);
                if (speed != 0.0) {
                    double uo = V3(usav, k, i, j), vo = V3(vsav, k, i, j);
                    double old = sqrt(uo * uo + vo * vo);
                    double d = fabs((old - speed) / speed);
                    if (whichresid == 2) { count++; r = r + d; }
                    else if (d > r) r = d;
                }
            }
    if (whichresid == 2 && count > 0) r = r / count;
    return r;
}

/* load_vector_gravity, HO_ASSEMBLE_TAUD_LOCAL (velo_higher:4397-4428) */
void orc_load_gravity_local(const orc_grid *g, const double *stagwbndsigma,
                            double dx, double dy, const int32_t *active_vertex,
                            const double *stagthck, const double *dusrf_dx,
                            const double *dusrf_dy, double *loadu, double *loadv)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            if (!S2(active_vertex, i, j)) continue;
            for (int k = 1; k <= nz; ++k) {
                double dz = S2(stagthck, i, j) * (stagwbndsigma[k] - stagwbndsigma[k - 1]);
                V3(loadu, k, i, j) = V3(loadu, k, i, j) - RHOI * GRAV * dx * dy * dz / VOL0 * S2(dusrf_dx, i, j);
                V3(loadv, k, i, j) = V3(loadv, k, i, j) - RHOI * GRAV * dx * dy * dz / VOL0 * S2(dusrf_dy, i, j);
            }
        }
}

/* load_vector_gravity, standard FE assembly (velo_higher:4436-4528) */
void orc_load_gravity_standard(const orc_grid *g, const double *sigma, const int32_t *active_cell,
                               const double *xV, const double *yV,
                               const double *stagusrf, const double *stagthck,
                               const double *dusrf_dx, const double *dusrf_dy,
                               double *loadu, double *loadv)
{
    build_tables();
    const int nx = g->nx, ny = g->ny, nz = g->nz, nh = g->nhalo;
    for (int j = nh + 1; j <= ny - nh + 1; ++j)
        for (int i = nh + 1; i <= nx - nh + 1; ++i) {
            if (!C2(active_cell, i, j)) continue;
            for (int k = 1; k <= nz - 1; ++k) {
                double x[8], y[8], z[8], sx[8], sy[8];
                for (int n = 0; n < 8; ++n) {
                    int in = i - 1 + NODE_IX[n], jn = j - 1 + NODE_JY[n], kn = k + NODE_KL[n];
                    x[n] = S2(xV, in, jn); y[n] = S2(yV, in, jn);
                    z[n] = S2(stagusrf, in, jn) - sigma[kn - 1] * S2(stagthck, in, jn);
                    sx[n] = S2(dusrf_dx, in, jn); sy[n] = S2(dusrf_dy, in, jn);
                }
                for (int p = 0; p < 8; ++p) {
                    double ddx[8], ddy[8], ddz[8], detJ;
                    orc_basis_derivatives(x, y, z, p, ddx, ddy, ddz, &detJ);
                    double qx = 0.0, qy = 0.0;
                    for (int n = 0; n < 8; ++n) { qx = qx + T_phi[p][n] * sx[n]; qy = qy + T_phi[p][n] * sy[n]; }
                    for (int n = 0; n < 8; ++n) {
                        int in = i - 1 + NODE_IX[n], jn = j - 1 + NODE_JY[n], kn = k + NODE_KL[n];
                        V3(loadu, kn, in, jn) = V3(loadu, kn, in, jn) - RHOI * GRAV * 1.0 * detJ / VOL0 * qx * T_phi[p][n];
                        V3(loadv, kn, in, jn) = V3(loadv, kn, in, jn) - RHOI * GRAV * 1.0 * detJ / VOL0 * qy * T_phi[p][n];
                    }
                }
            }
        }
}


/* ------------------------------------------------------------------------ */
/* load_vector_lateral_bc + lateral_shelf_bc (velo_higher:4536-4950): net    */
/* hydrostatic pressure on vertical ice faces that border ice-free ocean,    */
/* integrated with 2x2 Gauss points over each layer of the face.            */
/* ------------------------------------------------------------------------ */
static void shelf_face(const orc_grid *g, const double *sigma, int face, int ic, int jc,
                       const double *susrf, const double *sthck, const double *xV, const double *yV,
                       double *loadu, double *loadv)
{
    const int nx = g->nx, nz = g->nz;
    int in[4], jn[4];
    double x[4], y[4], s[4], h[4];
    /* face 0 west, 1 east, 2 south, 3 north (nodes 1,2 along the face, 3,4 above them) */
    if (face == 0) { in[0] = ic - 1; jn[0] = jc;     in[1] = ic - 1; jn[1] = jc - 1; }
    else if (face == 1) { in[0] = ic; jn[0] = jc - 1; in[1] = ic;     jn[1] = jc; }
    else if (face == 2) { in[0] = ic - 1; jn[0] = jc - 1; in[1] = ic; jn[1] = jc - 1; }
    else { in[0] = ic; jn[0] = jc; in[1] = ic - 1; jn[1] = jc; }
    for (int n = 0; n < 2; ++n)
        x[n] = face <= 1 ? S2(yV, in[n], jn[n]) : S2(xV, in[n], jn[n]);
    in[2] = in[1]; jn[2] = jn[1]; in[3] = in[0]; jn[3] = jn[0];
    x[2] = x[1]; x[3] = x[0];
    s[0] = S2(susrf, in[0], jn[0]); s[1] = S2(susrf, in[1], jn[1]); s[2] = s[1]; s[3] = s[0];
    h[0] = S2(sthck, in[0], jn[0]); h[1] = S2(sthck, in[1], jn[1]); h[2] = h[1]; h[3] = h[0];
    const double r = 1.0 / sqrt(3.0);
    static const int SX[4] = {-1, 1, 1, -1}, SY[4] = {-1, -1, 1, 1};
    for (int k = 1; k <= nz - 1; ++k) {
        y[0] = s[0] - sigma[k] * h[0];
        y[1] = s[1] - sigma[k] * h[1];
        y[2] = s[2] - sigma[k - 1] * h[2];
        y[3] = s[3] - sigma[k - 1] * h[3];
        const int kn[4] = {k + 1, k + 1, k, k};
        for (int p = 0; p < 4; ++p) {
            const double xq = SX[p] * r, yq = SY[p] * r;
            double phi[4], dxr[4], dyr[4];
            for (int n = 0; n < 4; ++n) {
                const double ax = SX[n] > 0 ? 1.0 + xq : 1.0 - xq;
                const double ay = SY[n] > 0 ? 1.0 + yq : 1.0 - yq;
                phi[n] = ax * ay / 4.0;
                dxr[n] = SX[n] * (ay / 4.0);
                dyr[n] = SY[n] * (ax / 4.0);
            }
            double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
            for (int n = 0; n < 4; ++n) {
                J11 = J11 + dxr[n] * x[n];
                J12 = J12 + dxr[n] * y[n];
                J21 = J21 + dyr[n] * x[n];
                J22 = J22 + dyr[n] * y[n];
            }
            const double detJ = fabs(J11 * J22 - J12 * J21);
            double hq = 0.0, sq = 0.0;
            for (int n = 0; n < 4; ++n) { hq = hq + phi[n] * h[n]; sq = sq + phi[n] * s[n]; }
            const double ratio = sq / hq < 1.0 ? sq / hq : 1.0;
            const double p_av = 0.5 * RHOI * GRAV * hq - 0.5 * RHOO * GRAV * hq * ((1.0 - ratio) * (1.0 - ratio));
            for (int n = 0; n < 4; ++n) {
                const double t = p_av * 1.0 * detJ / VOL0 * phi[n];
                if (face == 0) V3(loadu, kn[n], in[n], jn[n]) = V3(loadu, kn[n], in[n], jn[n]) - t;
                else if (face == 1) V3(loadu, kn[n], in[n], jn[n]) = V3(loadu, kn[n], in[n], jn[n]) + t;
                else if (face == 2) V3(loadv, kn[n], in[n], jn[n]) = V3(loadv, kn[n], in[n], jn[n]) - t;
                else V3(loadv, kn[n], in[n], jn[n]) = V3(loadv, kn[n], in[n], jn[n]) + t;
            }
        }
    }
}

void orc_load_lateral_bc(const orc_grid *g, const double *sigma, const int32_t *ocean_mask,
                         const int32_t *calving_front_mask, const int32_t *active_cell,
                         const double *xV, const double *yV, const double *stagusrf,
                         const double *stagthck, double *loadu, double *loadv)
{
    const int nx = g->nx, ny = g->ny, nh = g->nhalo;
    for (int j = nh + 1; j <= ny - nh + 1; ++j)
        for (int i = nh + 1; i <= nx - nh + 1; ++i) {
            if (!C2(active_cell, i, j)) continue;
            const int ni[4] = {i - 1, i + 1, i, i}, nj[4] = {j, j, j - 1, j + 1};
            for (int f = 0; f < 4; ++f)
                if (C2(ocean_mask, ni[f], nj[f]) == 1 ||
                    (C2(calving_front_mask, ni[f], nj[f]) == 1 && !C2(active_cell, ni[f], nj[f])))
                    shelf_face(g, sigma, f, i, j, stagusrf, stagthck, xV, yV, loadu, loadv);
        }
}

/* matvec_multiply_structured_3d (pcg:2533-2647) */
void orc_matvec_3d(const orc_grid *g, const int32_t *active_vertex,
                   const double *Auu, const double *Auv,
                   const double *Avu, const double *Avv,
                   const double *xu, const double *xv, double *yu, double *yv)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    size_t nv = (size_t)nz * (nx - 1) * (ny - 1);
    memset(yu, 0, nv * sizeof(double));
    memset(yv, 0, nv * sizeof(double));
    ORC_PAR_FOR
    for (int j = g->jlo; j <= g->jhi; ++j)
        for (int i = g->ilo; i <= g->ihi; ++i) {
            if (!S2(active_vertex, i, j)) continue;
            for (int k = 1; k <= nz; ++k)
                row_product(nx, ny, nz, k, i, j, Auu, Auv, Avu, Avv, xu, xv,
                            &V3(yu, k, i, j), &V3(yv, k, i, j));
        }
}

/* global_sum_staggered_3d_real8 (pcg:2333-2379): owned vertices, all k, w1 + w2 */
static double gsum2(const orc_grid *g, const double *w1, const double *w2)
{
    const int nx = g->nx, nz = g->nz;
    double s = 0.0;
    ORC_PAR_SUM_S
    for (int j = g->jlo; j <= g->jhi; ++j)
        for (int i = g->ilo; i <= g->ihi; ++i)
            for (int k = 1; k <= nz; ++k) s = s + V3(w1, k, i, j) + V3(w2, k, i, j);
    return s;
}

/* tridiag_solver (pcg:2852-2904), Thomas algorithm without pivoting */
void orc_tridiag(int n, const double *sb, const double *d, const double *sp,
                 const double *rhs, double *soln)
{
    double gam[64];
    double beta = d[0];
    soln[0] = rhs[0] / beta;
    for (int k = 1; k < n; ++k) {
        gam[k] = sp[k - 1] / beta;
        beta = d[k] - sb[k] * gam[k];
        soln[k] = (rhs[k] - sb[k] * soln[k - 1]) / beta;
    }
    for (int k = n - 2; k >= 0; --k) soln[k] = soln[k] - gam[k + 1] * soln[k + 1];
}

/* setup_preconditioner_3d (pcg:2150-2265) + the precond branches of the PCG loops
 * (pcg:300-365) + easy_sia_solver (pcg:2759-2848).  M arrays are (-1:1, nz, nx-1, ny-1). */
typedef struct { double *Adu, *Adv, *Muu, *Mvv; } orc_pc;

static void pc_setup(const orc_grid *g, int precond, const double *Auu, const double *Avv, orc_pc *pc)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    size_t nv = (size_t)nz * (nx - 1) * (ny - 1);
    pc->Adu = calloc(nv, sizeof(double)); pc->Adv = calloc(nv, sizeof(double));
    pc->Muu = calloc(3 * nv, sizeof(double)); pc->Mvv = calloc(3 * nv, sizeof(double));
    for (size_t q = 0; q < nv; ++q) {
        if (precond == 1) {
            pc->Adu[q] = Auu[27 * q + IDX_A(0, 0, 0) - 1];
            pc->Adv[q] = Avv[27 * q + IDX_A(0, 0, 0) - 1];
        } else if (precond == 2) {
            for (int o = -1; o <= 1; ++o) {
                pc->Muu[3 * q + o + 1] = Auu[27 * q + IDX_A(0, 0, o) - 1];
                pc->Mvv[3 * q + o + 1] = Avv[27 * q + IDX_A(0, 0, o) - 1];
            }
        }
    }
}

static void pc_free(orc_pc *pc) { free(pc->Adu); free(pc->Adv); free(pc->Muu); free(pc->Mvv); }

static void sia_solve(const orc_grid *g, const int32_t *active_vertex, const double *M,
                      const double *r, double *z)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    ORC_PAR_FOR
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            double sb[64], d[64], sp[64], rhs[64], soln[64];
            if (!S2(active_vertex, i, j)) continue;
            size_t base = (size_t)nz * ((i - 1) + (size_t)(nx - 1) * (j - 1));
            for (int k = 0; k < nz; ++k) {
                rhs[k] = r[base + k];
                sb[k] = (k == 0) ? 0.0 : M[3 * (base + k) + 0];
                d[k] = M[3 * (base + k) + 1];
                sp[k] = (k == nz - 1) ? 0.0 : M[3 * (base + k) + 2];
            }
            orc_tridiag(nz, sb, d, sp, rhs, soln);
            for (int k = 0; k < nz; ++k) z[base + k] = soln[k];
        }
}

static void pc_apply(const orc_grid *g, int precond, const int32_t *active_vertex, const orc_pc *pc,
                     const double *ru, const double *rv, double *zu, double *zv)
{
    size_t nv = (size_t)g->nz * (g->nx - 1) * (g->ny - 1);
    if (precond == 0) {
        memcpy(zu, ru, nv * sizeof(double));
        memcpy(zv, rv, nv * sizeof(double));
    } else if (precond == 1) {
        for (size_t q = 0; q < nv; ++q) {
            zu[q] = pc->Adu[q] != 0.0 ? ru[q] / pc->Adu[q] : 0.0;
            zv[q] = pc->Adv[q] != 0.0 ? rv[q] / pc->Adv[q] : 0.0;
        }
    } else if (precond == 2) {
        sia_solve(g, active_vertex, pc->Muu, ru, zu);
        sia_solve(g, active_vertex, pc->Mvv, rv, zv);
    }
}

void orc_precond_apply(const orc_grid *g, const int32_t *active_vertex, int precond,
                       const double *Auu, const double *Avv, const double *ru, const double *rv,
                       double *zu, double *zv)
{
    orc_pc pc;
    pc_setup(g, precond, Auu, Avv, &pc);
    pc_apply(g, precond, active_vertex, &pc, ru, rv, zu, zv);
    pc_free(&pc);
}

/* pcg_solver_standard_3d (pcg:84-549) */
int orc_pcg_standard_3d(const orc_grid *g, const int32_t *av,
                        const double *Auu, const double *Auv,
                        const double *Avu, const double *Avv,
                        const double *bu, const double *bv,
                        double *xu, double *xv, int precond,
                        int maxiters, double tolerance,
                        double *err_out, int *niters_out)
{
    size_t nv = (size_t)g->nz * (g->nx - 1) * (g->ny - 1);
    double *ru = calloc(nv, 8), *rv = calloc(nv, 8), *du = calloc(nv, 8), *dv = calloc(nv, 8);
    double *qu = calloc(nv, 8), *qv = calloc(nv, 8), *zu = calloc(nv, 8), *zv = calloc(nv, 8);
    double *wu = calloc(nv, 8), *wv = calloc(nv, 8);
    orc_pc pc;
    pc_setup(g, precond, Auu, Avv, &pc);
    int status = ORC_OK;
    double err = 0.0;
    orc_halo(g, g->nz, xu); orc_halo(g, g->nz, xv);
    orc_matvec_3d(g, av, Auu, Auv, Avu, Avv, xu, xv, zu, zv);
    for (size_t q = 0; q < nv; ++q) { ru[q] = bu[q] - zu[q]; rv[q] = bv[q] - zv[q]; }
    int niters = maxiters;
    double eta0 = 1.0, eta1, eta2, alpha, beta;
    memset(du, 0, nv * 8); memset(dv, 0, nv * 8);
    memset(zu, 0, nv * 8); memset(zv, 0, nv * 8);
    for (size_t q = 0; q < nv; ++q) { wu[q] = bu[q] * bu[q]; wv[q] = bv[q] * bv[q]; }
    double L2_rhs = sqrt(gsum2(g, wu, wv));
    for (int n = 1; n <= maxiters; ++n) {
        pc_apply(g, precond, av, &pc, ru, rv, zu, zv);
        for (size_t q = 0; q < nv; ++q) { wu[q] = ru[q] * zu[q]; wv[q] = rv[q] * zv[q]; }
        eta1 = gsum2(g, wu, wv);
        if (eta1 != eta1) { status = ORC_ERR_NAN; break; }
        beta = eta1 / eta0;
        for (size_t q = 0; q < nv; ++q) { du[q] = zu[q] + beta * du[q]; dv[q] = zv[q] + beta * dv[q]; }
        orc_halo(g, g->nz, du); orc_halo(g, g->nz, dv);
        orc_matvec_3d(g, av, Auu, Auv, Avu, Avv, du, dv, qu, qv);
        eta0 = eta1;
        for (size_t q = 0; q < nv; ++q) { wu[q] = du[q] * qu[q]; wv[q] = dv[q] * qv[q]; }
        eta2 = gsum2(g, wu, wv);
        alpha = eta1 / eta2;
        if (alpha != alpha) { status = ORC_ERR_NAN; break; }
        for (size_t q = 0; q < nv; ++q) {
            xu[q] = xu[q] + alpha * du[q]; xv[q] = xv[q] + alpha * dv[q];
            ru[q] = ru[q] - alpha * qu[q]; rv[q] = rv[q] - alpha * qv[q];
        }
        if (n % 5 == 0) {
            orc_halo(g, g->nz, xu); orc_halo(g, g->nz, xv);
            orc_matvec_3d(g, av, Auu, Auv, Avu, Avv, xu, xv, zu, zv);
            for (size_t q = 0; q < nv; ++q) { ru[q] = bu[q] - zu[q]; rv[q] = bv[q] - zv[q]; }
            for (size_t q = 0; q < nv; ++q) { wu[q] = ru[q] * ru[q]; wv[q] = rv[q] * rv[q]; }
            err = sqrt(gsum2(g, wu, wv)) / L2_rhs;
            if (err < tolerance) { niters = n; break; }
        }
    }
    *err_out = err;
    *niters_out = niters;
    pc_free(&pc);
    free(ru); free(rv); free(du); free(dv); free(qu); free(qv); free(zu); free(zv); free(wu); free(wv);
    return status;
}

/* pcg_solver_chrongear_3d (pcg:965-1610) */
int orc_pcg_chrongear_3d(const orc_grid *g, const int32_t *av,
                         const double *Auu, const double *Auv,
                         const double *Avu, const double *Avv,
                         const double *bu, const double *bv,
                         double *xu, double *xv, int precond,
                         int maxiters, double tolerance,
                         double *err_out, int *niters_out)
{
    const int nz = g->nz;
    size_t nv = (size_t)nz * (g->nx - 1) * (g->ny - 1);
    double *ru = calloc(nv, 8), *rv = calloc(nv, 8), *du = calloc(nv, 8), *dv = calloc(nv, 8);
    double *zu = calloc(nv, 8), *zv = calloc(nv, 8), *qu = calloc(nv, 8), *qv = calloc(nv, 8);
    double *Azu = calloc(nv, 8), *Azv = calloc(nv, 8), *wu = calloc(nv, 8), *wv = calloc(nv, 8);
    double *w2u = calloc(nv, 8), *w2v = calloc(nv, 8);
    orc_pc pc;
    pc_setup(g, precond, Auu, Avv, &pc);
    int status = ORC_OK, niters = maxiters;
    double err = 0.0;
    ORC_PAR_FOR
    for (size_t q = 0; q < nv; ++q) { wu[q] = bu[q] * bu[q]; wv[q] = bv[q] * bv[q]; }
    double L2_rhs = sqrt(gsum2(g, wu, wv));
    orc_halo(g, nz, xu); orc_halo(g, nz, xv);
    orc_matvec_3d(g, av, Auu, Auv, Avu, Avv, xu, xv, zu, zv);
    ORC_PAR_FOR
    for (size_t q = 0; q < nv; ++q) { ru[q] = bu[q] - zu[q]; rv[q] = bv[q] - zv[q]; }
    orc_halo(g, nz, ru); orc_halo(g, nz, rv);
    pc_apply(g, precond, av, &pc, ru, rv, zu, zv);
    ORC_PAR_FOR
    for (size_t q = 0; q < nv; ++q) { wu[q] = ru[q] * zu[q]; wv[q] = rv[q] * zv[q]; }
    memcpy(du, zu, nv * 8); memcpy(dv, zv, nv * 8);
    orc_matvec_3d(g, av, Auu, Auv, Avu, Avv, du, dv, qu, qv);
    ORC_PAR_FOR
    for (size_t q = 0; q < nv; ++q) { w2u[q] = du[q] * qu[q]; w2v[q] = dv[q] * qv[q]; }
    double rho_old = gsum2(g, wu, wv);
    double sigma = gsum2(g, w2u, w2v);
    orc_halo(g, nz, qu); orc_halo(g, nz, qv);
    double alpha = rho_old / sigma;
    if (alpha != alpha) status = ORC_ERR_NAN;
    if (!status) {
        ORC_PAR_FOR
    for (size_t q = 0; q < nv; ++q) {
            xu[q] = xu[q] + alpha * du[q]; xv[q] = xv[q] + alpha * dv[q];
            ru[q] = ru[q] - alpha * qu[q]; rv[q] = rv[q] - alpha * qv[q];
        }
        for (int n = 2; n <= maxiters; ++n) {
            pc_apply(g, precond, av, &pc, ru, rv, zu, zv);
            orc_matvec_3d(g, av, Auu, Auv, Avu, Avv, zu, zv, Azu, Azv);
            ORC_PAR_FOR
    for (size_t q = 0; q < nv; ++q) {
                wu[q] = ru[q] * zu[q]; wv[q] = rv[q] * zv[q];
                w2u[q] = Azu[q] * zu[q]; w2v[q] = Azv[q] * zv[q];
            }
            double rho = gsum2(g, wu, wv), delta = gsum2(g, w2u, w2v);
            orc_halo(g, nz, Azu); orc_halo(g, nz, Azv);
            double beta = rho / rho_old;
            sigma = delta - beta * rho / alpha;
            alpha = rho / sigma;
            rho_old = rho;
            if (alpha != alpha) { status = ORC_ERR_NAN; break; }
            ORC_PAR_FOR
    for (size_t q = 0; q < nv; ++q) {
                du[q] = zu[q] + beta * du[q]; dv[q] = zv[q] + beta * dv[q];
                qu[q] = Azu[q] + beta * qu[q]; qv[q] = Azv[q] + beta * qv[q];
                xu[q] = xu[q] + alpha * du[q]; xv[q] = xv[q] + alpha * dv[q];
                ru[q] = ru[q] - alpha * qu[q]; rv[q] = rv[q] - alpha * qv[q];
            }
            if (n % 5 == 0) {
                orc_matvec_3d(g, av, Auu, Auv, Avu, Avv, xu, xv, zu, zv);
                ORC_PAR_FOR
    for (size_t q = 0; q < nv; ++q) { ru[q] = bu[q] - zu[q]; rv[q] = bv[q] - zv[q]; }
                ORC_PAR_FOR
    for (size_t q = 0; q < nv; ++q) { wu[q] = ru[q] * ru[q]; wv[q] = rv[q] * rv[q]; }
                err = sqrt(gsum2(g, wu, wv)) / L2_rhs;
                if (err < tolerance) { niters = n; break; }
                orc_halo(g, nz, ru); orc_halo(g, nz, rv);
            }
        }
    }
    *err_out = err;
    *niters_out = niters;
    pc_free(&pc);
    free(ru); free(rv); free(du); free(dv); free(zu); free(zv); free(qu); free(qv);
    free(Azu); free(Azv); free(wu); free(wv); free(w2u); free(w2v);
    return status;
}

/* ------------------------------------------------------------------------ */
/* Picard loop, 3D BP branch of glissade_velo_higher_solve                  */
/* (velo_higher:2003-2016, 2206-3727)                                       */
/* ------------------------------------------------------------------------ */
int orc_picard_solve(const orc_grid *g, orc_picard_opts *o,
                     const double *sigma, const int32_t *active_cell,
                     const int32_t *active_vertex,
                     const double *xV, const double *yV,
                     const double *stagusrf, const double *stagthck,
                     const double *flwafact, const double *beta_internal,
                     const int32_t *umask, const int32_t *vmask,
                     const double *loadu, const double *loadv,
                     double *uvel, double *vvel, double *efvs,
                     double *bu, double *bv, double *resid_u, double *resid_v)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    size_t nv = (size_t)nz * (nx - 1) * (ny - 1), na = 27 * nv;
    double *Auu = malloc(na * 8), *Auv = malloc(na * 8), *Avu = malloc(na * 8), *Avv = malloc(na * 8);
    double *usav = malloc(nv * 8), *vsav = malloc(nv * 8);
    const double L2_target = 1.0e-4, L2_target_rel = 1.0e-7, resid_target = 1.0e-4;
    double crit = 1.0e10, target = 1.0e-12;
    double L2 = 1.0e20, L2rel = 1.0e20, resid_velo = 1.0;
    int counter = 0, status = ORC_OK;
    const size_t ncol = (size_t)(nx - 1) * (ny - 1);
    double *beta_work = NULL, *ubas = NULL, *vbas = NULL;
    if (o->bp) { beta_work = malloc(ncol * 8); ubas = malloc(ncol * 8); vbas = malloc(ncol * 8); }
    o->total_linear_iters = 0;
    while (crit >= target && counter < o->maxiter_nonlinear) {
        counter++;
        memcpy(usav, uvel, nv * 8); memcpy(vsav, vvel, nv * 8);
        if (o->bp) {
            for (int j = 1; j <= ny - 1; ++j)
                for (int i = 1; i <= nx - 1; ++i) {
                    S2(ubas, i, j) = V3(uvel, nz, i, j);
                    S2(vbas, i, j) = V3(vvel, nz, i, j);
                }
            status = orc_calcbeta(g, o->bp, ubas, vbas, o->beta_external, o->effecpress_stag,
                                  o->f_ground, beta_work);
            if (status) break;
            beta_internal = beta_work;
        }
        status = orc_assemble_3d(g, sigma, active_cell, xV, yV, uvel, vvel, stagusrf, stagthck,
                                 flwafact, o->whichapprox, efvs, o->whichefvs, o->efvs_constant,
                                 Auu, Auv, Avu, Avv);
        if (status) break;
        if (o->whichbabc != 6) {   /* HO_BABC_NO_SLIP */
            if (o->assemble_beta_standard)
                orc_basal_sliding_standard(g, active_cell, xV, yV, beta_internal, Auu, Avv);
            else
                orc_basal_sliding_local(g, o->dx, o->dy, active_vertex, beta_internal, Auu, Avv);
        }
        memcpy(bu, loadu, nv * 8); memcpy(bv, loadv, nv * 8);
        orc_dirichlet_3d(g, active_vertex, umask, vmask, uvel, vvel, Auu, Auv, Avu, Avv, bu, bv);
        orc_halo(g, 27 * nz, Auu); orc_halo(g, 27 * nz, Auv);
        orc_halo(g, 27 * nz, Avu); orc_halo(g, 27 * nz, Avv);
        orc_halo(g, nz, bu); orc_halo(g, nz, bv);
        orc_check_symmetry_3d(g, active_vertex, Auu, Auv, Avu, Avv, NULL);
        long nnz = orc_count_nonzeros_3d(g, active_vertex, Auu, Auv, Avu, Avv);
        if (nnz == 0) {
            memset(resid_u, 0, nv * 8); memset(resid_v, 0, nv * 8);
            memset(bu, 0, nv * 8); memset(bv, 0, nv * 8);
            memset(uvel, 0, nv * 8); memset(vvel, 0, nv * 8);
            L2 = 0.0;
            break;
        }
        orc_residual_vector_3d(g, active_vertex, Auu, Auv, Avu, Avv, bu, bv, uvel, vvel,
                               resid_u, resid_v, &L2, &L2rel);
        if (o->L2_history && counter <= o->L2_history_cap) o->L2_history[counter - 1] = L2;
        double err; int niters;
        if (o->whichsparse == 3)
            status = orc_pcg_chrongear_3d(g, active_vertex, Auu, Auv, Avu, Avv, bu, bv, uvel, vvel,
                                          o->whichprecond, o->linear_maxiters, o->linear_tolerance,
                                          &err, &niters);
        else
            status = orc_pcg_standard_3d(g, active_vertex, Auu, Auv, Avu, Avv, bu, bv, uvel, vvel,
                                         o->whichprecond, o->linear_maxiters, o->linear_tolerance,
                                         &err, &niters);
        if (status) break;
        o->total_linear_iters += niters;
        orc_halo(g, nz, uvel); orc_halo(g, nz, vvel);
        resid_velo = orc_residual_velocity_3d(g, o->whichresid, uvel, vvel, usav, vsav);
        if (o->whichresid == 3) { crit = L2; target = L2_target; }
        else if (o->whichresid == 4) { crit = L2rel; target = L2_target_rel; }
        else { crit = resid_velo; target = resid_target; }
        if (o->target_override > 0) target = o->target_override;
    }
    o->converged = counter < o->maxiter_nonlinear;
    o->picard_iters = counter;
    o->final_L2_norm = L2;
    o->final_resid_velo = resid_velo;
    o->status = status;
    free(Auu); free(Auv); free(Avu); free(Avv); free(usav); free(vsav);
    free(beta_work); free(ubas); free(vbas);
    return status;
}

/* ------------------------------------------------------------------------ */
/* Post-solve diagnostics                                                   */
/* ------------------------------------------------------------------------ */

/* parallel_halo for a cell field (nx, ny), nhalo = 2 (parallel_slap.F90:1794-1860):
 * periodic: west halo <- the last two owned columns, east halo <- the first two,
 * then the same for full rows; outflow: halo cells zero. */
void orc_cell_halo(const orc_grid *g, double *a)
{
    const int nx = g->nx, ny = g->ny, h = g->nhalo;
    if (!g->periodic) {
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                if (i <= h || i > nx - h || j <= h || j > ny - h) C2(a, i, j) = 0.0;
        return;
    }
    for (int j = h + 1; j <= ny - h; ++j)
        for (int t = 1; t <= h; ++t) {
            const double e = C2(a, nx - 2 * h + t, j), w = C2(a, h + t, j);
            C2(a, t, j) = e;
            C2(a, nx - h + t, j) = w;
        }
    for (int t = 1; t <= h; ++t)
        for (int i = 1; i <= nx; ++i) {
            const double n = C2(a, i, ny - 2 * h + t), s = C2(a, i, h + t);
            C2(a, i, t) = n;
            C2(a, i, ny - h + t) = s;
        }
}

/* compute_internal_stress (velo_higher:6968-7195): element-mean stress components
 * from the quadrature points; efvs is used for L1L2 only (whichapprox 3). */
int orc_internal_stress(const orc_grid *g, const double *sigma, const int32_t *active_cell,
                        const double *xV, const double *yV, const double *stagusrf,
                        const double *stagthck, const double *flwafact, const double *efvs,
                        int whichefvs, double efvs_constant, int whichapprox,
                        const double *uvel, const double *vvel,
                        double *tau_xz, double *tau_yz, double *tau_xx, double *tau_yy,
                        double *tau_xy, double *tau_eff)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz, h = g->nhalo;
    const size_t ne = (size_t)(nz - 1) * nx * ny;
    memset(tau_xz, 0, ne * sizeof(double));
    memset(tau_yz, 0, ne * sizeof(double));
    memset(tau_xx, 0, ne * sizeof(double));
    memset(tau_yy, 0, ne * sizeof(double));
    memset(tau_xy, 0, ne * sizeof(double));
    memset(tau_eff, 0, ne * sizeof(double));
    for (int j = h + 1; j <= ny - h + 1; ++j)
        for (int i = h + 1; i <= nx - h + 1; ++i) {
            if (!C2(active_cell, i, j)) continue;
            for (int k = 1; k <= nz - 1; ++k) {
                double x[8], y[8], z[8], u[8], v[8];
                for (int n = 0; n < 8; ++n) {
                    const int in = i - 1 + NODE_IX[n], jn = j - 1 + NODE_JY[n], kn = k + NODE_KL[n];
                    x[n] = S2(xV, in, jn);
                    y[n] = S2(yV, in, jn);
                    z[n] = S2(stagusrf, in, jn) - sigma[kn - 1] * S2(stagthck, in, jn);
                    u[n] = V3(uvel, kn, in, jn);
                    v[n] = V3(vvel, kn, in, jn);
                }
                double sxz = 0, syz = 0, sxx = 0, syy = 0, sxy = 0;
                for (int p = 0; p < 8; ++p) {
                    double dx[8], dy[8], dz[8], detJ;
                    const int st = orc_basis_derivatives(x, y, z, p, dx, dy, dz, &detJ);
                    if (st) return st;
                    double ux = 0, uy = 0, uz = 0, vx = 0, vy = 0, vz = 0;
                    if (whichapprox == 0) {
                        for (int n = 0; n < 8; ++n) { uz = uz + dz[n] * u[n]; vz = vz + dz[n] * v[n]; }
                    } else if (whichapprox == 1) {
                        for (int n = 0; n < 8; ++n) {
                            ux = ux + dx[n] * u[n]; uy = uy + dy[n] * u[n];
                            vx = vx + dx[n] * v[n]; vy = vy + dy[n] * v[n];
                        }
                    } else {
                        for (int n = 0; n < 8; ++n) {
                            ux = ux + dx[n] * u[n]; uy = uy + dy[n] * u[n]; uz = uz + dz[n] * u[n];
                            vx = vx + dx[n] * v[n]; vy = vy + dy[n] * v[n]; vz = vz + dz[n] * v[n];
                        }
                    }
                    const double ev = whichapprox == 3
                        ? F3(efvs, k, i, j)
                        : orc_efvs_qp(whichefvs, whichapprox, efvs_constant, dx, dy, dz, u, v,
                                      F3(flwafact, k, i, j));
                    sxz = sxz + ev * uz;
                    syz = syz + ev * vz;
                    sxx = sxx + 2.0 * ev * ux;
                    syy = syy + 2.0 * ev * vy;
                    sxy = sxy + ev * (vx + uy);
                }
                sxz = sxz / 8; syz = syz / 8; sxx = sxx / 8; syy = syy / 8; sxy = sxy / 8;
                F3(tau_xz, k, i, j) = sxz;
                F3(tau_yz, k, i, j) = syz;
                F3(tau_xx, k, i, j) = sxx;
                F3(tau_yy, k, i, j) = syy;
                F3(tau_xy, k, i, j) = sxy;
                F3(tau_eff, k, i, j) = sqrt(sxx * sxx + syy * syy + sxx * syy + sxy * sxy
                                            + sxz * sxz + syz * syz);
            }
        }
    return ORC_OK;
}

/* compute_basal_friction_heatflx (velo_higher:6755-6964).  ubas, vbas, beta are
 * vertex fields; out: cell field (W/m^2) with its halo filled.  whichassemble_bfric
 * 0: 2x2-Gauss finite-element mean over each active cell; 1 (default): vertex
 * values averaged to cell centres (glissade_unstagger, glissade_grid_operators.F90:152-231). */
void orc_basal_friction_heatflx(const orc_grid *g, const int32_t *active_cell,
                                const int32_t *active_vertex, const double *ubas,
                                const double *vbas, const double *beta, int whichassemble_bfric,
                                double *bfricflx)
{
    const int nx = g->nx, ny = g->ny, h = g->nhalo;
    memset(bfricflx, 0, (size_t)nx * ny * sizeof(double));
    if (whichassemble_bfric == 0) {
        const double r = 1.0 / sqrt(3.0);
        const double xq[4] = {-r, r, r, -r}, yq[4] = {-r, -r, r, r};
        /* nodes counter-clockwise from the SW vertex (i-1, j-1) */
        const int di[4] = {-1, 0, 0, -1}, dj[4] = {-1, -1, 0, 0};
        for (int j = h + 1; j <= ny - h; ++j)
            for (int i = h + 1; i <= nx - h; ++i) {
                if (!C2(active_cell, i, j)) continue;
                double u[4], v[4], b[4];
                for (int n = 0; n < 4; ++n) {
                    u[n] = S2(ubas, i + di[n], j + dj[n]);
                    v[n] = S2(vbas, i + di[n], j + dj[n]);
                    b[n] = S2(beta, i + di[n], j + dj[n]);
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
                C2(bfricflx, i, j) = acc / (sw * SCYR);
            }
    } else {
        double *st = (double *)calloc((size_t)(nx - 1) * (ny - 1), sizeof(double));
        for (int j = 1; j <= ny - 1; ++j)
            for (int i = 1; i <= nx - 1; ++i)
                if (S2(active_vertex, i, j)) {
                    double q = S2(beta, i, j) * (S2(ubas, i, j) * S2(ubas, i, j)
                                                 + S2(vbas, i, j) * S2(vbas, i, j));
                    S2(st, i, j) = q / SCYR;
                }
        for (int j = 2; j <= ny - 1; ++j)
            for (int i = 2; i <= nx - 1; ++i)
                C2(bfricflx, i, j) = (S2(st, i, j) + S2(st, i - 1, j) + S2(st, i, j - 1)
                                      + S2(st, i - 1, j - 1)) / 4.0;
        free(st);
    }
    orc_cell_halo(g, bfricflx);
}

/* basal traction (velo_higher:3893-3894): beta * velocity at the bed, vertex fields */
void orc_basal_traction(const orc_grid *g, const double *beta, const double *uvel,
                        const double *vvel, double *btractx, double *btracty)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            S2(btractx, i, j) = S2(beta, i, j) * V3(uvel, nz, i, j);
            S2(btracty, i, j) = S2(beta, i, j) * V3(vvel, nz, i, j);
        }
}

/* glissade_grounded_fraction (glissade_grounding_line.F90:55-178), options
 * HO_GROUND_NO_GLP (0) and HO_GROUND_ALL (2); the GLP (1) is fatal in the
 * reference and ORC_ERR_ARG here.  The vertex mask tests land_mask(i+1,j+1) twice
 * and never land_mask(i+1,j), as the reference's condition does (:112-115). */
int orc_grounded_fraction(int nx, int ny, const int32_t *ice, const int32_t *floating,
                          const int32_t *land, int whichground, double *f_ground)
{
    if (whichground != 0 && whichground != 2) return ORC_ERR_ARG;
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            const int vm = C2(ice, i, j + 1) == 1 || C2(ice, i + 1, j + 1) == 1 || C2(ice, i, j) == 1 ||
                           C2(ice, i + 1, j) == 1 || C2(land, i, j + 1) == 1 || C2(land, i + 1, j + 1) == 1 ||
                           C2(land, i, j) == 1 || C2(land, i + 1, j + 1) == 1;
            double fg = 0.0;
            if (vm) {
                if (whichground == 2) {
                    fg = 1.0;
                } else {
#define CG(ii, jj) ((C2(ice, ii, jj) == 1 && C2(floating, ii, jj) == 0) || C2(land, ii, jj) == 1)
                    fg = (CG(i, j + 1) || CG(i + 1, j + 1) || CG(i, j) || CG(i + 1, j)) ? 1.0 : 0.0;
#undef CG
                }
            }
            S2(f_ground, i, j) = fg;
        }
    return ORC_OK;
}

/* calcbeta (glissade_basal_traction.F90:67-598) for the options that need no
 * thermal or hydrology state: 0 constant, 4 large, 5 external, 9 power law,
 * 11 Coulomb-power law (Schoof), 12 Coulomb-power law (Tsai), 14 simple; then
 * beta *= f_ground, the beta_grounded_min floor where f_ground > 0, the check that
 * every beta is >= 0 (ORC_ERR_NAN otherwise) and the staggered halo.  ubas, vbas:
 * the basal velocity on vertices; effecpress_stag: N on vertices (options 11, 12);
 * g: the solve's grid (nx, ny cells). */
int orc_calcbeta(const orc_grid *g, const orc_basal_physics *bp, const double *ubas,
                 const double *vbas, const double *beta_external, const double *effecpress_stag,
                 const double *f_ground, double *beta)
{
    const int nx = g->nx, ny = g->ny;
    const double smallnum = 1.0e-2;
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            const double u = S2(ubas, i, j), v = S2(vbas, i, j);
            double speed = sqrt(u * u + v * v + smallnum * smallnum);
            if (bp->beta_powerlaw_umax > 0.0) speed = fmin(speed, bp->beta_powerlaw_umax);
            double b;
            switch (bp->whichbabc) {
            case 0: b = bp->ho_beta_const; break;
            case 4: b = bp->ho_beta_large; break;
            case 5: b = S2(beta_external, i, j); break;
            case 9: b = bp->powerlaw_c * pow(speed, 1.0 / bp->powerlaw_m - 1.0); break;
            case 11: {
                const double m = bp->powerlaw_m, N = S2(effecpress_stag, i, j);
                const double num = bp->powerlaw_c * bp->coulomb_c * N;
                const double den = pow(pow(bp->powerlaw_c, m) * speed + pow(bp->coulomb_c * N, m), 1.0 / m);
                b = (num / den) * pow(speed, 1.0 / m - 1.0);
                if (b > 1.0e8) b = 1.0e8;
                break;
            }
            case 12: {
                const double tp = bp->powerlaw_c * pow(speed, 1.0 / bp->powerlaw_m);
                const double tc = bp->coulomb_c * S2(effecpress_stag, i, j);
                b = tc <= tp ? tc / speed : tp / speed;
                break;
            }
            case 14: b = (j >= 5 && j <= ny - 5) ? 100.0 : 1.0e4; break;
            default: return ORC_ERR_ARG;
            }
            S2(beta, i, j) = b;
        }
    if (bp->whichbabc == 5) {
        double mx = -1.0e300;
        for (int j = 1; j <= ny - 1; ++j)
            for (int i = 1; i <= nx - 1; ++i) mx = fmax(mx, S2(beta, i, j));
        if (mx < 0.0) return ORC_ERR_ARG;
    }
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            double b = S2(beta, i, j) * S2(f_ground, i, j);
            if (S2(f_ground, i, j) > 0.0 && b < bp->beta_grounded_min) b = bp->beta_grounded_min;
            if (!(b >= 0.0)) return ORC_ERR_NAN;
            S2(beta, i, j) = b;
        }
    orc_halo(g, 1, beta);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* 2D (SSA) solve.  The 2D system lives in a one-level grid (nz = 1): its     */
/* nine couplings (iA, jA) occupy the 27-slot layout's kA = 0 slots          */
/* IDX_A(iA, jA, 0) (indxA_2d(iA,jA) + 9), so the 3D Dirichlet, halo,          */
/* symmetry, nnz, residual and PCG routines above are exactly the 2D ones      */
/* (pcg:553-961, 1614-2146, 2651-2755; velo_higher:8582-8764, 8931-9107).      */
/* ------------------------------------------------------------------------ */

/* glissade_vertical_average without a mask (glissade_grid_operators.F90:1172-1236) */
void orc_vertical_average(int nx, int ny, int nz, const double *sigma, const double *var,
                          double *var_2d)
{
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            double s = 0.0;
            for (int k = 1; k <= nz - 1; ++k)
                s = s + F3(var, k, i, j) * (sigma[k] - sigma[k - 1]);
            C2(var_2d, i, j) = s;
        }
}

/* assemble_stiffness_matrix_2d, SSA branch (velo_higher:5208-5595) with
 * get_basis_function_derivatives_2d (:6583-6751), compute_effective_viscosity on
 * the four element nodes (:7199-7378), compute_element_matrix SSA (:7868-7986) and
 * element_to_global_matrix_2d (:8058-8121).  g: the one-level grid; nz3, sigma,
 * flwafact (nz3-1, nx, ny) and efvs (nz3-1, nx, ny, out) belong to the 3D grid. */
int orc_assemble_2d_ssa(const orc_grid *g, int nz3, const double *sigma,
                        const int32_t *active_cell, const double *xV, const double *yV,
                        const double *uvel, const double *vvel, const double *stagthck,
                        const double *flwafact, int whichefvs, double efvs_constant,
                        double *efvs, double *Auu, double *Auv, double *Avu, double *Avv)
{
    const int nx = g->nx, ny = g->ny, h = g->nhalo;
    const int nz = 1;   /* for the V3 / A4 accessors of the one-level system */
    const size_t na = 27 * (size_t)(nx - 1) * (ny - 1);
    memset(Auu, 0, na * 8); memset(Auv, 0, na * 8); memset(Avu, 0, na * 8); memset(Avv, 0, na * 8);
    memset(efvs, 0, (size_t)(nz3 - 1) * nx * ny * 8);
    double *ff2 = (double *)malloc((size_t)nx * ny * 8);
    orc_vertical_average(nx, ny, nz3, sigma, flwafact, ff2);
    const double r = 1.0 / sqrt(3.0);
    const double xq[4] = {-r, r, r, -r}, yq[4] = {-r, -r, r, r};
    /* lower-face nodes SW, SE, NE, NW relative to the element's NE vertex (ishift(3,n)) */
    const int di[4] = {-1, 0, 0, -1}, dj[4] = {-1, -1, 0, 0};
    const double p_eff = (1.0 - (double)GN) / (double)GN, p2 = p_eff / 2, emin = 1.0e-8;
    int status = ORC_OK;
    for (int j = h + 1; j <= ny - h + 1 && !status; ++j)
        for (int i = h + 1; i <= nx - h + 1 && !status; ++i) {
            if (!C2(active_cell, i, j)) continue;
            double x[4], y[4], u[4], v[4], hn[4];
            for (int n = 0; n < 4; ++n) {
                x[n] = S2(xV, i + di[n], j + dj[n]);
                y[n] = S2(yV, i + di[n], j + dj[n]);
                u[n] = S2(uvel, i + di[n], j + dj[n]);
                v[n] = S2(vvel, i + di[n], j + dj[n]);
                hn[n] = S2(stagthck, i + di[n], j + dj[n]);
            }
            double Kuu[16] = {0}, Kuv[16] = {0}, Kvu[16] = {0}, Kvv[16] = {0}, evq[4];
            for (int p = 0; p < 4; ++p) {
                double dxr[4], dyr[4], phi[4];
                dxr[0] = -(1.0 - yq[p]) / 4.0; dxr[1] = (1.0 - yq[p]) / 4.0;
                dxr[2] = (1.0 + yq[p]) / 4.0;  dxr[3] = -(1.0 + yq[p]) / 4.0;
                dyr[0] = -(1.0 - xq[p]) / 4.0; dyr[1] = -(1.0 + xq[p]) / 4.0;
                dyr[2] = (1.0 + xq[p]) / 4.0;  dyr[3] = (1.0 - xq[p]) / 4.0;
                phi[0] = (1.0 - xq[p]) * (1.0 - yq[p]) / 4.0;
                phi[1] = (1.0 + xq[p]) * (1.0 - yq[p]) / 4.0;
                phi[2] = (1.0 + xq[p]) * (1.0 + yq[p]) / 4.0;
                phi[3] = (1.0 - xq[p]) * (1.0 + yq[p]) / 4.0;
                double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
                for (int n = 0; n < 4; ++n) {
                    J11 = J11 + dxr[n] * x[n];
                    J12 = J12 + dxr[n] * y[n];
                    J21 = J21 + dyr[n] * x[n];
                    J22 = J22 + dyr[n] * y[n];
                }
                const double det = J11 * J22 - J12 * J21;
                if (!(fabs(det) > 0.0)) { status = ORC_ERR_SINGULAR; break; }
                const double i11 = J22 / det, i12 = -J12 / det, i21 = -J21 / det, i22 = J11 / det;
                double dx[4], dy[4];
                for (int n = 0; n < 4; ++n) {
                    dx[n] = i11 * dxr[n] + i12 * dyr[n];
                    dy[n] = i21 * dxr[n] + i22 * dyr[n];
                }
                double ev;
                const double ff = C2(ff2, i, j);
                if (whichefvs == 0) {
                    ev = efvs_constant;
                } else if (whichefvs == 1) {
                    ev = ff * pow(VEL_SCALE / LEN_SCALE * SCYR, p_eff);
                } else {
                    double ux = 0, uy = 0, vx = 0, vy = 0;
                    for (int n = 0; n < 4; ++n) {
                        ux = ux + dx[n] * u[n]; uy = uy + dy[n] * u[n];
                        vx = vx + dx[n] * v[n]; vy = vy + dy[n] * v[n];
                    }
                    const double sq = emin * emin + (ux * ux + vy * vy + ux * vy + 0.25 * ((vx + uy) * (vx + uy)));
                    ev = ff * pow(sq, p2);
                }
                evq[p] = ev;
                double hq = 0.0;
                for (int n = 0; n < 4; ++n) hq = hq + phi[n] * hn[n];
                const double f = hq * ev * 1.0 * det / VOL0;
                for (int nc = 0; nc < 4; ++nc)
                    for (int nr = 0; nr < 4; ++nr) {
                        const int e = nr + 4 * nc;
                        Kuu[e] = Kuu[e] + f * (4.0 * dx[nr] * dx[nc] + dy[nr] * dy[nc]);
                        Kuv[e] = Kuv[e] + f * (2.0 * dx[nr] * dy[nc] + dy[nr] * dx[nc]);
                        Kvu[e] = Kvu[e] + f * (2.0 * dy[nr] * dx[nc] + dx[nr] * dy[nc]);
                        Kvv[e] = Kvv[e] + f * (4.0 * dy[nr] * dy[nc] + dx[nr] * dx[nc]);
                    }
            }
            if (status) break;
            for (int k = 1; k <= nz3 - 1; ++k) {
                double s = 0.0;
                for (int p = 0; p < 4; ++p) s = s + evq[p];
                efvs[(k - 1) + (size_t)(nz3 - 1) * ((i - 1) + (size_t)nx * (j - 1))] = s / 4;
            }
            for (int a = 0; a < 4 && !status; ++a)
                for (int b = 0; b < 4; ++b)
                    if (fabs(Kuu[a + 4 * b] - Kuu[b + 4 * a]) > 1.0e-10 ||
                        fabs(Kvv[a + 4 * b] - Kvv[b + 4 * a]) > 1.0e-10 ||
                        fabs(Kuv[a + 4 * b] - Kvu[b + 4 * a]) > 1.0e-10) {
                        status = ORC_ERR_ASYMMETRIC;
                        break;
                    }
            for (int nr = 0; nr < 4; ++nr) {
                const int iv = i + di[nr], jv = j + dj[nr];
                for (int nc = 0; nc < 4; ++nc) {
                    const int iA = di[nc] - di[nr], jA = dj[nc] - dj[nr];
                    const int m = IDX_A(iA, jA, 0);
                    const int e = nr + 4 * nc;
                    A4(Auu, m, 1, iv, jv) = A4(Auu, m, 1, iv, jv) + Kuu[e];
                    A4(Auv, m, 1, iv, jv) = A4(Auv, m, 1, iv, jv) + Kuv[e];
                    A4(Avu, m, 1, iv, jv) = A4(Avu, m, 1, iv, jv) + Kvu[e];
                    A4(Avv, m, 1, iv, jv) = A4(Avv, m, 1, iv, jv) + Kvv[e];
                }
            }
        }
    free(ff2);
    return status;
}

/* The SSA branch of the Picard loop (velo_higher:2206-3767 with solve_2d): the 2D
 * velocity starts from the bed level, the 2D load is the column sum of the 3D load,
 * the Dirichlet masks are the bed level's, and every level of the 3D velocity takes
 * the 2D solution at the end.  g: the 3D grid; outputs bu, bv, resid_u, resid_v are
 * one-level ((nx-1) x (ny-1)). */
int orc_picard_solve_2d(const orc_grid *g, orc_picard_opts *o,
                        const double *sigma, const int32_t *active_cell,
                        const int32_t *active_vertex,
                        const double *xV, const double *yV,
                        const double *stagusrf, const double *stagthck,
                        const double *flwafact, const double *beta_internal,
                        const int32_t *umask, const int32_t *vmask,
                        const double *loadu, const double *loadv,
                        double *uvel, double *vvel, double *efvs,
                        double *bu, double *bv, double *resid_u, double *resid_v)
{
    (void)stagusrf;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    orc_grid g2;
    orc_grid_init(&g2, nx, ny, 1, g->periodic);
    const size_t nv = (size_t)(nx - 1) * (ny - 1), na = 27 * nv;
    double *Auu = malloc(na * 8), *Auv = malloc(na * 8), *Avu = malloc(na * 8), *Avv = malloc(na * 8);
    double *u2 = malloc(nv * 8), *v2 = malloc(nv * 8), *us = malloc(nv * 8), *vs = malloc(nv * 8);
    double *l2u = calloc(nv, 8), *l2v = calloc(nv, 8);
    int32_t *um2 = malloc(nv * 4), *vm2 = malloc(nv * 4);
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            const size_t c = (i - 1) + (size_t)(nx - 1) * (j - 1);
            u2[c] = V3(uvel, nz, i, j);
            v2[c] = V3(vvel, nz, i, j);
            um2[c] = V3(umask, nz, i, j);
            vm2[c] = V3(vmask, nz, i, j);
            for (int k = 1; k <= nz; ++k) {
                l2u[c] = l2u[c] + V3(loadu, k, i, j);
                l2v[c] = l2v[c] + V3(loadv, k, i, j);
            }
        }
    const double L2_target = 1.0e-4, L2_target_rel = 1.0e-7, resid_target = 1.0e-4;
    double crit = 1.0e10, target = 1.0e-12;
    double L2 = 1.0e20, L2rel = 1.0e20, resid_velo = 1.0;
    int counter = 0, status = ORC_OK;
    double *beta_work = o->bp ? malloc(nv * 8) : NULL;
    o->total_linear_iters = 0;
    while (crit >= target && counter < o->maxiter_nonlinear) {
        counter++;
        memcpy(us, u2, nv * 8); memcpy(vs, v2, nv * 8);
        if (o->bp) {   /* the basal velocity of the 2D solve is the 2D velocity (:2241-2243) */
            status = orc_calcbeta(g, o->bp, u2, v2, o->beta_external, o->effecpress_stag, o->f_ground,
                                  beta_work);
            if (status) break;
            beta_internal = beta_work;
        }
        status = orc_assemble_2d_ssa(&g2, nz, sigma, active_cell, xV, yV, u2, v2, stagthck, flwafact,
                                     o->whichefvs, o->efvs_constant, efvs, Auu, Auv, Avu, Avv);
        if (status) break;
        /* the 2D branch applies the basal friction for every which_ho_babc (:2832-2852) */
        if (o->assemble_beta_standard)
            orc_basal_sliding_standard(&g2, active_cell, xV, yV, beta_internal, Auu, Avv);
        else
            orc_basal_sliding_local(&g2, o->dx, o->dy, active_vertex, beta_internal, Auu, Avv);
        memcpy(bu, l2u, nv * 8); memcpy(bv, l2v, nv * 8);
        orc_dirichlet_3d(&g2, active_vertex, um2, vm2, u2, v2, Auu, Auv, Avu, Avv, bu, bv);
        orc_halo(&g2, 27, Auu); orc_halo(&g2, 27, Auv); orc_halo(&g2, 27, Avu); orc_halo(&g2, 27, Avv);
        orc_halo(&g2, 1, bu); orc_halo(&g2, 1, bv);
        orc_check_symmetry_3d(&g2, active_vertex, Auu, Auv, Avu, Avv, NULL);
        long nnz = orc_count_nonzeros_3d(&g2, active_vertex, Auu, Auv, Avu, Avv);
        if (nnz == 0) {
            memset(resid_u, 0, nv * 8); memset(resid_v, 0, nv * 8);
            memset(bu, 0, nv * 8); memset(bv, 0, nv * 8);
            memset(u2, 0, nv * 8); memset(v2, 0, nv * 8);
            L2 = 0.0;
            break;
        }
        if (o->whichbabc == 6) {   /* no-slip: the SSA velocity is zero (:3234-3245) */
            memset(u2, 0, nv * 8); memset(v2, 0, nv * 8);
            memset(resid_u, 0, nv * 8); memset(resid_v, 0, nv * 8);
            L2 = 0.0;
            L2rel = 0.0;
        } else {
            orc_residual_vector_3d(&g2, active_vertex, Auu, Auv, Avu, Avv, bu, bv, u2, v2, resid_u, resid_v,
                                   &L2, &L2rel);
            if (o->L2_history && counter <= o->L2_history_cap) o->L2_history[counter - 1] = L2;
            double err; int niters;
            if (o->whichsparse == 3)
                status = orc_pcg_chrongear_3d(&g2, active_vertex, Auu, Auv, Avu, Avv, bu, bv, u2, v2,
                                              o->whichprecond, o->linear_maxiters, o->linear_tolerance,
                                              &err, &niters);
            else
                status = orc_pcg_standard_3d(&g2, active_vertex, Auu, Auv, Avu, Avv, bu, bv, u2, v2,
                                             o->whichprecond, o->linear_maxiters, o->linear_tolerance,
                                             &err, &niters);
            if (status) break;
            o->total_linear_iters += niters;
        }
        orc_halo(&g2, 1, u2); orc_halo(&g2, 1, v2);
        /* compute_residual_velocity_2d: the 2D field has no separate bed level */
        resid_velo = orc_residual_velocity_3d(&g2, o->whichresid == 1 ? 0 : o->whichresid, u2, v2, us, vs);
        if (o->whichresid == 3) { crit = L2; target = L2_target; }
        else if (o->whichresid == 4) { crit = L2rel; target = L2_target_rel; }
        else { crit = resid_velo; target = resid_target; }
        if (o->target_override > 0) target = o->target_override;
    }
    o->converged = counter < o->maxiter_nonlinear;
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            const size_t c = (i - 1) + (size_t)(nx - 1) * (j - 1);
            for (int k = 1; k <= nz; ++k) {
                V3(uvel, k, i, j) = u2[c];
                V3(vvel, k, i, j) = v2[c];
            }
        }
    o->picard_iters = counter;
    o->final_L2_norm = L2;
    o->final_resid_velo = resid_velo;
    o->status = status;
    free(Auu); free(Auv); free(Avu); free(Avv); free(u2); free(v2); free(us); free(vs);
    free(l2u); free(l2v); free(um2); free(vm2); free(beta_work);
    return status;
}

/* ------------------------------------------------------------------------ */
/* Depth-integrated 2D solvers: L1L2 (which_ho_approx = 3) and DIVA (= 4)   */
/* (velo_higher:1117-1173, 2703-2830, 3655-3700, 3773-3830, 5208-6356,      */
/* 7382-7864; glissade_grid_operators.F90:60-410).                          */
/* ------------------------------------------------------------------------ */

/* glissade_stagger (glissade_grid_operators.F90:60-148): margin 0 averages the four
 * cells around each vertex, margin 1 only those with mask = 1 (zero where none has) */
void orc_stagger(int nx, int ny, const double *var, double *stagvar, const int32_t *mask, int margin)
{
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            double s = 0.0;
            if (margin == 0) {
                s = (C2(var, i, j + 1) + C2(var, i + 1, j + 1) + C2(var, i, j) + C2(var, i + 1, j)) / 4.0;
            } else {
                const double sv = (double)C2(mask, i, j + 1) * C2(var, i, j + 1) +
                                  (double)C2(mask, i + 1, j + 1) * C2(var, i + 1, j + 1) +
                                  (double)C2(mask, i, j) * C2(var, i, j) +
                                  (double)C2(mask, i + 1, j) * C2(var, i + 1, j);
                const double sm = (double)(C2(mask, i, j + 1) + C2(mask, i + 1, j + 1) + C2(mask, i, j) +
                                           C2(mask, i + 1, j));
                if (sm > 0.0) s = sv / sm;
            }
            S2(stagvar, i, j) = s;
        }
}

/* glissade_gradient (glissade_grid_operators.F90:235-410): the vertex gradient is the
 * mean of the two adjacent edge differences; with margin 1 an edge counts only when
 * both its cells have mask = 1 */
void orc_gradient(int nx, int ny, double dx, double dy, const double *f, double *dfdx, double *dfdy,
                  const int32_t *mask, int margin)
{
#define EX(i, j) (margin == 0 || (C2(mask, i, j) == 1 && C2(mask, (i) + 1, j) == 1))
#define EY(i, j) (margin == 0 || (C2(mask, i, j) == 1 && C2(mask, i, (j) + 1) == 1))
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            const double xn = (C2(f, i + 1, j + 1) - C2(f, i, j + 1)) / dx;
            const double xs = (C2(f, i + 1, j) - C2(f, i, j)) / dx;
            double gx = 0.0, gy = 0.0;
            if (EX(i, j) && EX(i, j + 1)) gx = (xn + xs) / 2.0;
            else if (EX(i, j)) gx = xs;
            else if (EX(i, j + 1)) gx = xn;
            const double ye = (C2(f, i + 1, j + 1) - C2(f, i + 1, j)) / dy;
            const double yw = (C2(f, i, j + 1) - C2(f, i, j)) / dy;
            if (EY(i, j) && EY(i + 1, j)) gy = (ye + yw) / 2.0;
            else if (EY(i, j)) gy = yw;
            else if (EY(i + 1, j)) gy = ye;
            S2(dfdx, i, j) = gx;
            S2(dfdy, i, j) = gy;
        }
#undef EX
#undef EY
}

/* get_basis_function_derivatives_2d (velo_higher:6583-6751) for reference-element
 * derivatives dxr, dyr; returns nonzero for a singular Jacobian */
static int basis_2d(const double *x, const double *y, const double *dxr, const double *dyr,
                    double *dx, double *dy, double *detJ)
{
    double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
    for (int n = 0; n < 4; ++n) {
        J11 = J11 + dxr[n] * x[n];
        J12 = J12 + dxr[n] * y[n];
        J21 = J21 + dyr[n] * x[n];
        J22 = J22 + dyr[n] * y[n];
    }
    const double det = J11 * J22 - J12 * J21;
    *detJ = det;
    if (!(fabs(det) > 0.0)) return 1;
    const double i11 = J22 / det, i12 = -J12 / det, i21 = -J21 / det, i22 = J11 / det;
    for (int n = 0; n < 4; ++n) {
        dx[n] = i11 * dxr[n] + i12 * dyr[n];
        dy[n] = i21 * dxr[n] + i22 * dyr[n];
    }
    return 0;
}

/* reference-element tables of the 2D quadrature point p (0..3) or, for p = 4, the cell
 * center (velo_higher:560-600) */
static void tables_2d(int p, double *phi, double *dxr, double *dyr)
{
    const double r = 1.0 / sqrt(3.0);
    const double xq = p == 4 ? 0.0 : ((p == 0 || p == 3) ? -r : r);
    const double yq = p == 4 ? 0.0 : (p < 2 ? -r : r);
    dxr[0] = -(1.0 - yq) / 4.0; dxr[1] = (1.0 - yq) / 4.0;
    dxr[2] = (1.0 + yq) / 4.0;  dxr[3] = -(1.0 + yq) / 4.0;
    dyr[0] = -(1.0 - xq) / 4.0; dyr[1] = -(1.0 + xq) / 4.0;
    dyr[2] = (1.0 + xq) / 4.0;  dyr[3] = (1.0 - xq) / 4.0;
    phi[0] = (1.0 - xq) * (1.0 - yq) / 4.0;
    phi[1] = (1.0 + xq) * (1.0 - yq) / 4.0;
    phi[2] = (1.0 + xq) * (1.0 + yq) / 4.0;
    phi[3] = (1.0 - xq) * (1.0 + yq) / 4.0;
}

/* compute_effective_viscosity_L1L2 (velo_higher:7382-7618): the viscosity of every
 * layer at one quadrature point.  Nonlinear case: tau_parallel is the real root of
 * x^3 + a x + b = 0 (a = tau_perp^2, b = -effstrain/A, Cardano's form with the
 * first-order branch for small a), efvs = 1 / (2 A (tau_parallel^2 + tau_perp^2)). */
void orc_efvs_l1l2(int whichefvs, double efvs_constant, int nz3, const double *sigma,
                   const double *phi, const double *dphi_dx, const double *dphi_dy,
                   const double *u, const double *v, const double *h, const double *dsdx,
                   const double *dsdy, const double *flwa, const double *flwafact, double *efvs)
{
    const double p_effstr = (1.0 - (double)GN) / (double)GN, emin = 1.0e-8;
    if (whichefvs == 0) {
        for (int k = 0; k < nz3 - 1; ++k) efvs[k] = efvs_constant;
        return;
    }
    if (whichefvs == 1) {
        const double effstrain = VEL_SCALE / LEN_SCALE * SCYR;
        for (int k = 0; k < nz3 - 1; ++k) efvs[k] = flwafact[k] * pow(effstrain, p_effstr);
        return;
    }
    double ux = 0, uy = 0, vx = 0, vy = 0, sx = 0, sy = 0, th = 0;
    for (int n = 0; n < 4; ++n) {
        ux = ux + dphi_dx[n] * u[n];
        uy = uy + dphi_dy[n] * u[n];
        vx = vx + dphi_dx[n] * v[n];
        vy = vy + dphi_dy[n] * v[n];
        sx = sx + phi[n] * dsdx[n];
        sy = sy + phi[n] * dsdy[n];
        th = th + phi[n] * h[n];
    }
    const double sq = emin * emin + ux * ux + vy * vy + ux * vy + 0.25 * ((vx + uy) * (vx + uy));
    const double effstrain = sqrt(sq);
    for (int k = 0; k < nz3 - 1; ++k) {
        const double depth = th * sigma[k + 1];
        const double grads = sqrt(sx * sx + sy * sy);
        const double tau_perp = RHOI * GRAV * depth * grads;
        const double a = tau_perp * tau_perp;
        const double b = -effstrain / flwa[k];
        const double c = sqrt(b * b / 4.0 + a * a * a / 27.0);
        const double rootA = pow(-b / 2.0 + c, 1.0 / 3.0);
        double rootB;
        if (a * a * a / 27.0 > 1.0e-6 * (b * b / 4.0))
            rootB = -pow(b / 2.0 + c, 1.0 / 3.0);
        else
            rootB = -a / (3.0 * pow(fabs(b), 1.0 / 3.0));
        const double tp = rootA + rootB;
        efvs[k] = 1.0 / (2.0 * flwa[k] * (tp * tp + tau_perp * tau_perp));
    }
}

/* compute_effective_viscosity_diva (velo_higher:7622-7864) with cubic = .false.: the
 * vertical shear of each layer from the previous efvs (efvs in/out; a zero efvs starts
 * from the minimum strain rate's) */
void orc_efvs_diva(int whichefvs, double efvs_constant, int nz3, const double *stagsigma,
                   const double *phi, const double *dphi_dx, const double *dphi_dy,
                   const double *u, const double *v, const double *bx, const double *by,
                   const double *flwafact, double *efvs)
{
    const double p_effstr = (1.0 - (double)GN) / (double)GN, p2_effstr = p_effstr / 2, emin = 1.0e-8;
    if (whichefvs == 0) {
        for (int k = 0; k < nz3 - 1; ++k) efvs[k] = efvs_constant;
        return;
    }
    if (whichefvs == 1) {
        const double effstrain = VEL_SCALE / LEN_SCALE * SCYR;
        for (int k = 0; k < nz3 - 1; ++k) efvs[k] = flwafact[k] * pow(effstrain, p_effstr);
        return;
    }
    double ux = 0, uy = 0, vx = 0, vy = 0, tx = 0, ty = 0;
    for (int n = 0; n < 4; ++n) {
        ux = ux + dphi_dx[n] * u[n];
        uy = uy + dphi_dy[n] * u[n];
        vx = vx + dphi_dx[n] * v[n];
        vy = vy + dphi_dy[n] * v[n];
        tx = tx + phi[n] * bx[n];
        ty = ty + phi[n] * by[n];
    }
    for (int k = 0; k < nz3 - 1; ++k) {
        if (efvs[k] == 0.0) efvs[k] = flwafact[k] * pow(emin, p_effstr);
        const double uz = tx * stagsigma[k] / efvs[k];
        const double vz = ty * stagsigma[k] / efvs[k];
        const double sq = emin * emin + ux * ux + vy * vy + ux * vy + 0.25 * ((vx + uy) * (vx + uy)) +
                          0.25 * (uz * uz + vz * vz);
        efvs[k] = flwafact[k] * pow(sq, p2_effstr);
    }
}

/* compute_integrals_diva (velo_higher:5601-5726): omega_k(k) (nz3 levels) is the
 * quadrature-point mean of the integral from the surface of depth/efvs over the layers
 * below level k; omega the trapezoid integral of omega_k over sigma.  efvs_qp is
 * (nz3-1, 4). */
void orc_integrals_diva(int nz3, const double *sigma, double thck, const double *efvs_qp,
                        double *omega_k, double *omega)
{
    double okp[64][4];
    for (int p = 0; p < 4; ++p) {
        okp[nz3 - 1][p] = 0.0;
        for (int k = nz3 - 2; k >= 0; --k) {
            const double depth = 0.5 * (sigma[k] + sigma[k + 1]);
            const double dz = (sigma[k + 1] - sigma[k]) * thck;
            okp[k][p] = okp[k + 1][p] + depth / efvs_qp[k + (size_t)(nz3 - 1) * p] * dz;
        }
    }
    for (int k = 0; k < nz3; ++k) omega_k[k] = (okp[k][0] + okp[k][1] + okp[k][2] + okp[k][3]) / 4;
    double om = 0.0;
    for (int k = 0; k < nz3 - 1; ++k) {
        const double layer_avg = 0.5 * (omega_k[k] + omega_k[k + 1]);
        om = om + layer_avg * (sigma[k + 1] - sigma[k]);
    }
    *omega = om;
}

/* assemble_stiffness_matrix_2d, L1L2 and DIVA branches (velo_higher:5208-5595).
 * g: the one-level grid.  Per active cell and quadrature point the layer viscosities,
 * their sigma-weighted column mean, and the element matrix of thickness x that mean
 * (compute_element_matrix's Blatter-Pattyn form with dphi_dz = 0); efvs (nz3-1, nx, ny)
 * is the quadrature-point mean.  DIVA: efvs_qp3 (nz3-1, 4, nx, ny) in/out, omega_k
 * (nz3, nx, ny) and omega (nx, ny) out for active cells. */
int orc_assemble_2d_depth(const orc_grid *g, int whichapprox, int nz3, const double *sigma,
                          const int32_t *active_cell, const double *xV, const double *yV,
                          const double *uvel, const double *vvel, const double *stagthck,
                          const double *dusrf_dx, const double *dusrf_dy, const double *thck,
                          const double *flwa, const double *flwafact, const double *btractx,
                          const double *btracty, int whichefvs, double efvs_constant, double *efvs,
                          double *efvs_qp3, double *omega_k, double *omega,
                          double *Auu, double *Auv, double *Avu, double *Avv)
{
    const int nx = g->nx, ny = g->ny, h = g->nhalo;
    const int nz = 1;
    const int L = nz3 - 1;
    const size_t na = 27 * (size_t)(nx - 1) * (ny - 1);
    memset(Auu, 0, na * 8); memset(Auv, 0, na * 8); memset(Avu, 0, na * 8); memset(Avv, 0, na * 8);
    memset(efvs, 0, (size_t)L * nx * ny * 8);
    const int di[4] = {-1, 0, 0, -1}, dj[4] = {-1, -1, 0, 0};
    double stagsigma[64];
    for (int k = 0; k < L; ++k) stagsigma[k] = (sigma[k] + sigma[k + 1]) / 2.0;
    int status = ORC_OK;
    for (int j = h + 1; j <= ny - h + 1 && !status; ++j)
        for (int i = h + 1; i <= nx - h + 1 && !status; ++i) {
            if (!C2(active_cell, i, j)) continue;
            double x[4], y[4], u[4], v[4], hn[4], sx[4] = {0}, sy[4] = {0}, bx[4] = {0}, by[4] = {0};
            for (int n = 0; n < 4; ++n) {
                const int iv = i + di[n], jv = j + dj[n];
                x[n] = S2(xV, iv, jv);
                y[n] = S2(yV, iv, jv);
                u[n] = S2(uvel, iv, jv);
                v[n] = S2(vvel, iv, jv);
                hn[n] = S2(stagthck, iv, jv);
                if (whichapprox == 3) { sx[n] = S2(dusrf_dx, iv, jv); sy[n] = S2(dusrf_dy, iv, jv); }
                if (whichapprox == 4) { bx[n] = S2(btractx, iv, jv); by[n] = S2(btracty, iv, jv); }
            }
            const size_t cell = (i - 1) + (size_t)nx * (j - 1);
            const double *fl = flwa + (size_t)L * cell, *ff = flwafact + (size_t)L * cell;
            double eqp[64 * 4];
            if (whichapprox == 4) memcpy(eqp, efvs_qp3 + (size_t)L * 4 * cell, sizeof(double) * L * 4);
            double Kuu[16] = {0}, Kuv[16] = {0}, Kvu[16] = {0}, Kvv[16] = {0};
            for (int p = 0; p < 4; ++p) {
                double phi[4], dxr[4], dyr[4], dx[4], dy[4], det;
                tables_2d(p, phi, dxr, dyr);
                if (basis_2d(x, y, dxr, dyr, dx, dy, &det)) { status = ORC_ERR_SINGULAR; break; }
                double *e = eqp + (size_t)L * p;
                if (whichapprox == 3)
                    orc_efvs_l1l2(whichefvs, efvs_constant, nz3, sigma, phi, dx, dy, u, v, hn, sx, sy, fl, ff, e);
                else
                    orc_efvs_diva(whichefvs, efvs_constant, nz3, stagsigma, phi, dx, dy, u, v, bx, by, ff, e);
                double avg = 0.0;
                for (int k = 0; k < L; ++k) avg = avg + e[k] * (sigma[k + 1] - sigma[k]);
                double hq = 0.0;
                for (int n = 0; n < 4; ++n) hq = hq + phi[n] * hn[n];
                const double f = hq * avg * 1.0 * det / VOL0;
                for (int nc = 0; nc < 4; ++nc)
                    for (int nr = 0; nr < 4; ++nr) {
                        const int q = nr + 4 * nc;
                        Kuu[q] = Kuu[q] + f * (4.0 * dx[nr] * dx[nc] + dy[nr] * dy[nc] + 0.0 * 0.0);
                        Kuv[q] = Kuv[q] + f * (2.0 * dx[nr] * dy[nc] + dy[nr] * dx[nc]);
                        Kvu[q] = Kvu[q] + f * (2.0 * dy[nr] * dx[nc] + dx[nr] * dy[nc]);
                        Kvv[q] = Kvv[q] + f * (4.0 * dy[nr] * dy[nc] + dx[nr] * dx[nc] + 0.0 * 0.0);
                    }
            }
            if (status) break;
            if (whichapprox == 4) {
                memcpy(efvs_qp3 + (size_t)L * 4 * cell, eqp, sizeof(double) * L * 4);
                orc_integrals_diva(nz3, sigma, C2(thck, i, j), eqp, omega_k + (size_t)nz3 * cell,
                                   &C2(omega, i, j));
            }
            for (int p = 0; p < 4; ++p)
                for (int k = 0; k < L; ++k) efvs[k + (size_t)L * cell] = efvs[k + (size_t)L * cell] + eqp[k + (size_t)L * p];
            for (int k = 0; k < L; ++k) efvs[k + (size_t)L * cell] = efvs[k + (size_t)L * cell] / 4;
            for (int a = 0; a < 4 && !status; ++a)
                for (int b = 0; b < 4; ++b)
                    if (fabs(Kuu[a + 4 * b] - Kuu[b + 4 * a]) > 1.0e-10 ||
                        fabs(Kvv[a + 4 * b] - Kvv[b + 4 * a]) > 1.0e-10 ||
                        fabs(Kuv[a + 4 * b] - Kvu[b + 4 * a]) > 1.0e-10) {
                        status = ORC_ERR_ASYMMETRIC;
                        break;
                    }
            for (int nr = 0; nr < 4; ++nr) {
                const int iv = i + di[nr], jv = j + dj[nr];
                for (int nc = 0; nc < 4; ++nc) {
                    const int m = IDX_A(di[nc] - di[nr], dj[nc] - dj[nr], 0);
                    const int q = nr + 4 * nc;
                    A4(Auu, m, 1, iv, jv) = A4(Auu, m, 1, iv, jv) + Kuu[q];
                    A4(Auv, m, 1, iv, jv) = A4(Auv, m, 1, iv, jv) + Kuv[q];
                    A4(Avu, m, 1, iv, jv) = A4(Avu, m, 1, iv, jv) + Kvu[q];
                    A4(Avv, m, 1, iv, jv) = A4(Avv, m, 1, iv, jv) + Kvv[q];
                }
            }
        }
    return status;
}

/* compute_3d_velocity_L1L2 (velo_higher:5834-6356), method 1 (edge_velocity = .false.):
 * g: the 3D grid (nz levels).  On input uvel/vvel hold the basal (2D) velocity at level
 * nz; the levels above are integrated from the bed with the vertical shear stresses
 * tau_xz, tau_yz (the SIA stress plus centered differences of the depth-integrated
 * membrane stresses) and tau_eff^2 = tau_parallel^2 + tau_xz^2 + tau_yz^2 on owned
 * active vertices; every other vertex keeps zero above the bed.  umask2/vmask2 are the
 * bed level's Dirichlet masks ((nx-1) x (ny-1)); no halo update (the caller's). */
int orc_velocity_3d_l1l2(const orc_grid *g, const double *sigma, double dx, double dy,
                         const int32_t *ice_mask, const int32_t *active_cell,
                         const int32_t *active_vertex, const int32_t *umask2, const int32_t *vmask2,
                         const double *xV, const double *yV, const double *thck, const double *stagthck,
                         const double *dusrf_dx, const double *dusrf_dy, const double *flwa,
                         const double *efvs, double *uvel, double *vvel)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz, h = g->nhalo;
    const size_t nc = (size_t)nx * ny, nv = (size_t)(nx - 1) * (ny - 1);
    double *ux = calloc(nc, 8), *uy = calloc(nc, 8), *vx = calloc(nc, 8), *vy = calloc(nc, 8);
    double *tpar = calloc((nz - 1) * nc, 8), *eint = calloc((nz - 1) * nc, 8);
    double *txz = calloc((nz - 1) * nv, 8), *tyz = calloc((nz - 1) * nv, 8);
    double *w1 = malloc(nc * 8), *w2 = malloc(nc * 8), *w3 = malloc(nc * 8);
    double *g1x = malloc(nv * 8), *g1y = malloc(nv * 8), *g2x = malloc(nv * 8), *g2y = malloc(nv * 8);
    double *g3x = malloc(nv * 8), *g3y = malloc(nv * 8), *st = malloc(nv * 8), *sf = malloc(nv * 8);
    double *wk = malloc(nc * 8);
    const int di[4] = {-1, 0, 0, -1}, dj[4] = {-1, -1, 0, 0};
    int status = ORC_OK;
    for (int j = h + 1; j <= ny - h + 1 && !status; ++j)
        for (int i = h + 1; i <= nx - h + 1; ++i) {
            if (!C2(active_cell, i, j)) continue;
            double x[4], y[4], u[4], v[4], phi[4], dxr[4], dyr[4], ddx[4], ddy[4], det;
            for (int n = 0; n < 4; ++n) {
                x[n] = S2(xV, i + di[n], j + dj[n]);
                y[n] = S2(yV, i + di[n], j + dj[n]);
                u[n] = V3(uvel, nz, i + di[n], j + dj[n]);
                v[n] = V3(vvel, nz, i + di[n], j + dj[n]);
            }
            tables_2d(4, phi, dxr, dyr);
            if (basis_2d(x, y, dxr, dyr, ddx, ddy, &det)) { status = ORC_ERR_SINGULAR; break; }
            for (int n = 0; n < 4; ++n) {
                C2(ux, i, j) = C2(ux, i, j) + ddx[n] * u[n];
                C2(uy, i, j) = C2(uy, i, j) + ddy[n] * u[n];
                C2(vx, i, j) = C2(vx, i, j) + ddx[n] * v[n];
                C2(vy, i, j) = C2(vy, i, j) + ddy[n] * v[n];
            }
            const double a = C2(ux, i, j), b = C2(vy, i, j), t = C2(vx, i, j) + C2(uy, i, j);
            const double eps = sqrt(a * a + b * b + a * b + 0.25 * (t * t));
            for (int k = 1; k <= nz - 1; ++k) F3(tpar, k, i, j) = 2.0 * F3(efvs, k, i, j) * eps;
            F3(eint, 1, i, j) = F3(efvs, 1, i, j) * (sigma[1] - sigma[0]) * C2(thck, i, j);
            for (int k = 2; k <= nz - 1; ++k)
                F3(eint, k, i, j) = F3(eint, k - 1, i, j) + F3(efvs, k, i, j) * (sigma[k] - sigma[k - 1]) * C2(thck, i, j);
        }
    if (!status) {
        for (int k = 1; k <= nz - 1; ++k) {
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    const double e = F3(eint, k, i, j);
                    C2(w1, i, j) = e * (2.0 * C2(ux, i, j) + C2(vy, i, j));
                    C2(w2, i, j) = e * (C2(uy, i, j) + C2(vx, i, j));
                    C2(w3, i, j) = e * (2.0 * C2(vy, i, j) + C2(ux, i, j));
                }
            orc_gradient(nx, ny, dx, dy, w1, g1x, g1y, ice_mask, 1);
            orc_gradient(nx, ny, dx, dy, w2, g2x, g2y, ice_mask, 1);
            orc_gradient(nx, ny, dx, dy, w3, g3x, g3y, ice_mask, 1);
            for (int j = g->jlo; j <= g->jhi; ++j)
                for (int i = g->ilo; i <= g->ihi; ++i) {
                    if (!S2(active_vertex, i, j)) continue;
                    const double depth = 0.5 * (sigma[k - 1] + sigma[k]) * S2(stagthck, i, j);
                    const size_t q = (k - 1) + (size_t)(nz - 1) * ((i - 1) + (size_t)(nx - 1) * (j - 1));
                    txz[q] = -(RHOI * GRAV * depth * S2(dusrf_dx, i, j)) + 2.0 * S2(g1x, i, j) + S2(g2y, i, j);
                    tyz[q] = -(RHOI * GRAV * depth * S2(dusrf_dy, i, j)) + S2(g2x, i, j) + 2.0 * S2(g3y, i, j);
                }
        }
        for (int j = 1; j <= ny - 1; ++j)
            for (int i = 1; i <= nx - 1; ++i)
                for (int k = 1; k <= nz - 1; ++k) { V3(uvel, k, i, j) = 0.0; V3(vvel, k, i, j) = 0.0; }
        for (int k = nz - 1; k >= 1; --k) {
            for (size_t c = 0; c < nc; ++c) wk[c] = tpar[(k - 1) + (size_t)(nz - 1) * c];
            orc_stagger(nx, ny, wk, st, ice_mask, 1);
            for (size_t c = 0; c < nv; ++c) st[c] = st[c] * st[c];
            for (size_t c = 0; c < nc; ++c) wk[c] = flwa[(k - 1) + (size_t)(nz - 1) * c];
            orc_stagger(nx, ny, wk, sf, ice_mask, 1);
            for (int j = g->jlo; j <= g->jhi; ++j)
                for (int i = g->ilo; i <= g->ihi; ++i) {
                    if (!S2(active_vertex, i, j)) continue;
                    const size_t q = (k - 1) + (size_t)(nz - 1) * ((i - 1) + (size_t)(nx - 1) * (j - 1));
                    const double tes = S2(st, i, j) + txz[q] * txz[q] + tyz[q] * tyz[q];
                    /* tau_eff_sq**((gn-1)/2) with gn = 3 is tau_eff_sq itself */
                    const double fact = 2.0 * S2(sf, i, j) * tes * (sigma[k] - sigma[k - 1]) * S2(stagthck, i, j);
                    V3(uvel, k, i, j) = S2(umask2, i, j) == 1 ? V3(uvel, nz, i, j) : V3(uvel, k + 1, i, j) + fact * txz[q];
                    V3(vvel, k, i, j) = S2(vmask2, i, j) == 1 ? V3(vvel, nz, i, j) : V3(vvel, k + 1, i, j) + fact * tyz[q];
                }
        }
    }
    free(ux); free(uy); free(vx); free(vy); free(tpar); free(eint); free(txz); free(tyz);
    free(w1); free(w2); free(w3); free(g1x); free(g1y); free(g2x); free(g2y); free(g3x); free(g3y);
    free(st); free(sf); free(wk);
    return status;
}

/* compute_3d_velocity_diva (velo_higher:5730-5830), diva_level_index = 0: on active
 * vertices the bed velocity u_b = u_mean - tau_b stag_omega and u(k) = u_b + tau_b
 * stag_omega_k(k) (stag_omega_k is (nz, nx-1, ny-1)) */
void orc_velocity_3d_diva(const orc_grid *g, const int32_t *active_vertex, const double *stag_omega_k,
                          const double *stag_omega, const double *btx, const double *bty,
                          const double *u2, const double *v2, double *uvel, double *vvel)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            if (!S2(active_vertex, i, j)) continue;
            V3(uvel, nz, i, j) = S2(u2, i, j) - S2(btx, i, j) * S2(stag_omega, i, j);
            V3(vvel, nz, i, j) = S2(v2, i, j) - S2(bty, i, j) * S2(stag_omega, i, j);
            for (int k = 1; k <= nz - 1; ++k) {
                V3(uvel, k, i, j) = V3(uvel, nz, i, j) + S2(btx, i, j) * V3(stag_omega_k, k, i, j);
                V3(vvel, k, i, j) = V3(vvel, nz, i, j) + S2(bty, i, j) * V3(stag_omega_k, k, i, j);
            }
        }
}

/* The Picard loop of glissade_velo_higher_solve for which_ho_approx = 3 (L1L2) and 4
 * (DIVA) (velo_higher:1117-1173, 1274-1324, 2206-3767 with solve_2d).  g: the 3D grid.
 *   L1L2: the 2D velocity starts from the bed level of uvel/vvel; beta_internal is the
 *         basal friction; after the loop the 3D velocity is integrated from the bed
 *         (compute_3d_velocity_L1L2) and halo-updated.  With a no-slip bed the velocity
 *         is zero and the loop ends at once (velo_higher:3234-3245).
 *   DIVA: the 2D velocity is the depth-mean velocity, from df->uvel_2d/vvel_2d; the
 *         layer viscosities of each quadrature point start from efvs (in/out); after
 *         each assembly omega (halo-updated) is staggered over df->ice_plus_land_mask and
 *         the basal friction is beta_eff = beta / (1 + beta omega) (1/omega under
 *         no-slip); after each solve the basal traction beta_eff u and the 3D velocity
 *         (compute_3d_velocity_diva) follow.
 * The 2D load is the column sum of loadu/loadv; the Dirichlet masks are the bed level's.
 * Outputs: uvel/vvel (3D), df->uvel_2d/vvel_2d, df->btractx/btracty (DIVA), efvs (the
 * layer viscosities of the last assembly), bu/bv/resid_u/resid_v one-level. */
int orc_picard_solve_depth(const orc_grid *g, orc_picard_opts *o, const orc_depth_fields *df,
                           const double *sigma, const int32_t *active_cell,
                           const int32_t *active_vertex, const double *xV, const double *yV,
                           const double *stagthck, const double *flwafact, const double *beta_internal,
                           const int32_t *umask, const int32_t *vmask,
                           const double *loadu, const double *loadv,
                           double *uvel, double *vvel, double *efvs,
                           double *bu, double *bv, double *resid_u, double *resid_v)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz, L = nz - 1;
    const int approx = o->whichapprox;
    if ((approx != 3 && approx != 4) || nz < 2 || nz > 64) return ORC_ERR_ARG;
    orc_grid g2;
    orc_grid_init(&g2, nx, ny, 1, g->periodic);
    const size_t nv = (size_t)(nx - 1) * (ny - 1), nc = (size_t)nx * ny, na = 27 * nv;
    double *Auu = malloc(na * 8), *Auv = malloc(na * 8), *Avu = malloc(na * 8), *Avv = malloc(na * 8);
    double *u2 = malloc(nv * 8), *v2 = malloc(nv * 8), *us = malloc(nv * 8), *vs = malloc(nv * 8);
    double *l2u = calloc(nv, 8), *l2v = calloc(nv, 8);
    int32_t *um2 = malloc(nv * 4), *vm2 = malloc(nv * 4);
    double *eqp = NULL, *omk = NULL, *om = NULL, *som = NULL, *somk = NULL, *beff = NULL, *wk = NULL;
    double *s1 = NULL;
    if (approx == 4) {
        s1 = malloc(nv * 8);
        eqp = malloc((size_t)L * 4 * nc * 8);
        omk = calloc((size_t)nz * nc, 8);
        om = calloc(nc, 8);
        som = calloc(nv, 8);
        somk = calloc((size_t)nz * nv, 8);
        beff = calloc(nv, 8);
        wk = malloc(nc * 8);
        for (size_t c = 0; c < nc; ++c)
            for (int p = 0; p < 4; ++p)
                memcpy(eqp + (size_t)L * (p + 4 * c), efvs + (size_t)L * c, sizeof(double) * L);
    }
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            const size_t c = (i - 1) + (size_t)(nx - 1) * (j - 1);
            u2[c] = approx == 4 ? df->uvel_2d[c] : V3(uvel, nz, i, j);
            v2[c] = approx == 4 ? df->vvel_2d[c] : V3(vvel, nz, i, j);
            um2[c] = V3(umask, nz, i, j);
            vm2[c] = V3(vmask, nz, i, j);
            for (int k = 1; k <= nz; ++k) {
                l2u[c] = l2u[c] + V3(loadu, k, i, j);
                l2v[c] = l2v[c] + V3(loadv, k, i, j);
            }
        }
    /* velocity of inactive owned vertices is zero (velo_higher:1960-1972) */
    for (int j = g->jlo; j <= g->jhi; ++j)
        for (int i = g->ilo; i <= g->ihi; ++i)
            if (!S2(active_vertex, i, j)) {
                S2(u2, i, j) = 0.0;
                S2(v2, i, j) = 0.0;
                for (int k = 1; k <= nz; ++k) { V3(uvel, k, i, j) = 0.0; V3(vvel, k, i, j) = 0.0; }
            }
    const double L2_target = 1.0e-4, L2_target_rel = 1.0e-7, resid_target = 1.0e-4;
    double crit = 1.0e10, target = 1.0e-12;
    double L2 = 1.0e20, L2rel = 1.0e20, resid_velo = 1.0;
    int counter = 0, status = ORC_OK;
    double *beta_work = o->bp ? malloc(nv * 8) : NULL;
    double *ub = o->bp ? malloc(nv * 8) : NULL, *vb = o->bp ? malloc(nv * 8) : NULL;
    const int noslip = o->whichbabc == 6;
    o->total_linear_iters = 0;
    while (crit >= target && counter < o->maxiter_nonlinear) {
        counter++;
        memcpy(us, u2, nv * 8); memcpy(vs, v2, nv * 8);
        const double *beta = beta_internal;
        if (o->bp) {   /* the basal velocity: the 2D one (L1L2) or the bed level (DIVA), :2237-2243 */
            for (int j = 1; j <= ny - 1; ++j)
                for (int i = 1; i <= nx - 1; ++i) {
                    S2(ub, i, j) = approx == 4 ? V3(uvel, nz, i, j) : S2(u2, i, j);
                    S2(vb, i, j) = approx == 4 ? V3(vvel, nz, i, j) : S2(v2, i, j);
                }
            status = orc_calcbeta(g, o->bp, ub, vb, o->beta_external, o->effecpress_stag, o->f_ground,
                                  beta_work);
            if (status) break;
            beta = beta_work;
        }
        status = orc_assemble_2d_depth(&g2, approx, nz, sigma, active_cell, xV, yV, u2, v2, stagthck,
                                       df->dusrf_dx, df->dusrf_dy, df->thck, df->flwa, flwafact,
                                       df->btractx, df->btracty, o->whichefvs, o->efvs_constant, efvs,
                                       eqp, omk, om, Auu, Auv, Avu, Avv);
        if (status) break;
        if (approx == 4) {
            orc_cell_halo(g, om);
            orc_stagger(nx, ny, om, som, df->ice_plus_land_mask, 1);
            for (size_t c = 0; c < nv; ++c) {
                if (noslip) beff[c] = som[c] > 0.0 ? 1.0 / som[c] : 0.0;
                else beff[c] = beta[c] / (1.0 + beta[c] * som[c]);
            }
            beta = beff;
        }
        /* the 2D branch applies the basal friction for every which_ho_babc (:2832-2852) */
        if (o->assemble_beta_standard)
            orc_basal_sliding_standard(&g2, active_cell, xV, yV, beta, Auu, Avv);
        else
            orc_basal_sliding_local(&g2, o->dx, o->dy, active_vertex, beta, Auu, Avv);
        memcpy(bu, l2u, nv * 8); memcpy(bv, l2v, nv * 8);
        orc_dirichlet_3d(&g2, active_vertex, um2, vm2, u2, v2, Auu, Auv, Avu, Avv, bu, bv);
        orc_halo(&g2, 27, Auu); orc_halo(&g2, 27, Auv); orc_halo(&g2, 27, Avu); orc_halo(&g2, 27, Avv);
        orc_halo(&g2, 1, bu); orc_halo(&g2, 1, bv);
        orc_check_symmetry_3d(&g2, active_vertex, Auu, Auv, Avu, Avv, NULL);
        long nnz = orc_count_nonzeros_3d(&g2, active_vertex, Auu, Auv, Avu, Avv);
        if (nnz == 0) {
            memset(resid_u, 0, nv * 8); memset(resid_v, 0, nv * 8);
            memset(bu, 0, nv * 8); memset(bv, 0, nv * 8);
            memset(u2, 0, nv * 8); memset(v2, 0, nv * 8);
            L2 = 0.0;
            break;
        }
        if (approx == 3 && noslip) {
            memset(u2, 0, nv * 8); memset(v2, 0, nv * 8);
            memset(resid_u, 0, nv * 8); memset(resid_v, 0, nv * 8);
            L2 = 0.0;
            L2rel = 0.0;
        } else {
            orc_residual_vector_3d(&g2, active_vertex, Auu, Auv, Avu, Avv, bu, bv, u2, v2, resid_u, resid_v,
                                   &L2, &L2rel);
            if (o->L2_history && counter <= o->L2_history_cap) o->L2_history[counter - 1] = L2;
            double err; int niters;
            if (o->whichsparse == 3)
                status = orc_pcg_chrongear_3d(&g2, active_vertex, Auu, Auv, Avu, Avv, bu, bv, u2, v2,
                                              o->whichprecond, o->linear_maxiters, o->linear_tolerance,
                                              &err, &niters);
            else
                status = orc_pcg_standard_3d(&g2, active_vertex, Auu, Auv, Avu, Avv, bu, bv, u2, v2,
                                             o->whichprecond, o->linear_maxiters, o->linear_tolerance,
                                             &err, &niters);
            if (status) break;
            o->total_linear_iters += niters;
        }
        orc_halo(&g2, 1, u2); orc_halo(&g2, 1, v2);
        resid_velo = orc_residual_velocity_3d(&g2, o->whichresid == 1 ? 0 : o->whichresid, u2, v2, us, vs);
        if (approx == 4) {
            /* basal traction (Goldberg 2011 eq. 38-39) and the 3D velocity (velo_higher:3661-3700) */
            for (size_t c = 0; c < nv; ++c) {
                df->btractx[c] = beff[c] * u2[c];
                df->btracty[c] = beff[c] * v2[c];
            }
            for (int k = 1; k <= nz; ++k) {
                for (size_t c = 0; c < nc; ++c) wk[c] = omk[(k - 1) + (size_t)nz * c];
                orc_stagger(nx, ny, wk, s1, df->ice_plus_land_mask, 1);
                for (size_t c = 0; c < nv; ++c) somk[(k - 1) + (size_t)nz * c] = s1[c];
            }
            orc_velocity_3d_diva(g, active_vertex, somk, som, df->btractx, df->btracty, u2, v2, uvel, vvel);
            orc_halo(g, nz, uvel);
            orc_halo(g, nz, vvel);
        }
        if (o->whichresid == 3) { crit = L2; target = L2_target; }
        else if (o->whichresid == 4) { crit = L2rel; target = L2_target_rel; }
        else { crit = resid_velo; target = resid_target; }
        if (o->target_override > 0) target = o->target_override;
    }
    o->converged = counter < o->maxiter_nonlinear;
    memcpy(df->uvel_2d, u2, nv * 8);
    memcpy(df->vvel_2d, v2, nv * 8);
    if (approx == 3 && !status) {
        for (int j = 1; j <= ny - 1; ++j)
            for (int i = 1; i <= nx - 1; ++i) {
                V3(uvel, nz, i, j) = S2(u2, i, j);
                V3(vvel, nz, i, j) = S2(v2, i, j);
            }
        status = orc_velocity_3d_l1l2(g, sigma, o->dx, o->dy, df->ice_mask, active_cell, active_vertex, um2,
                                      vm2, xV, yV, df->thck, stagthck, df->dusrf_dx, df->dusrf_dy, df->flwa,
                                      efvs, uvel, vvel);
        orc_halo(g, nz, uvel);
        orc_halo(g, nz, vvel);
    }
    o->picard_iters = counter;
    o->final_L2_norm = L2;
    o->final_resid_velo = resid_velo;
    o->status = status;
    free(Auu); free(Auv); free(Avu); free(Avv); free(u2); free(v2); free(us); free(vs);
    free(l2u); free(l2v); free(um2); free(vm2); free(beta_work); free(ub); free(vb);
    free(eqp); free(omk); free(om); free(som); free(somk); free(beff); free(wk); free(s1);
    return status;
}

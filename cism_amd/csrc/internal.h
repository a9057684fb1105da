// internal.h -- shared definitions of libcism_gpu.so (gfx950 / CDNA4).
//
// Device data layout:
//   node q (0-based) = k + nz*(i + sx*j), sx = nx-1, sy = ny-1, k fastest --
//   the reference's vector order (nz, nx-1, ny-1).
//   Matrix (FP64): tiles of 64 nodes (one wavefront); inside a tile, slot-major,
//   then the block pair (Auu, Auv) or (Avu, Avv), then lane, then the block in the
//   pair (b = 0 Auu, 1 Auv, 2 Avu, 3 Avv):
//       A[aidx(b, m, q)],  aidx = ((q/64)*27 + m)*256 + (b/2)*128 + (q%64)*2 + b%2
//   so one wave's coefficients are one contiguous 27*64*4-value region, a lane
//   gets the four blocks of a slot with two 16-byte loads, and each of those load
//   instructions reads 1 KB contiguous over the wave.  The FP32 symmetric copy
//   (sidx) keeps a lane's four blocks together: one 16-byte load, 1 KB per wave.
//   npad (nz*sx*sy rounded up to 256) keeps tiles whole.
//   Vectors interleave u and v: x[2*q + c], c = 0 (u), 1 (v).
//   Element scratch: keidx(entry, e), tiles of 64 elements, entry-major inside a
//   tile; e = (k-1) + (nz-1)*((i-1) + nx*(j-1)), the reference's efvs layout
//   (nz-1, nx, ny).
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/cism_gpu.h"

#define CISM_NHALO 2
#define CISM_SLOTS 27
#define CISM_DIAG 13           // slot of indxA(0,0,0) = 14 (0-based 13)
#define CISM_KE_ENTRIES 136    // 36 (Kuu upper) + 36 (Kvv upper) + 64 (Kuv)
#define CISM_RED_BLOCKS 1024   // max partial-sum blocks of one reduction
#define CISM_BLOCK 256
#define CISM_MAX_NV 16         // widest per-block partial record of any reduction

// physical constants (glimmer_physcon.F90:64-75) and the matrix volume scale
#define CISM_RHOI 910.0
#define CISM_RHOO 1028.0
#define CISM_GRAV 9.81
#define CISM_SCYR 31536000.0
#define CISM_VOL0 1.0e9

struct GridDev {
    int nx, ny, nz, sx, sy;
    int ilo, ihi, jlo, jhi;   // owned staggered vertices, 0-based inclusive
    int periodic;
    long long nn, npad;       // nodes (nz*sx*sy), padded plane length
    long long ne, nepad;      // elements (nz-1)*nx*ny, padded
};

// device-resident scalars of the CG iterations (one per context)
struct DevScal {
    double rho_old, alpha, beta, sigma, eta0, eta1, eta2;
    double L2_rhs, err;
    double sums[8];
    int nan_flag, singular_flag;
    int zero_stop;   // inner solves: an exactly zero residual ends the iteration (alpha = 0)
    int beta_bad;    // calcbeta met a negative or NaN beta
};

struct cism_guess;   // the learned initial guess of a context (surrogate.hip)

// aggregation coarse correction of the mixed-precision inner solves (coarse.hip): per
// level an ax x ay grid of aggregates with its 9-point 2 x 2-block stencil in FP64 (the
// Galerkin sums) and FP32 (the V-cycle), the l1-Jacobi diagonal, and the V-cycle's
// residual, smoothed and corrected vectors (2 per aggregate each)
#define CG_MAXLEV 12
struct CoarseLevel {
    int ax, ay;
    double *A64;
    float *A, *dl1, *r, *e, *eo;
};
struct CoarseHier {
    int nlev;       // levels; the last one is solved with the dense inverse Ainv
    int ok;         // built for the current FP32 matrix
    CoarseLevel lev[CG_MAXLEV];
    int32_t *cidx;  // level-1 aggregate of every column (-1: none)
    float *colsum;  // (u, v) column sums of the residual, 2 per column
    float *colF;    // level-1 column blocks (9 column directions and the diagonal), 40 per column
    double *S;      // level-1 Galerkin partial sums, 40 per aggregate
    float *Ainv;    // dense inverse of the coarsest level
};

struct cism_ctx {
    GridDev g;
    int device;
    int maxiters;
    double tolerance;
    hipStream_t stream;
    // matrix, FP64 and (lazily) FP32
    double *A;
    float *A32;
    // (lazily) the FP64 matrix in symmetric storage, written by a symmetry pass
    // (a64s_pending) and valid (a64s_valid) once that pass left no pair unrepaired: every
    // coefficient pair of an owned active row then holds one value, and products with
    // A64s equal those with A bit for bit.  Every writer of A clears a64s_valid.
    double *A64s;
    int a64s_valid, a64s_pending;
    int sym64_on;   // cism_set_fp64_symmetric (default 1)
    // vectors (2*npad each): b x r d z q Az t(temp) load usav
    double *b, *x, *r, *d, *z, *q, *Az, *t, *load, *usav;
    double *vstage;   // 2*npad staging for (de)interleaving host vectors
    // FP32 vectors for the inner solves (2*npad each)
    float *x32, *b32, *r32, *d32, *z32, *q32, *Az32;
    // column factors of the SIA preconditioner (6*npad each)
    double *F64;
    float *F32;
    // optional per-column preconditioner scale of the mixed-precision solves (2*ncol)
    float *cscale;
    int cscale_on;
    // one-level (SSA) contexts: element scratch, column-mean flow factor, cell efvs,
    // sigma and the 3D flow factor of the caller's nz3 levels
    double *ssa_Ke, *ssa_ff2, *ssa_efvs2, *ssa_sigma, *ssa_flw3;
    int ssa_nz3;
    // L1L2 / DIVA on a one-level context (ssa.hip): flow factor A and cell thickness, the
    // layer viscosities (nz3-1 per cell; DIVA also per quadrature point), DIVA's column
    // integrals, stag_omega, beta_eff and basal traction, the 3D velocity (nz3 per
    // vertex), L1L2's cell strain rates, a basal-velocity stage, the ice and
    // ice-plus-land cell masks
    double *dep_flwa, *dep_thck, *dep_efvs3, *dep_eqp, *dep_omk, *dep_om, *dep_beff, *dep_som;
    double *dep_btr, *dep_u3, *dep_v3, *dep_cs, *dep_ub;
    int32_t *dep_ice, *dep_ipl;
    int dep_nz3;
    // basal physics of the Picard loop (cism_set_basal_physics): with bp_on, beta is
    // recomputed from the bed velocity at the start of every Picard iteration
    cism_basal_physics bp;
    int bp_on;
    double *bp_bext, *bp_N, *bp_fg;   // beta_external, effective pressure, f_ground (vertices)
    // Anderson acceleration of the Picard map: m columns of dX and dF, previous x and f
    double *aa, *aa_sums;
    int aa_m, aa_count, aa_head;
    double aa_fnorm_prev;
    // masks and geometry
    int32_t *av, *ac, *umask, *vmask;
    int dirichlet_any;   // 0: the Dirichlet masks as uploaded prescribe no node
    double *xV, *yV, *susrf, *sthck, *beta, *sigma, *wsig, *flwafact, *efvs;
    double *dsdx, *dsdy;
    // FP64 reference velocity (lazily) for error reports
    double *xref;
    // element scratch (lazily)
    double *Ke;
    // the active-vertex and active-cell masks of the last assembly (av_zero, ac_zero),
    // valid when a_zero_ok: an owned row with no active cell around it and an inactive
    // vertex, then and now, is still all zero, and the next assembly need not rewrite
    // it.  Any other writer of A clears a_zero_ok.
    int32_t *av_zero, *ac_zero;
    int a_zero_ok;
    // c->A holds the last assembly's matrix (0 after a mixed-path assembly that did not
    // write it); the basal terms of that assembly, for residual_from_elements
    int a64_valid;
    int asm_add_basal;
    double asm_basal_scale;
    // the same record for the FP32 symmetric matrix: the masks of the last symmetry
    // pass that wrote it from an assembled A (as_av_prev, as_ac_prev), valid when as_ok
    int32_t *as_av_prev, *as_ac_prev;
    int as_ok;
    // reductions
    double *partials;
    double *red_groups;   // the first pass of a two-pass reduction (RED_GROUPS x CISM_MAX_NV)
    DevScal *sc;
    DevScal *sc_host;   // pinned mirror
    // staging for layout conversion (27*nn doubles, lazily)
    double *stage;
    // learned initial guess (cism_learned_guess_load), or null
    cism_guess *guess;
    // coarse correction of the mixed-precision inner solves (lazily), or null
    CoarseHier *coarse;
    // cell masks of cism_glissade_velo_higher_solve as int32 planes (ice, floating,
    // ocean, land, calving front = 0), 5 * nx * ny, lazily
    int32_t *planes;
};

// A named host-side range for rocprofv3 --marker-trace, with the reference's timer names
// (t_startf/t_stopf regions of glissade_velo_higher.F90 and glissade_velo_higher_pcg.F90).
// The device work a range enqueues runs asynchronously; with the environment variable
// CISM_ROCTX_SYNC=1 a range given a stream waits for it before it closes, so that each
// range covers its own device work (a few microseconds per range).
bool roctx_sync_enabled();
struct TimerRange {
    explicit TimerRange(const char *name, hipStream_t s = nullptr, bool sync = false) : s_(s), sync_(sync) {
        // with CISM_ROCTX_SYNC=1 a synchronized range first drains the work queued before
        // it, so that it spans its own device work only
        if (sync_ && roctx_sync_enabled()) (void)hipStreamSynchronize(s_);
        roctxRangePushA(name);
    }
    ~TimerRange() {
        if (sync_ && roctx_sync_enabled()) (void)hipStreamSynchronize(s_);
        roctxRangePop();
    }
    TimerRange(const TimerRange &) = delete;
    TimerRange &operator=(const TimerRange &) = delete;

  private:
    hipStream_t s_;
    bool sync_;
};

// ---------------------------------------------------------------------------
// index helpers
// ---------------------------------------------------------------------------
__host__ __device__ inline int slot_of(int iA, int jA, int kA) {
    return (iA + 1) + 3 * (jA + 1) + 9 * (kA + 1);
}

// position of coefficient (block b, slot m) of node q in the tiled matrix layout;
// blocks b and b ^ 1 (b = 0, 2) are adjacent
__host__ __device__ inline long long aidx(int b, int m, long long q) {
    return (((q >> 6) * 27 + m) << 8) + ((long long)(b >> 1) << 7) + ((q & 63) << 1) + (b & 1);
}

// the four blocks (uu, uv, vu, vv) of slot m of node q: one 16-byte load (FP32) or two (FP64)
// FP32 symmetric storage of the mixed-precision solve: 14 slots per node, the
// diagonal block (slot CISM_DIAG) and the 13 couplings after it in indxA order,
// in tiles of 64 nodes like aidx.  A coupling before the diagonal, slot m of node
// q, is the transpose of slot 26 - m of the neighbour it points to.
#define CISM_SYM_SLOTS 14
__host__ __device__ inline long long sidx(int b, int s, long long q) {
    return (((q >> 6) * CISM_SYM_SLOTS + s) << 8) + ((q & 63) << 2) + b;
}
// FP64 symmetric storage (A64s: the same 14 slots, for the FP64 solver once the symmetry
// pass has left the matrix exactly symmetric) in aidx's tiled form: blocks b and b ^ 1
// adjacent, so each 16-byte load of a wave reads 1 KB contiguous
__host__ __device__ inline long long s64idx(int b, int s, long long q) {
    return (((q >> 6) * CISM_SYM_SLOTS + s) << 8) + ((long long)(b >> 1) << 7) + ((q & 63) << 1) + (b & 1);
}

__device__ inline void load_blocks(const float *__restrict__ A, long long q, int m, float &uu,
                                   float &uv, float &vu, float &vv) {
    const float4 a = *reinterpret_cast<const float4 *>(A + aidx(0, m, q));
    uu = a.x; uv = a.y; vu = a.z; vv = a.w;
}
__device__ inline void load_blocks(const double *__restrict__ A, long long q, int m, double &uu,
                                   double &uv, double &vu, double &vv) {
    const double2 a = *reinterpret_cast<const double2 *>(A + aidx(0, m, q));
    const double2 b = *reinterpret_cast<const double2 *>(A + aidx(2, m, q));
    uu = a.x; uv = a.y; vu = b.x; vv = b.y;
}
// The same with non-temporal loads, for kernels that stream the FP64 matrix once (matvec,
// residuals, nonzero count): the matrix does not displace the vectors from L2, and the
// FP64 matvec runs 2-3 % faster.  The symmetry pass reads each pair twice and keeps the
// cached form.
typedef double dvec2_nt __attribute__((ext_vector_type(2)));
__device__ inline void load_blocks_nt(const double *__restrict__ A, long long q, int m, double &uu,
                                      double &uv, double &vu, double &vv) {
    const dvec2_nt a = __builtin_nontemporal_load(reinterpret_cast<const dvec2_nt *>(A + aidx(0, m, q)));
    const dvec2_nt b = __builtin_nontemporal_load(reinterpret_cast<const dvec2_nt *>(A + aidx(2, m, q)));
    uu = a.x; uv = a.y; vu = b.x; vv = b.y;
}
__device__ inline void store_blocks_nt(double *__restrict__ A, long long q, int m, double uu, double uv, double vu,
                                       double vv) {
    const dvec2_nt a = {uu, uv}, b = {vu, vv};
    __builtin_nontemporal_store(a, reinterpret_cast<dvec2_nt *>(A + aidx(0, m, q)));
    __builtin_nontemporal_store(b, reinterpret_cast<dvec2_nt *>(A + aidx(2, m, q)));
}

// periodic / outflow image of a staggered index (parallel_slap.F90:2807-2925):
// returns the owned index whose value the halo entry takes, or -1 for zero
__device__ inline int halo_image(int i, int s, int periodic) {
    if (periodic) {
        if (i < 2) return i + s - 3;
        if (i == s - 1) return 2;
        return i;
    }
    if (i == 0 || i == s - 1) return -1;
    return i;
}

__device__ inline bool is_owned(const GridDev &g, int i, int j) {
    return i >= g.ilo && i <= g.ihi && j >= g.jlo && j <= g.jhi;
}

// whether staggered vertex (i, j) (0-based) touches an active cell of ac among the
// cells the assembly visits (cells nhalo+1 .. n-nhalo+1, 1-based)
__device__ inline bool touches_active_cell(const GridDev &g, const int32_t *__restrict__ ac, int i, int j) {
    bool any = false;
#pragma unroll
    for (int dJ = 0; dJ < 2; ++dJ)
#pragma unroll
        for (int dI = 0; dI < 2; ++dI) {
            const int ic = i + dI, jc = j + dJ;
            if (ic >= CISM_NHALO && ic <= g.nx - CISM_NHALO && jc >= CISM_NHALO && jc <= g.ny - CISM_NHALO)
                any = any || ac[ic + (long long)g.nx * jc] != 0;
        }
    return any;
}

// ---------------------------------------------------------------------------
// deterministic block reduction (fixed lane/wave order)
// ---------------------------------------------------------------------------
// the block's sums to partials[blk * NV + c] (blk: the block's index, or a node-block
// index of the kernel's own when it maps blocks to nodes in another order)
template <int NV>
__device__ inline void block_reduce_store(double (&v)[NV], double *partials, long long blk = -1) {
    static_assert(NV <= CISM_MAX_NV, "partials hold CISM_MAX_NV values per block");
    __shared__ double red[CISM_BLOCK / 64][NV];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        double s = v[c];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        if (lane == 0) red[wave][c] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w][threadIdx.x];
        partials[(size_t)(blk < 0 ? (long long)blockIdx.x : blk) * NV + threadIdx.x] = s;
    }
}

template <int NV>
__device__ inline void block_max_store(double (&v)[NV], double *partials) {
    static_assert(NV <= CISM_MAX_NV, "partials hold CISM_MAX_NV values per block");
    __shared__ double red[CISM_BLOCK / 64][NV];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        double s = v[c];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s = fmax(s, __shfl_down(s, off, 64));
        if (lane == 0) red[wave][c] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = red[0][threadIdx.x];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) s = fmax(s, red[w][threadIdx.x]);
        partials[(size_t)blockIdx.x * NV + threadIdx.x] = s;
    }
}

// grid-stride helpers
inline int nblocks_for(long long n, int block = CISM_BLOCK) {
    long long b = (n + block - 1) / block;
    return (int)(b < 1 ? 1 : b);
}
// XCD-aware launch of the FP32 symmetric matvec: workgroups are dealt round-robin
// to the 8 XCDs, so a grid of xcd_blocks_for(n) blocks (a multiple of 8 chunks) lets
// XCD x take chunks x, x+8, x+16, ... of CISM_XCD_CHUNK consecutive node blocks; the
// neighbours a node reads from the row before it were then read by the same XCD
// moments earlier and sit in its L2
#define CISM_XCD_CHUNK 128
inline int xcd_blocks_for(long long n) {
    const long long per = 8LL * CISM_XCD_CHUNK;
    return (int)((nblocks_for(n) + per - 1) / per * per);
}
__device__ inline long long xcd_block() {
    const unsigned b = blockIdx.x, x = b & 7u, k = b >> 3;
    return ((long long)(k / CISM_XCD_CHUNK) * 8 + x) * CISM_XCD_CHUNK + (k % CISM_XCD_CHUNK);
}

inline int red_blocks_for(long long n) {
    long long b = (n + CISM_BLOCK - 1) / CISM_BLOCK;
    if (b > CISM_RED_BLOCKS) b = CISM_RED_BLOCKS;
    return (int)(b < 1 ? 1 : b);
}

// ---------------------------------------------------------------------------
// error handling
// ---------------------------------------------------------------------------
void cism_set_error(const char *fmt, ...);
#define HIP_TRY(call)                                                             \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            cism_set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,            \
                           hipGetErrorString(e_));                                \
            return CISM_ERR_HIP;                                                  \
        }                                                                         \
    } while (0)
#define CISM_TRY(...)                  \
    do {                               \
        int s_ = (__VA_ARGS__);        \
        if (s_ != CISM_OK) return s_;  \
    } while (0)

// ---------------------------------------------------------------------------
// launchers shared between translation units (all enqueue on ctx->stream)
// ---------------------------------------------------------------------------
// reduce partials[nb][nv] -> sc->sums[0..nv), then apply a scalar update (mode)
enum ReduceMode {
    RED_SUMS = 0,       // just the sums
    RED_BB,             // L2_rhs = sqrt(s0)
    RED_STD_ETA1,       // eta1 = s0; NaN check; beta = eta1/eta0
    RED_STD_ETA2,       // eta2 = s0; eta0 = eta1; alpha = eta1/eta2; NaN check
    RED_CG_FIRST,       // rho_old = s0; sigma = s1; alpha = rho_old/sigma
    RED_CG_ITER,        // Chron-Gear scalars (pcg:1501-1512)
    RED_ERR,            // err = sqrt(s0)/L2_rhs
    RED_MAX             // max instead of sum (sums[c] = max)
};
int launch_reduce(cism_ctx *c, int nb, int nv, int mode);

int launch_matvec64(cism_ctx *c, const double *A, const double *x, double *y);
int launch_halo_vec(cism_ctx *c, double *v);      // 2-component vector
int column_scale_device(cism_ctx *c, double *out);  // out: 2*ncol doubles
// 2 J(x) = x.(A x - b) - b.x of the context's x on owned active rows, t = A x - b
int guess_energy_device(cism_ctx *c, const double *t, double *twoJ);
// SSA (one-level context, ssa.hip)
int ensure_ssa(cism_ctx *c, int nz3);
int ssa_prepare_device(cism_ctx *c);
int assemble2d_device(cism_ctx *c, int whichefvs, double efvs_constant, int add_basal, double dx, double dy);
int ssa_efvs_layers_device(cism_ctx *c, double *efvs3);
// L1L2 (approx 3) and DIVA (approx 4) on a one-level context (ssa.hip)
int ensure_depth(cism_ctx *c, int nz3);
int depth_prepare_device(cism_ctx *c, int approx);
int assemble_depth_device(cism_ctx *c, int approx, int whichefvs, double efvs_constant, int add_basal,
                          int noslip, double dx, double dy);
const double *depth_basal_beta(cism_ctx *c, int approx);   // beta_eff (DIVA) or beta
int diva_post_solve_device(cism_ctx *c);
int l1l2_velocity_device(cism_ctx *c, double dx, double dy);
// calcbeta from a basal velocity (u at ub[stride*col], v at vb[stride*col])
int calcbeta_from_device(cism_ctx *c, const double *ub, const double *vb, long long stride);
int calcbeta_device(cism_ctx *c);
// post-solve diagnostics from the context's geometry and velocity c->x
int internal_stress_device(cism_ctx *c, int whichapprox, int whichefvs, double efvs_constant,
                           double *tau);   // tau: 6 * ne (xz yz xx yy xy eff)
// basal velocity from xb (u at xb[cs*col], v at xb[cs*col+1]); xb null: the bed layer of c->x
int basal_diagnostics_device(cism_ctx *c, int whichassemble_bfric, double *bfric /* nx*ny */,
                             double *btract /* 2*ncol */, const double *xb = nullptr,
                             long long cs = 0);
// Anderson-accelerated update of c->x (the Picard map's output) from c->usav (its
// input); m <= AA_MAX columns; restarts the history when the fixed-point residual grows.
#define AA_MAX 4
int anderson_reset(cism_ctx *c);
int anderson_step(cism_ctx *c, int m, double *fnorm);
int launch_halo_mat(cism_ctx *c, double *A);      // 108 planes
int launch_cg_matvec(cism_ctx *c, int which);
int launch_cg_update32(cism_ctx *c);   // one FP32 fused update (timing)
int rel_err_device(cism_ctx *c, double *rel);
int solver_pcg64(cism_ctx *c, int chrongear, int precond, int maxiters, double tol,
                 double *err, int *niters);
// r_ready: c->r already holds b - A x with sums[2] = |r|^2 and (on the host too)
// sums[3] = |b|^2, as residual_vector_device leaves them with rs = c->r
// stalled (optional): set to 1 when the refinement is given up (refine_policy.h);
// unconverged (optional): set to 1 when its budget ran out before the tolerance
// skip_final = 1: end after the first refinement step without the confirming FP64 residual
// when the inner solve's own estimate meets half the tolerance (cism_solve_opts)
int solver_mixed(cism_ctx *c, int precond, int maxiters, double tol, double inner_rtol,
                 double *err, int *niters, int *n64, int *n32, int r_ready = 0,
                 int *stalled = nullptr, int *unconverged = nullptr, int skip_final = 0);

// As (optional): also write the FP32 symmetric matrix, the Picard residual of the context's
// x (c->t = A x - b, c->r = b - A x with b the load vector, sums[0..3] as
// residual_vector_device) and the nonzero count (sums[4]) in the gather, for a matrix that
// no later pass changes except the halo update (no Dirichlet rows, basal friction local);
// x must carry its halo; launch_a32_halo and launch_halo_vec(c->r) complete the halo rows
// store64 = 0 (with As only): the FP64 rows are not written (c->a64_valid = 0);
// residual_from_elements then gives the FP64 residual of that matrix
int assemble_device(cism_ctx *c, int whichapprox, int whichefvs, double efvs_constant,
                    int add_basal, double dx, double dy, float *As = nullptr, int store64 = 1);
int launch_a32_halo(cism_ctx *c);
int residual_from_elements(cism_ctx *c);   // c->r = b - A x, sums[0] = |r|^2 (owned)
int dirichlet_device(cism_ctx *c);
// basal friction by the standard finite-element assembly onto the assembled c->A (c->ac,
// xV, yV, beta); the local assembly is part of the gather (assemble_device's add_basal)
int basal_standard_device(cism_ctx *c, const double *beta = nullptr);   // null: c->beta
int symmetry_device(cism_ctx *c, float *As, double *As64 = nullptr);
// after the synchronization that follows symmetry_device: A64s is valid when that pass
// wrote it and counted no unrepaired pair (sums[0] == 0)
void sym64_commit(cism_ctx *c);
int ensure_a64s(cism_ctx *c);   // CISM_ERR_NOMEM leaves the FP64 solver on the full storage
int count_nonzeros_device(cism_ctx *c);          // result in sc->sums[0]
// r = A x - b on owned active rows: sums[0] = |r|^2, sums[1] = |b|^2 over them.  With
// rs: also rs = b - A x, halo rows from their images; sums[2] = |rs|^2, sums[3] = |b|^2
// over owned rows
int residual_vector_device(cism_ctx *c, double *ru, double *rs);
int residual_velocity_device(cism_ctx *c, int whichresid);       // sums[0]
int load_gravity_device(cism_ctx *c, int whichassemble_taud, double dx, double dy);
int load_lateral_device(cism_ctx *c, const int32_t *ocean, const int32_t *cfront);

// the learned guess into c->x from the context's av, sthck, dsdx, dsdy, beta, flwafact
int learned_guess_device(cism_ctx *c);
void learned_guess_free(cism_ctx *c);
// the solve's inputs from cell fields (host arrays) into the context; mask bits in c->umask
int velo_inputs_device(cism_ctx *c, double dx, double dy, double thklim, double eus, double max_slope,
                       const double *thck, const double *usrf, const double *topg, const double *flwa);
int velo_inputs_upload(cism_ctx *c, const double *thck, const double *usrf, const double *topg,
                       const double *flwa);
int velo_inputs_compute(cism_ctx *c, double dx, double dy, double thklim, double eus, double max_slope);
// the mask bits left in c->umask by velo_inputs_device as int32 planes in c->planes
int mask_planes_device(cism_ctx *c);
// grounded fraction and basal physics of the Picard loop from the context's mask planes
int basal_setup_device(cism_ctx *c, const int32_t *planes, int whichground, const cism_basal_physics *bp);
// the upload half of basal_setup_device (beta_external and N to the device); checks bp
int basal_fields_upload(cism_ctx *c, const cism_basal_physics *bp, const double *beta_external,
                        const double *effecpress_stag);

// the aggregation coarse correction (coarse.hip).  coarse_setup builds the hierarchy from
// the FP32 symmetric matrix unless it is current; coarse_invalidate marks it stale (the
// symmetry pass that rewrites A32 calls it); coarse_colsum forms the column sums of r that
// coarse_apply restricts (the fused FP32 update writes them itself); coarse_apply adds
// P V(P^T r) to z.
// Sf, cs: the SIA preconditioner's FP32 factors are formed in the same pass (scaled by cs
// when not null); they stay current with the hierarchy
int coarse_setup(cism_ctx *c, const float *As, float *Sf, const float *cs);
void coarse_invalidate(cism_ctx *c);
int coarse_colsum(cism_ctx *c, const float *r);
int coarse_apply(cism_ctx *c, float *z);
void coarse_free(cism_ctx *c);

int ensure_stage(cism_ctx *c);
int ensure_ke(cism_ctx *c);
int ensure_a32(cism_ctx *c);
int upload_matrix(cism_ctx *c, const double *Auu, const double *Auv, const double *Avu,
                  const double *Avv);
int download_matrix(cism_ctx *c, double *Auu, double *Auv, double *Avu, double *Avv);
int upload_vec(cism_ctx *c, double *dst, const double *hu, const double *hv);
int download_vec(cism_ctx *c, const double *src, double *hu, double *hv);
int sync_scalars(cism_ctx *c);   // copies *sc to *sc_host, synchronizes

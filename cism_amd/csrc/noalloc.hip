// noalloc.hip -- compute_residual_vector_3d on caller-owned device arrays, with no
// device allocation (cism_compute_residual_vector_3d_noalloc in include/cism_gpu.h).
//
// The context entry point cism_compute_residual_vector_3d uploads host arrays into
// context buffers and reduces the two norms through the context's partial-sum
// buffer.  This variant takes device arrays in the reference's Fortran layouts and
// keeps its only scratch -- the per-workgroup partial sums of |r|^2 and |b|^2 -- in a
// file-scope __device__ array.  It calls no hipMalloc, hipFree or allocating library
// routine, so code under HIP graph capture or a serving loop that must not allocate
// can call it.  Every call on one GPU uses that GPU's copy of the array: calls that
// overlap on one GPU (other streams or threads) race on it.
#include "internal.h"

#include <mutex>

// at most this many workgroups; each leaves 2 partial sums (|r|^2, |b|^2)
#define NOALLOC_BLOCKS 1024
#define NOALLOC_THREADS 256
#define NOALLOC_MAX_DEVICES 64
// grids above this node count are rejected (the kernels' row and column arithmetic
// is 32-bit inside a column of the grid; 2^30 nodes is also 8 GB per vector)
#define NOALLOC_MAX_NODES (1LL << 30)

// the scratch: 2 * NOALLOC_BLOCKS partial sums (16 KiB); every call writes every
// entry it later reads, so nothing needs clearing between calls
static __device__ __align__(256) double scratch[2 * NOALLOC_BLOCKS];

// device address of `scratch` on each device, looked up once per device with that
// device current (each GPU holds its own copy of the array)
static double *g_scratch[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

static int scratch_on_device(int device, double **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            cism_set_error("residual_vector_3d_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s",
                           device, hipGetErrorString(e));
            return CISM_ERR_HIP;
        }
        g_scratch[device] = (double *)p;
    }
    *out = g_scratch[device];
    return CISM_OK;
}

struct FGrid {
    int nz, sx, sy;            // levels, staggered extents (nx-1, ny-1)
    int ilo, ihi, jlo, jhi;    // owned staggered vertices, 0-based inclusive
    long long nn;
};

// one lane per node (grid-stride): r = A x - b on owned active rows, 0 elsewhere; the
// 27 slots summed in indxA order as the reference's row product (pcg:2607-2645), so the
// vector equals the context entry point's bit for bit.  part[2 * block + {0, 1}] =
// the block's |r|^2 and |b|^2 over owned active rows, in a fixed order.
__global__ __launch_bounds__(NOALLOC_THREADS) void k_resvec_fortran(
    FGrid g, const int32_t *__restrict__ av, const double *__restrict__ Auu, const double *__restrict__ Auv,
    const double *__restrict__ Avu, const double *__restrict__ Avv, const double *__restrict__ bu,
    const double *__restrict__ bv, const double *__restrict__ xu, const double *__restrict__ xv,
    double *__restrict__ ru, double *__restrict__ rv, double *__restrict__ part) {
    double acc[2] = {0.0, 0.0};
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long cst = g.nz, rst = (long long)g.nz * g.sx;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < g.nn; q += stride) {
        const int k = (int)(q % g.nz);
        const long long col = q / g.nz;
        const int i = (int)(col % g.sx), j = (int)(col / g.sx);
        double r0 = 0.0, r1 = 0.0;
        if (i >= g.ilo && i <= g.ihi && j >= g.jlo && j <= g.jhi && av[col]) {
            double su = 0.0, sv = 0.0;
            const double *au = Auu + 27 * q, *auv = Auv + 27 * q, *avu = Avu + 27 * q, *avv = Avv + 27 * q;
#pragma unroll
            for (int m = 0; m < 27; ++m) {
                const int iA = m % 3 - 1, jA = (m / 3) % 3 - 1, kA = m / 9 - 1;
                if (k + kA < 0 || k + kA >= g.nz) continue;
                const long long qn = q + kA + iA * cst + jA * rst;
                const double x0 = xu[qn], x1 = xv[qn];
                su = su + au[m] * x0 + auv[m] * x1;
                sv = sv + avu[m] * x0 + avv[m] * x1;
            }
            r0 = su - bu[q];
            r1 = sv - bv[q];
            acc[0] += r0 * r0 + r1 * r1;
            acc[1] += bu[q] * bu[q] + bv[q] * bv[q];
        }
        ru[q] = r0;
        rv[q] = r1;
    }
    // fixed-order block reduction: waves by shuffles, then wave 0 over the waves
    __shared__ double red[NOALLOC_THREADS / 64][2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        double s = acc[c];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        if (lane == 0) red[wave][c] = s;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        double s = 0.0;
        for (int w = 0; w < NOALLOC_THREADS / 64; ++w) s += red[w][threadIdx.x];
        part[2 * blockIdx.x + threadIdx.x] = s;
    }
}

// one workgroup: the nb partial pairs in a fixed order -> norms[0] = |r|, norms[1] =
// |r| / |b| (0 when |b| = 0), the reference's L2_norm and L2_norm_relative
__global__ __launch_bounds__(NOALLOC_THREADS) void k_resvec_norms(int nb, const double *__restrict__ part,
                                                                  double *__restrict__ norms) {
    __shared__ double s[2][NOALLOC_THREADS];
    double a = 0.0, b = 0.0;
    for (int t = threadIdx.x; t < nb; t += NOALLOC_THREADS) {
        a += part[2 * t];
        b += part[2 * t + 1];
    }
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    __syncthreads();
    for (int w = NOALLOC_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s[0][threadIdx.x] += s[0][threadIdx.x + w];
            s[1][threadIdx.x] += s[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double L2 = sqrt(s[0][0]), rhs = sqrt(s[1][0]);
        norms[0] = L2;
        norms[1] = rhs > 0.0 ? L2 / rhs : 0.0;
    }
}

extern "C" int cism_compute_residual_vector_3d_noalloc(int nx, int ny, int nz, int periodic, int device,
                                                       void *stream, const int32_t *active_vertex,
                                                       const double *Auu, const double *Auv,
                                                       const double *Avu, const double *Avv,
                                                       const double *bu, const double *bv,
                                                       const double *uvel, const double *vvel,
                                                       double *resid_u, double *resid_v, double *norms) {
    // arguments first: nothing touches the GPU before they pass
    if (nx < 7 || ny < 7 || nz < 1 || device < 0 || device >= NOALLOC_MAX_DEVICES) {
        cism_set_error("residual_vector_3d_noalloc: invalid grid nx=%d ny=%d nz=%d or device %d", nx, ny, nz,
                       device);
        return CISM_ERR_ARG;
    }
    const long long nn = (long long)nz * (nx - 1) * (ny - 1);
    if (nn > NOALLOC_MAX_NODES) {
        cism_set_error("residual_vector_3d_noalloc: %lld nodes exceed the maximum of %lld", nn, NOALLOC_MAX_NODES);
        return CISM_ERR_ARG;
    }
    if (!active_vertex || !Auu || !Auv || !Avu || !Avv || !bu || !bv || !uvel || !vvel || !resid_u ||
        !resid_v || !norms) {
        cism_set_error("residual_vector_3d_noalloc: null array");
        return CISM_ERR_ARG;
    }
    HIP_TRY(hipSetDevice(device));
    double *part = nullptr;
    CISM_TRY(scratch_on_device(device, &part));
    FGrid g;
    g.nz = nz; g.sx = nx - 1; g.sy = ny - 1; g.nn = nn;
    // parallel_slap.F90:661-673 (0-based): outflow domains own the west/south boundary rows
    g.ilo = (periodic ? 3 : 2) - 1;
    g.jlo = (periodic ? 3 : 2) - 1;
    g.ihi = nx - 3;
    g.jhi = ny - 3;
    long long nb = (nn + NOALLOC_THREADS - 1) / NOALLOC_THREADS;
    if (nb > NOALLOC_BLOCKS) nb = NOALLOC_BLOCKS;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_resvec_fortran, dim3((unsigned)nb), dim3(NOALLOC_THREADS), 0, s, g, active_vertex, Auu,
                       Auv, Avu, Avv, bu, bv, uvel, vvel, resid_u, resid_v, part);
    hipLaunchKernelGGL(k_resvec_norms, dim3(1), dim3(NOALLOC_THREADS), 0, s, (int)nb, part, norms);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

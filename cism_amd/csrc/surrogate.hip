// surrogate.hip -- learned initial guess for the velocity solve (no reference
// counterpart): a fully convolutional 4-level encoder-decoder (~1.0M
// parameters) mapping 2D ice-sheet fields on the vertex grid to u, v at every
// level plus a per-column preconditioner scale, with hand-written forward and
// backward kernels and Adam.
//
// Precision: FP32 master weights and activations; every convolution runs as an
// implicit GEMM on gfx950's MFMA (v_mfma_f32_16x16x32_bf16) with BF16 operands and FP32
// accumulation.  Activations are NHWC (channels fastest); 3x3 convolutions use
// periodic (circular) padding, matching the periodic grids of the solver.
//
//   enc0  3x3 16->32, 32->32          (H x W; 5 input channels zero-padded to 16)
//   enc1  pool, 3x3 32->64, 64->64    (H/2)
//   enc2  pool, 3x3 64->128, 128->128 (H/4)
//   mid   pool, 3x3 128->128 x2       (H/8)
//   dec2  up, cat enc2, 3x3 256->128
//   dec1  up, cat enc1, 3x3 192->64
//   dec0  up, cat enc0, 3x3 96->32
//   head  1x1 32->32 (24 used: u, v at nz<=11 levels and two log-scales)
// ReLU after every convolution except the head.
#include "internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline short f2bf(float f) {
    // round to nearest even
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

#define SG_IN_RAW 5
#define SG_IN 16
#define SG_OUT 24
#define SG_OUT_PAD 32

// ---------------------------------------------------------------------------
// implicit-GEMM convolution: out[p][co] = act(bias[co] + sum_k W[co][k] * X(p, k))
// k = tap*Ci + ci, tap = ky*3 + kx (ks = 3) or 0 (ks = 1); X(p, k) = in at pixel p
// shifted by (ky-1, kx-1) with periodic wrap.  Requires Ci % 16 == 0.
// Block tile: 64 output channels x 64 pixels, 4 waves of 2x2 16x16 MFMA tiles.
// ---------------------------------------------------------------------------
struct ConvDims {
    int B, H, W, Ci, Co, ks, relu;
};

__device__ inline long long shifted_pixel(long long p, int tap, int ks, int H, int W) {
    if (ks == 1) return p;
    const int HW = H * W;
    const long long b = p / HW;
    const int r = (int)(p - b * HW);
    int y = r / W, x = r - (r / W) * W;
    y += tap / 3 - 1;
    x += tap % 3 - 1;
    y = y < 0 ? y + H : (y >= H ? y - H : y);
    x = x < 0 ? x + W : (x >= W ? x - W : x);
    return b * HW + (long long)y * W + x;
}

// K values staged in LDS per step of the forward convolution: 2 MFMA k-steps of 32
// between two barriers; a thread stages 16 consecutive k of one weight row and of one
// pixel (Ci % 16 == 0, so the 16 lie in one tap and their 3x3 shift is computed once).
// A K that is not a multiple of 32 (3x3 over 16 channels: 144) ends in staged zeros.
// Lane l of an MFMA holds row (l & 15) of A and column (l & 15) of B with 8 consecutive
// k from 8 (l >> 4); its 4 results are rows 4 (l >> 4) + i of column (l & 15).
#define CONV_CK 64

__global__ __launch_bounds__(256) void k_conv_mfma(ConvDims d, const float *__restrict__ in,
                                                   const short *__restrict__ Wb,
                                                   const float *__restrict__ bias,
                                                   float *__restrict__ out) {
    alignas(16) __shared__ short As[64][CONV_CK + 8];   // weights [co][k]
    alignas(16) __shared__ short Bs[64][CONV_CK + 8];   // inputs [pixel][k]
    const int K = d.ks * d.ks * d.Ci;
    const int HW = d.H * d.W;
    const long long P = (long long)d.B * HW;
    const int co0 = blockIdx.y * 64;
    const long long p0 = (long long)blockIdx.x * 64;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wc = (wave >> 1) * 32, wp = (wave & 1) * 32;
    f32x4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
    const int lr = t >> 2, lk = (t & 3) * 16;   // loader: row / pixel lr, 16 k values from lk
    const int co = co0 + lr;
    const long long p = p0 + lr;
    const bool pok = p < P, cok = co < d.Co;
    // the loader pixel's image coordinates, once
    long long pb = 0;
    int py = 0, px = 0;
    if (pok) {
        pb = p / HW;
        const int r = (int)(p - pb * HW);
        py = r / d.W;
        px = r - py * d.W;
    }
    for (int k0 = 0; k0 < K; k0 += CONV_CK) {
        const int k = k0 + lk;
        uint4 a0 = {0u, 0u, 0u, 0u}, a1 = {0u, 0u, 0u, 0u};
        if (cok && k < K) {
            const uint4 *src = reinterpret_cast<const uint4 *>(Wb + (long long)co * K + k);
            a0 = src[0];
            a1 = src[1];
        }
        *reinterpret_cast<uint4 *>(&As[lr][lk]) = a0;
        *reinterpret_cast<uint4 *>(&As[lr][lk + 8]) = a1;
        if (pok && k < K) {
            const int tap = k / d.Ci, ci = k - tap * d.Ci;
            long long src = p;
            if (d.ks == 3) {
                int y = py + tap / 3 - 1, x = px + tap % 3 - 1;
                y = y < 0 ? y + d.H : (y >= d.H ? y - d.H : y);
                x = x < 0 ? x + d.W : (x >= d.W ? x - d.W : x);
                src = pb * HW + (long long)y * d.W + x;
            }
            // 16 BF16 values packed two per word: two 16-byte LDS stores
            const float4 *v = reinterpret_cast<const float4 *>(in + src * d.Ci + ci);
            unsigned w[8];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float4 f = v[q];
                w[2 * q] = (unsigned)(unsigned short)f2bf(f.x) | ((unsigned)(unsigned short)f2bf(f.y) << 16);
                w[2 * q + 1] = (unsigned)(unsigned short)f2bf(f.z) | ((unsigned)(unsigned short)f2bf(f.w) << 16);
            }
            *reinterpret_cast<uint4 *>(&Bs[lr][lk]) = make_uint4(w[0], w[1], w[2], w[3]);
            *reinterpret_cast<uint4 *>(&Bs[lr][lk + 8]) = make_uint4(w[4], w[5], w[6], w[7]);
        } else {
            *reinterpret_cast<uint4 *>(&Bs[lr][lk]) = make_uint4(0u, 0u, 0u, 0u);
            *reinterpret_cast<uint4 *>(&Bs[lr][lk + 8]) = make_uint4(0u, 0u, 0u, 0u);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < CONV_CK; kk += 32) {
            if (k0 + kk >= K) break;   // the same for the whole block
            bf16x8 af[2], bfm[2];
            for (int a = 0; a < 2; ++a) {
                const int row = wc + a * 16 + (lane & 15);
                af[a] = *reinterpret_cast<const bf16x8 *>(&As[row][kk + 8 * (lane >> 4)]);
            }
            for (int b = 0; b < 2; ++b) {
                const int col = wp + b * 16 + (lane & 15);
                bfm[b] = *reinterpret_cast<const bf16x8 *>(&Bs[col][kk + 8 * (lane >> 4)]);
            }
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfm[b], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
    }
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int i = 0; i < 4; ++i) {
                const int oc = co0 + wc + a * 16 + 4 * (lane >> 4) + i;
                const long long op = p0 + wp + b * 16 + (lane & 15);
                if (oc < d.Co && op < P) {
                    float v = acc[a][b][i] + (bias ? bias[oc] : 0.f);
                    if (d.relu) v = v > 0.f ? v : 0.f;
                    out[op * d.Co + oc] = v;
                }
            }
}

// weight gradient: dW[co][k] = sum_p dY[p][co] * X(p, k); pixel range split over
// gridDim.z, partial tiles reduced by k_wgrad_reduce (fixed order).
__global__ __launch_bounds__(256) void k_wgrad_mfma(ConvDims d, const float *__restrict__ dY,
                                                    const float *__restrict__ X,
                                                    float *__restrict__ part, long long pchunk) {
    alignas(16) __shared__ short As[64][32 + 8];   // [co][pix]
    alignas(16) __shared__ short Bs[64][32 + 8];   // [k][pix]
    const int K = d.ks * d.ks * d.Ci;
    const long long P = (long long)d.B * d.H * d.W;
    const int co0 = blockIdx.y * 64, kq0 = blockIdx.x * 64;
    const long long pbeg = (long long)blockIdx.z * pchunk;
    const long long pend = pbeg + pchunk < P ? pbeg + pchunk : P;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wc = (wave >> 1) * 32, wk = (wave & 1) * 32;
    f32x4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
    const int lp = t >> 4, lc = (t & 15) * 4;     // loader: pixels lp and lp + 16, 4 values from lc
    // the B loader's tap and channel are fixed for the block
    const int kl = kq0 + lc;
    const int ltap = kl < K ? kl / d.Ci : 0, lci = kl < K ? kl - ltap * d.Ci : 0;
    for (long long q0 = pbeg; q0 < pend; q0 += 32) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int pl = lp + 16 * h;
            const long long p = q0 + pl;
            const bool okp = p < pend;
            {   // A: dY[p][co0+lc..+3] -> As[co][pix]
                float4 v = {0.f, 0.f, 0.f, 0.f};
                if (okp && co0 + lc < d.Co) v = *reinterpret_cast<const float4 *>(dY + p * d.Co + co0 + lc);
                As[lc + 0][pl] = f2bf(v.x);
                As[lc + 1][pl] = f2bf(v.y);
                As[lc + 2][pl] = f2bf(v.z);
                As[lc + 3][pl] = f2bf(v.w);
            }
            {   // B: X(p, kq0+lc..+3) -> Bs[k][pix]
                float4 v = {0.f, 0.f, 0.f, 0.f};
                if (okp && kl < K) {
                    const long long src = shifted_pixel(p, ltap, d.ks, d.H, d.W);
                    v = *reinterpret_cast<const float4 *>(X + src * d.Ci + lci);
                }
                Bs[lc + 0][pl] = f2bf(v.x);
                Bs[lc + 1][pl] = f2bf(v.y);
                Bs[lc + 2][pl] = f2bf(v.z);
                Bs[lc + 3][pl] = f2bf(v.w);
            }
        }
        __syncthreads();
        bf16x8 af[2], bfm[2];
        for (int a = 0; a < 2; ++a) {
            const int row = wc + a * 16 + (lane & 15);
            af[a] = *reinterpret_cast<const bf16x8 *>(&As[row][8 * (lane >> 4)]);
        }
        for (int b = 0; b < 2; ++b) {
            const int col = wk + b * 16 + (lane & 15);
            bfm[b] = *reinterpret_cast<const bf16x8 *>(&Bs[col][8 * (lane >> 4)]);
        }
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfm[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
    float *pz = part + (long long)blockIdx.z * d.Co * K;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int i = 0; i < 4; ++i) {
                const int co = co0 + wc + a * 16 + 4 * (lane >> 4) + i;
                const int k = kq0 + wk + b * 16 + (lane & 15);
                if (co < d.Co && k < K) pz[(long long)co * K + k] = acc[a][b][i];
            }
}

__global__ void k_wgrad_reduce(const float *__restrict__ part, int nz, long long n, float *__restrict__ dW) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float s = 0.f;
    for (int z = 0; z < nz; ++z) s += part[(long long)z * n + t];
    dW[t] = s;
}

// bias gradient: db[co] = sum_p dY[p][co] (one block per channel, fixed order)
__global__ __launch_bounds__(256) void k_bgrad(const float *__restrict__ dY, long long P, int Co,
                                               float *__restrict__ db) {
    __shared__ float red[256];
    const int co = blockIdx.x;
    float s = 0.f;
    for (long long p = threadIdx.x; p < P; p += 256) s += dY[p * Co + co];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[co] = red[0];
}

__global__ void k_relu_bwd(long long n, const float *__restrict__ y, float *__restrict__ dy) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n && !(y[t] > 0.f)) dy[t] = 0.f;
}

// 2x2 average pool and nearest 2x upsample, NHWC, with their adjoints
__global__ void k_pool(int B, int H, int W, int C, const float *__restrict__ in, float *__restrict__ out) {
    const long long n = (long long)B * (H / 2) * (W / 2) * C;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int c = (int)(t % C);
    long long r = t / C;
    const int x = (int)(r % (W / 2));
    r /= (W / 2);
    const int y = (int)(r % (H / 2));
    const long long b = r / (H / 2);
    auto at = [&](int yy, int xx) { return in[((b * H + yy) * W + xx) * C + c]; };
    out[t] = 0.25f * (at(2 * y, 2 * x) + at(2 * y, 2 * x + 1) + at(2 * y + 1, 2 * x) + at(2 * y + 1, 2 * x + 1));
}

__global__ void k_pool_bwd(int B, int H, int W, int C, const float *__restrict__ dout, float *__restrict__ din) {
    const long long n = (long long)B * H * W * C;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int c = (int)(t % C);
    long long r = t / C;
    const int x = (int)(r % W);
    r /= W;
    const int y = (int)(r % H);
    const long long b = r / H;
    din[t] = 0.25f * dout[((b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
}

// upsample a (H/2 x W/2, Ca) to H x W into channels [0, Ca) of out (Ctot) and copy
// skip (H x W, Cs) into channels [Ca, Ca + Cs): the decoder's up + concat
__global__ void k_up_cat(int B, int H, int W, int Ca, int Cs, const float *__restrict__ a,
                         const float *__restrict__ skip, float *__restrict__ out) {
    const int Ct = Ca + Cs;
    const long long n = (long long)B * H * W * Ct;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int c = (int)(t % Ct);
    const long long p = t / Ct;
    if (c < Ca) {
        const int x = (int)(p % W);
        const long long r = p / W;
        const int y = (int)(r % H);
        const long long b = r / H;
        out[t] = a[((b * (H / 2) + y / 2) * (W / 2) + x / 2) * Ca + c];
    } else {
        out[t] = skip[p * Cs + (c - Ca)];
    }
}

// adjoint of k_up_cat: da (H/2 x W/2, Ca) = sum of the 4 children; dskip += tail
__global__ void k_up_cat_bwd_a(int B, int H, int W, int Ca, int Cs, const float *__restrict__ dout,
                               float *__restrict__ da) {
    const int Ct = Ca + Cs, h = H / 2, w = W / 2;
    const long long n = (long long)B * h * w * Ca;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int c = (int)(t % Ca);
    long long r = t / Ca;
    const int x = (int)(r % w);
    r /= w;
    const int y = (int)(r % h);
    const long long b = r / h;
    float s = 0.f;
    for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) s += dout[((b * H + 2 * y + dy) * W + 2 * x + dx) * Ct + c];
    da[t] = s;
}

__global__ void k_up_cat_bwd_skip(long long P, int Ca, int Cs, const float *__restrict__ dout,
                                  float *__restrict__ dskip) {
    const long long n = P * Cs;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long long p = t / Cs;
    const int c = (int)(t % Cs);
    dskip[t] += dout[p * (Ca + Cs) + Ca + c];
}

// masked mean squared error over the SG_OUT used channels; out has SG_OUT_PAD channels
__global__ __launch_bounds__(256) void k_loss(long long P, int nch, const float *__restrict__ out,
                                              const float *__restrict__ tgt,
                                              const float *__restrict__ mask,
                                              const float *__restrict__ chw,
                                              double *__restrict__ partials) {
    __shared__ double red[256];
    double s = 0.0, m = 0.0;
    for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < P; p += (long long)gridDim.x * 256) {
        const float mk = mask[p];
        if (mk == 0.f) continue;
        m += mk;
        for (int c = 0; c < nch; ++c) {
            const double e = (double)out[p * SG_OUT_PAD + c] - tgt[p * SG_OUT + c];
            s += mk * chw[c] * e * e;
        }
    }
    for (int pass = 0; pass < 2; ++pass) {
        red[threadIdx.x] = pass == 0 ? s : m;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) partials[2 * blockIdx.x + pass] = red[0];
        __syncthreads();
    }
}

__global__ void k_loss_final(const double *__restrict__ partials, int nb, double *__restrict__ res) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0, m = 0.0;
    for (int b = 0; b < nb; ++b) { s += partials[2 * b]; m += partials[2 * b + 1]; }
    res[0] = m > 0 ? s / m : 0.0;
    res[1] = m;
}

__global__ void k_loss_grad(long long P, int nch, const float *__restrict__ out,
                            const float *__restrict__ tgt, const float *__restrict__ mask,
                            const float *__restrict__ chw, const double *__restrict__ res,
                            float *__restrict__ dout) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P * SG_OUT_PAD) return;
    const long long p = t / SG_OUT_PAD;
    const int c = (int)(t % SG_OUT_PAD);
    const double m = res[1];
    float g = 0.f;
    if (c < nch && m > 0) g = (float)(2.0 * mask[p] * chw[c] * ((double)out[t] - tgt[p * SG_OUT + c]) / m);
    dout[t] = g;
}

// Adam on FP32 masters
__global__ void k_adam(long long n, float *__restrict__ w, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gg = g[t];
    const float mm = b1 * m[t] + (1.f - b1) * gg;
    const float vv = b2 * v[t] + (1.f - b2) * gg * gg;
    m[t] = mm;
    v[t] = vv;
    w[t] = w[t] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// BF16 copy of the forward weights W[co][k] and of the data-gradient weights
// W'[ci][k'] (k' = tap'*Co + co, tap' = 8 - tap): dX = conv(dY, W') with the
// flipped kernel, periodic padding making it exact
__global__ void k_prep_weights(int Co, int Ci, int ks, const float *__restrict__ W,
                               short *__restrict__ Wb, short *__restrict__ Wt) {
    const int K = ks * ks * Ci;
    const long long n = (long long)Co * K;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int co = (int)(t / K), k = (int)(t % K);
    const int tap = k / Ci, ci = k % Ci;
    const short b = f2bf(W[t]);
    Wb[t] = b;
    const int tapf = ks == 3 ? 8 - tap : 0;
    Wt[(long long)ci * (ks * ks * Co) + (long long)tapf * Co + co] = b;
}

// ---------------------------------------------------------------------------
// network container
// ---------------------------------------------------------------------------
struct ConvLayer {
    int ks, Ci, Co, relu;
    long long woff, boff;   // offsets into the parameter vector
    short *Wb, *Wt;         // BF16 forward / data-gradient weights
};

struct Act {
    int H, W, C;
    float *y, *dy;
    long long n(int B) const { return (long long)B * H * W * C; }
};

struct cism_surrogate {
    int B, H, W;
    std::vector<ConvLayer> L;
    long long nparams;
    float *w, *g, *m, *v;   // params, grads, Adam moments
    long long step;
    // activations: 0 in; conv outputs and pool/cat buffers
    std::vector<Act> A;
    float *raw, *tgt, *mask, *chw, *wpart;
    long long wpart_n;
    double *lpart, *lres;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    float last_forward_ms;   // device time of the last forward pass (kernels only)
    int wprep_ok;            // the BF16 copies Wb, Wt match w (cleared whenever w changes)
};

// activation indices
enum {
    A_IN, A_E0a, A_E0b, A_P0, A_E1a, A_E1b, A_P1, A_E2a, A_E2b, A_P2, A_M0, A_M1, A_C2, A_D2,
    A_C1, A_D1, A_C0, A_D0, A_OUT, A_N
};
// layer indices
enum { L_E0a, L_E0b, L_E1a, L_E1b, L_E2a, L_E2b, L_M0, L_M1, L_D2, L_D1, L_D0, L_HEAD, L_N };

static const int LAYER_IN[L_N] = {A_IN, A_E0a, A_P0, A_E1a, A_P1, A_E2a, A_P2, A_M0, A_C2, A_C1, A_C0, A_D0};
static const int LAYER_OUT[L_N] = {A_E0a, A_E0b, A_E1a, A_E1b, A_E2a, A_E2b, A_M0, A_M1, A_D2, A_D1, A_D0, A_OUT};

static int sg_alloc(float **p, long long n) {
    hipError_t e = hipMalloc(p, sizeof(float) * (n > 0 ? n : 1));
    if (e != hipSuccess) {
        cism_set_error("surrogate allocation failed: %s", hipGetErrorString(e));
        return CISM_ERR_NOMEM;
    }
    // null-stream memset: complete it before anything runs on the non-blocking stream
    HIP_TRY(hipMemset(*p, 0, sizeof(float) * (n > 0 ? n : 1)));
    HIP_TRY(hipDeviceSynchronize());
    return CISM_OK;
}

extern "C" int cism_surrogate_destroy(cism_surrogate *s);

extern "C" int cism_surrogate_create(int H, int W, int batch, int device, cism_surrogate **out) {
    if (!out || H % 8 || W % 8 || H < 8 || W < 8 || batch < 1) {
        cism_set_error("surrogate: H and W must be positive multiples of 8 (got %d x %d)", H, W);
        return CISM_ERR_ARG;
    }
    HIP_TRY(hipSetDevice(device));
    cism_surrogate *s = new cism_surrogate();
    s->B = batch; s->H = H; s->W = W;
    const int spec[L_N][4] = {{3, SG_IN, 32, 1}, {3, 32, 32, 1}, {3, 32, 64, 1}, {3, 64, 64, 1},
                              {3, 64, 128, 1}, {3, 128, 128, 1}, {3, 128, 128, 1}, {3, 128, 128, 1},
                              {3, 256, 128, 1}, {3, 192, 64, 1}, {3, 96, 32, 1}, {1, 32, SG_OUT_PAD, 0}};
    long long off = 0;
    for (int l = 0; l < L_N; ++l) {
        ConvLayer c;
        c.ks = spec[l][0]; c.Ci = spec[l][1]; c.Co = spec[l][2]; c.relu = spec[l][3];
        c.woff = off;
        off += (long long)c.Co * c.ks * c.ks * c.Ci;
        c.boff = off;
        off += c.Co;
        c.Wb = c.Wt = nullptr;
        s->L.push_back(c);
    }
    s->nparams = off;
    const int dims[A_N][3] = {{H, W, SG_IN}, {H, W, 32}, {H, W, 32}, {H / 2, W / 2, 32},
                              {H / 2, W / 2, 64}, {H / 2, W / 2, 64}, {H / 4, W / 4, 64},
                              {H / 4, W / 4, 128}, {H / 4, W / 4, 128}, {H / 8, W / 8, 128},
                              {H / 8, W / 8, 128}, {H / 8, W / 8, 128}, {H / 4, W / 4, 256},
                              {H / 4, W / 4, 128}, {H / 2, W / 2, 192}, {H / 2, W / 2, 64},
                              {H, W, 96}, {H, W, 32}, {H, W, SG_OUT_PAD}};
    int st = CISM_OK;
    for (int a = 0; a < A_N; ++a) {
        Act x;
        x.H = dims[a][0]; x.W = dims[a][1]; x.C = dims[a][2];
        x.y = x.dy = nullptr;
        s->A.push_back(x);
    }
    for (int a = 0; a < A_N && !st; ++a) {
        st = sg_alloc(&s->A[a].y, s->A[a].n(batch));
        if (!st) st = sg_alloc(&s->A[a].dy, s->A[a].n(batch));
    }
    for (int l = 0; l < L_N && !st; ++l) {
        const long long nw = (long long)s->L[l].Co * s->L[l].ks * s->L[l].ks * s->L[l].Ci;
        if (hipMalloc(&s->L[l].Wb, sizeof(short) * nw) != hipSuccess ||
            hipMalloc(&s->L[l].Wt, sizeof(short) * nw) != hipSuccess)
            st = CISM_ERR_NOMEM;
    }
    if (!st) st = sg_alloc(&s->w, s->nparams);
    if (!st) st = sg_alloc(&s->g, s->nparams);
    if (!st) st = sg_alloc(&s->m, s->nparams);
    if (!st) st = sg_alloc(&s->v, s->nparams);
    const long long P = (long long)batch * H * W;
    if (!st) st = sg_alloc(&s->raw, P * SG_IN_RAW);
    if (!st) st = sg_alloc(&s->tgt, P * SG_OUT);
    if (!st) st = sg_alloc(&s->mask, P);
    if (!st) st = sg_alloc(&s->chw, SG_OUT);
    // weight-gradient partials: the largest Co*K times the pixel split
    s->wpart_n = 0;
    for (int l = 0; l < L_N; ++l) {
        const long long n = (long long)s->L[l].Co * s->L[l].ks * s->L[l].ks * s->L[l].Ci * 64;
        if (n > s->wpart_n) s->wpart_n = n;
    }
    if (!st) st = sg_alloc(&s->wpart, s->wpart_n);
    if (!st && hipMalloc(&s->lpart, sizeof(double) * 2 * 1024) != hipSuccess) st = CISM_ERR_NOMEM;
    if (!st && hipMalloc(&s->lres, sizeof(double) * 2) != hipSuccess) st = CISM_ERR_NOMEM;
    if (!st && hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) st = CISM_ERR_HIP;
    if (!st && (hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess))
        st = CISM_ERR_HIP;
    if (st) {
        cism_surrogate_destroy(s);
        return st;
    }
    // default initialisation: He-normal from a fixed LCG, zero biases, zero head
    std::vector<float> h(s->nparams, 0.f);
    unsigned long long rs = 0x9E3779B97F4A7C15ull;
    auto unif = [&]() {
        rs = rs * 6364136223846793005ull + 1442695040888963407ull;
        return ((rs >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < L_N; ++l) {
        const ConvLayer &c = s->L[l];
        const long long nw = (long long)c.Co * c.ks * c.ks * c.Ci;
        const double sd = l == L_HEAD ? 1e-3 : sqrt(2.0 / (c.ks * c.ks * c.Ci));
        for (long long i = 0; i < nw; ++i) {
            const double u1 = unif(), u2 = unif();
            h[c.woff + i] = (float)(sd * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
        }
    }
    float ones[SG_OUT];
    for (int c = 0; c < SG_OUT; ++c) ones[c] = 1.f;
    HIP_TRY(hipMemcpyAsync(s->chw, ones, sizeof ones, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->w, h.data(), sizeof(float) * s->nparams, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->step = 0;
    s->wprep_ok = 0;
    *out = s;
    return CISM_OK;
}

extern "C" int cism_surrogate_destroy(cism_surrogate *s) {
    if (!s) return CISM_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto &a : s->A) {
        if (a.y) (void)hipFree(a.y);
        if (a.dy) (void)hipFree(a.dy);
    }
    for (auto &l : s->L) {
        if (l.Wb) (void)hipFree(l.Wb);
        if (l.Wt) (void)hipFree(l.Wt);
    }
    void *ptrs[] = {s->w, s->g, s->m, s->v, s->raw, s->tgt, s->mask, s->chw, s->wpart, s->lpart, s->lres};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return CISM_OK;
}

extern "C" long long cism_surrogate_num_params(const cism_surrogate *s) { return s ? s->nparams : 0; }

extern "C" double cism_surrogate_last_forward_ms(const cism_surrogate *s) {
    return s ? (double)s->last_forward_ms : -1.0;
}

// Host transfers of the weights go through the network's own (non-blocking) stream and
// wait for it, so they are ordered after every kernel launched before them.
extern "C" int cism_surrogate_get_params(cism_surrogate *s, float *host) {
    HIP_TRY(hipMemcpyAsync(host, s->w, sizeof(float) * s->nparams, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return CISM_OK;
}

extern "C" int cism_surrogate_set_params(cism_surrogate *s, const float *host) {
    s->wprep_ok = 0;
    HIP_TRY(hipMemcpyAsync(s->w, host, sizeof(float) * s->nparams, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemsetAsync(s->m, 0, sizeof(float) * s->nparams, s->stream));
    HIP_TRY(hipMemsetAsync(s->v, 0, sizeof(float) * s->nparams, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->step = 0;
    return CISM_OK;
}

// BF16 copies of the weights, rebuilt only after the weights changed (creation,
// cism_surrogate_set_params, a training step)
static int prep_weights(cism_surrogate *s) {
    if (s->wprep_ok) return CISM_OK;
    for (auto &c : s->L) {
        const long long n = (long long)c.Co * c.ks * c.ks * c.Ci;
        hipLaunchKernelGGL(k_prep_weights, dim3(nblocks_for(n)), dim3(256), 0, s->stream, c.Co, c.Ci, c.ks,
                           s->w + c.woff, c.Wb, c.Wt);
    }
    HIP_TRY(hipGetLastError());
    s->wprep_ok = 1;
    return CISM_OK;
}

static int conv_launch(cism_surrogate *s, const ConvDims &d, const float *in, const short *Wb,
                       const float *bias, float *out) {
    if (d.Ci % 16 != 0 || (d.ks != 1 && d.ks != 3)) {
        cism_set_error("surrogate convolution: Ci %% 16 must be 0 and the kernel 1x1 or 3x3");
        return CISM_ERR_ARG;
    }
    const long long P = (long long)d.B * d.H * d.W;
    dim3 grid((unsigned)((P + 63) / 64), (unsigned)((d.Co + 63) / 64));
    hipLaunchKernelGGL(k_conv_mfma, grid, dim3(256), 0, s->stream, d, in, Wb, bias, out);
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

static int forward(cism_surrogate *s) {
    const int B = s->B;
    CISM_TRY(prep_weights(s));
    auto conv = [&](int l) -> int {
        const ConvLayer &c = s->L[l];
        const Act &in = s->A[LAYER_IN[l]], &out = s->A[LAYER_OUT[l]];
        ConvDims d = {B, in.H, in.W, c.Ci, c.Co, c.ks, c.relu};
        return conv_launch(s, d, in.y, c.Wb, s->w + c.boff, out.y);
    };
    auto pool = [&](int a_in, int a_out) {
        const Act &x = s->A[a_in];
        hipLaunchKernelGGL(k_pool, dim3(nblocks_for(s->A[a_out].n(B))), dim3(256), 0, s->stream, B, x.H, x.W,
                           x.C, x.y, s->A[a_out].y);
    };
    auto upcat = [&](int a_low, int a_skip, int a_out) {
        const Act &o = s->A[a_out];
        hipLaunchKernelGGL(k_up_cat, dim3(nblocks_for(o.n(B))), dim3(256), 0, s->stream, B, o.H, o.W,
                           s->A[a_low].C, s->A[a_skip].C, s->A[a_low].y, s->A[a_skip].y, o.y);
    };
    CISM_TRY(conv(L_E0a)); CISM_TRY(conv(L_E0b)); pool(A_E0b, A_P0);
    CISM_TRY(conv(L_E1a)); CISM_TRY(conv(L_E1b)); pool(A_E1b, A_P1);
    CISM_TRY(conv(L_E2a)); CISM_TRY(conv(L_E2b)); pool(A_E2b, A_P2);
    CISM_TRY(conv(L_M0)); CISM_TRY(conv(L_M1));
    upcat(A_M1, A_E2b, A_C2); CISM_TRY(conv(L_D2));
    upcat(A_D2, A_E1b, A_C1); CISM_TRY(conv(L_D1));
    upcat(A_D1, A_E0b, A_C0); CISM_TRY(conv(L_D0));
    CISM_TRY(conv(L_HEAD));
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

// backward of layer l given dY of its output (already ReLU-masked); adds dX into the
// input activation's gradient when want_dx
static int conv_backward(cism_surrogate *s, int l, bool want_dx) {
    const int B = s->B;
    const ConvLayer &c = s->L[l];
    const Act &in = s->A[LAYER_IN[l]], &out = s->A[LAYER_OUT[l]];
    const long long P = (long long)B * in.H * in.W;
    const int K = c.ks * c.ks * c.Ci;
    // weight gradient
    int nzs = (int)((P + 4095) / 4096);
    if (nzs > 64) nzs = 64;
    const long long pchunk = ((P + nzs - 1) / nzs + 15) / 16 * 16;
    ConvDims d = {B, in.H, in.W, c.Ci, c.Co, c.ks, 0};
    dim3 grid((unsigned)((K + 63) / 64), (unsigned)((c.Co + 63) / 64), (unsigned)nzs);
    hipLaunchKernelGGL(k_wgrad_mfma, grid, dim3(256), 0, s->stream, d, out.dy, in.y, s->wpart, pchunk);
    const long long nw = (long long)c.Co * K;
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(nblocks_for(nw)), dim3(256), 0, s->stream, s->wpart, nzs, nw,
                       s->g + c.woff);
    hipLaunchKernelGGL(k_bgrad, dim3(c.Co), dim3(256), 0, s->stream, out.dy, P, c.Co, s->g + c.boff);
    HIP_TRY(hipGetLastError());
    if (want_dx) {
        // dX = conv(dY, W') : channels Co -> Ci
        ConvDims dd = {B, in.H, in.W, c.Co, c.Ci, c.ks, 0};
        CISM_TRY(conv_launch(s, dd, out.dy, c.Wt, nullptr, in.dy));
    }
    return CISM_OK;
}

static int backward(cism_surrogate *s) {
    const int B = s->B;
    auto relu_bwd = [&](int a) {
        const Act &x = s->A[a];
        hipLaunchKernelGGL(k_relu_bwd, dim3(nblocks_for(x.n(B))), dim3(256), 0, s->stream, x.n(B), x.y, x.dy);
    };
    auto upcat_bwd = [&](int a_low, int a_skip, int a_cat) {
        // d(low) = sum of children of the first channels; d(skip) += tail channels
        const Act &o = s->A[a_cat];
        hipLaunchKernelGGL(k_up_cat_bwd_a, dim3(nblocks_for(s->A[a_low].n(B))), dim3(256), 0, s->stream, B,
                           o.H, o.W, s->A[a_low].C, s->A[a_skip].C, o.dy, s->A[a_low].dy);
        hipLaunchKernelGGL(k_up_cat_bwd_skip, dim3(nblocks_for(s->A[a_skip].n(B))), dim3(256), 0, s->stream,
                           (long long)B * o.H * o.W, s->A[a_low].C, s->A[a_skip].C, o.dy, s->A[a_skip].dy);
    };
    // head: dOUT set by the loss
    CISM_TRY(conv_backward(s, L_HEAD, true));          // -> dD0
    relu_bwd(A_D0); CISM_TRY(conv_backward(s, L_D0, true));   // -> dC0
    upcat_bwd(A_D1, A_E0b, A_C0);                       // dD1 = up^T, dE0b (skip) = tail (added below)
    relu_bwd(A_D1); CISM_TRY(conv_backward(s, L_D1, true));   // -> dC1
    upcat_bwd(A_D2, A_E1b, A_C1);
    relu_bwd(A_D2); CISM_TRY(conv_backward(s, L_D2, true));   // -> dC2
    upcat_bwd(A_M1, A_E2b, A_C2);
    relu_bwd(A_M1); CISM_TRY(conv_backward(s, L_M1, true));   // -> dM0
    relu_bwd(A_M0); CISM_TRY(conv_backward(s, L_M0, true));   // -> dP2
    // dE2b = pool^T dP2 + skip part; the skip part was added by upcat_bwd into dE2b
    // before pool_bwd would overwrite it, so accumulate via a temporary: pool_bwd
    // writes into dE2a's buffer first (free at this point), then add
    {
        const Act &x = s->A[A_E2b];
        float *tmp = s->A[A_E2a].dy;
        hipLaunchKernelGGL(k_pool_bwd, dim3(nblocks_for(x.n(B))), dim3(256), 0, s->stream, B, x.H, x.W, x.C,
                           s->A[A_P2].dy, tmp);
        hipLaunchKernelGGL(k_up_cat_bwd_skip, dim3(nblocks_for(x.n(B))), dim3(256), 0, s->stream,
                           (long long)B * x.H * x.W, 0, x.C, tmp, x.dy);
    }
    relu_bwd(A_E2b); CISM_TRY(conv_backward(s, L_E2b, true));   // -> dE2a
    relu_bwd(A_E2a); CISM_TRY(conv_backward(s, L_E2a, true));   // -> dP1
    {
        const Act &x = s->A[A_E1b];
        float *tmp = s->A[A_E1a].dy;
        hipLaunchKernelGGL(k_pool_bwd, dim3(nblocks_for(x.n(B))), dim3(256), 0, s->stream, B, x.H, x.W, x.C,
                           s->A[A_P1].dy, tmp);
        hipLaunchKernelGGL(k_up_cat_bwd_skip, dim3(nblocks_for(x.n(B))), dim3(256), 0, s->stream,
                           (long long)B * x.H * x.W, 0, x.C, tmp, x.dy);
    }
    relu_bwd(A_E1b); CISM_TRY(conv_backward(s, L_E1b, true));
    relu_bwd(A_E1a); CISM_TRY(conv_backward(s, L_E1a, true));   // -> dP0
    {
        const Act &x = s->A[A_E0b];
        float *tmp = s->A[A_E0a].dy;
        hipLaunchKernelGGL(k_pool_bwd, dim3(nblocks_for(x.n(B))), dim3(256), 0, s->stream, B, x.H, x.W, x.C,
                           s->A[A_P0].dy, tmp);
        hipLaunchKernelGGL(k_up_cat_bwd_skip, dim3(nblocks_for(x.n(B))), dim3(256), 0, s->stream,
                           (long long)B * x.H * x.W, 0, x.C, tmp, x.dy);
    }
    relu_bwd(A_E0b); CISM_TRY(conv_backward(s, L_E0b, true));
    relu_bwd(A_E0a); CISM_TRY(conv_backward(s, L_E0a, false));
    HIP_TRY(hipGetLastError());
    return CISM_OK;
}

static int zero_skip_grads(cism_surrogate *s) {
    // the skip activations accumulate gradient from two paths
    const int sk[3] = {A_E0b, A_E1b, A_E2b};
    for (int a : sk) HIP_TRY(hipMemsetAsync(s->A[a].dy, 0, sizeof(float) * s->A[a].n(s->B), s->stream));
    return CISM_OK;
}

// inputs arrive with SG_IN_RAW channels; the first layer reads SG_IN (zero padded)
__global__ void k_widen(long long P, const float *__restrict__ raw, float *__restrict__ in) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P * SG_IN) return;
    const long long p = t / SG_IN;
    const int c = (int)(t % SG_IN);
    in[t] = c < SG_IN_RAW ? raw[p * SG_IN_RAW + c] : 0.f;
}

static int upload_batch(cism_surrogate *s, const float *in, const float *tgt, const float *mask) {
    const long long P = (long long)s->B * s->H * s->W;
    HIP_TRY(hipMemcpyAsync(s->raw, in, sizeof(float) * P * SG_IN_RAW, hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_widen, dim3(nblocks_for(P * SG_IN)), dim3(256), 0, s->stream, P, s->raw,
                       s->A[A_IN].y);
    if (tgt) HIP_TRY(hipMemcpyAsync(s->tgt, tgt, sizeof(float) * P * SG_OUT, hipMemcpyHostToDevice, s->stream));
    if (mask) HIP_TRY(hipMemcpyAsync(s->mask, mask, sizeof(float) * P, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return CISM_OK;
}

static int loss_and_grad(cism_surrogate *s, int nch, double *loss) {
    const long long P = (long long)s->B * s->H * s->W;
    int nb = (int)((P + 255) / 256);
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, s->stream, P, nch, s->A[A_OUT].y, s->tgt, s->mask,
                       s->chw, s->lpart);
    hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(1), 0, s->stream, s->lpart, nb, s->lres);
    hipLaunchKernelGGL(k_loss_grad, dim3(nblocks_for(P * SG_OUT_PAD)), dim3(256), 0, s->stream, P, nch,
                       s->A[A_OUT].y, s->tgt, s->mask, s->chw, s->lres, s->A[A_OUT].dy);
    HIP_TRY(hipGetLastError());
    double res[2];
    HIP_TRY(hipMemcpyAsync(res, s->lres, sizeof res, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (loss) *loss = res[0];
    return CISM_OK;
}

__global__ void k_narrow(long long P, const float *__restrict__ o, float *__restrict__ out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P * SG_OUT) return;
    out[t] = o[(t / SG_OUT) * SG_OUT_PAD + t % SG_OUT];
}

// forward on host inputs (B x H x W x 5, NHWC) -> host outputs (B x H x W x 24)
extern "C" int cism_surrogate_forward(cism_surrogate *s, const float *in, float *out) {
    if (!s) return CISM_ERR_ARG;
    CISM_TRY(upload_batch(s, in, nullptr, nullptr));
    HIP_TRY(hipEventRecord(s->ev0, s->stream));
    CISM_TRY(forward(s));
    HIP_TRY(hipEventRecord(s->ev1, s->stream));
    const long long P = (long long)s->B * s->H * s->W;
    // narrow SG_OUT_PAD -> SG_OUT channels into the target buffer (free during inference)
    hipLaunchKernelGGL(k_narrow, dim3(nblocks_for(P * SG_OUT)), dim3(256), 0, s->stream, P, s->A[A_OUT].y,
                       s->tgt);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, s->tgt, sizeof(float) * P * SG_OUT, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipEventElapsedTime(&s->last_forward_ms, s->ev0, s->ev1));
    return CISM_OK;
}

// loss and gradient w.r.t. all parameters (grad may be NULL) without an update
extern "C" int cism_surrogate_loss_grad(cism_surrogate *s, const float *in, const float *tgt,
                                        const float *mask, int nch, double *loss, float *grad) {
    if (!s || nch < 1 || nch > SG_OUT) return CISM_ERR_ARG;
    CISM_TRY(upload_batch(s, in, tgt, mask));
    CISM_TRY(forward(s));
    CISM_TRY(loss_and_grad(s, nch, loss));
    CISM_TRY(zero_skip_grads(s));
    CISM_TRY(backward(s));
    if (grad) {
        HIP_TRY(hipMemcpyAsync(grad, s->g, sizeof(float) * s->nparams, hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
    }
    return CISM_OK;
}

extern "C" int cism_surrogate_train_step(cism_surrogate *s, const float *in, const float *tgt,
                                         const float *mask, int nch, float lr, double *loss) {
    CISM_TRY(cism_surrogate_loss_grad(s, in, tgt, mask, nch, loss, nullptr));
    s->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)s->step), c2 = 1.f - powf(b2, (float)s->step);
    hipLaunchKernelGGL(k_adam, dim3(nblocks_for(s->nparams)), dim3(256), 0, s->stream, s->nparams, s->w, s->g,
                       s->m, s->v, lr, b1, b2, 1e-8f, c1, c2);
    s->wprep_ok = 0;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    return CISM_OK;
}

extern "C" int cism_surrogate_set_channel_weights(cism_surrogate *s, const float *w24) {
    HIP_TRY(hipMemcpyAsync(s->chw, w24, sizeof(float) * SG_OUT, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return CISM_OK;
}

// ---------------------------------------------------------------------------
// The learned initial guess on a solver context (cism_learned_guess*): features
// from the context's resident fields, the network, and the decoded velocity written
// into the context's x, all on the device.
//
// Image layout: pixel (y, x) is owned vertex (i, j) = (x + ilo, y + jlo).  A periodic
// grid's owned vertices are the whole image (their count must be a multiple of 8 in
// each direction, the network's periodic padding then matches the grid's).  An
// outflow grid's owned vertices sit in the corner of a canvas at least 8 pixels wider
// and taller, rounded up to a multiple of 8, whose other pixels carry zero features
// (no ice) and separate the domain's edges across the periodic wrap.
// ---------------------------------------------------------------------------
#define SG_V_CH 11             // output channels: u at levels 0..10, v at 11..21
#define SG_NZ_MAX 11

struct cism_guess {
    cism_surrogate *net;
    int W, H, Wown, Hown, i0, j0;
    int nz, encoding;          // encoding 0: v = uscale o; 1: v = sign(o) uscale expm1(|o|)
    double uscale;
};

// features of one pixel (stagthck / 1000, 100 ds/dx, 100 ds/dy, (ln max(beta, 1e-3) - 7) / 2,
// log10 of the column-mean flow factor + 16; zero at inactive vertices and off the grid),
// the column-mean A of each of the vertex's four cells from flwafact = A^(-1/3) / 2
__global__ __launch_bounds__(256) void k_guess_features(GridDev g, int W, int H, int Wown, int Hown, int i0,
                                                        int j0, const int32_t *__restrict__ av,
                                                        const double *__restrict__ sthck,
                                                        const double *__restrict__ dsdx,
                                                        const double *__restrict__ dsdy,
                                                        const double *__restrict__ beta,
                                                        const double *__restrict__ ff, float *__restrict__ in) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)W * H) return;
    const int y = (int)(p / W), x = (int)(p % W);
    float f[SG_IN_RAW] = {0.f, 0.f, 0.f, 0.f, 0.f};
    if (x < Wown && y < Hown) {
        const int i = x + i0, j = y + j0;
        const long long v = i + (long long)g.sx * j;
        if (av[v]) {
            double vA = 0.0;
            for (int dj = 0; dj < 2; ++dj)
                for (int di = 0; di < 2; ++di) {
                    const long long cell = (i + di) + (long long)g.nx * (j + dj);
                    double colA = 0.0;
                    for (int k = 0; k < g.nz - 1; ++k) {
                        const double t = ff[k + (long long)(g.nz - 1) * cell];
                        colA += t > 0.0 ? 1.0 / (8.0 * t * t * t) : 0.0;
                    }
                    vA += colA / (g.nz - 1);
                }
            vA *= 0.25;
            f[0] = (float)(sthck[v] / 1000.0);
            f[1] = (float)(dsdx[v] * 100.0);
            f[2] = (float)(dsdy[v] * 100.0);
            f[3] = (float)((log(fmax(beta[v], 1e-3)) - 7.0) / 2.0);
            f[4] = (float)(log10(fmax(vA, 1e-30)) + 16.0);
        }
    }
#pragma unroll
    for (int c = 0; c < SG_IN; ++c) in[p * SG_IN + c] = c < SG_IN_RAW ? f[c] : 0.f;
}

// largest |output| decoded (network units), as the host decoder clips
__device__ inline double guess_decode(float o, int encoding, double uscale) {
    const double lim = encoding == 1 ? 10.0 : 1000.0;
    const double t = fmin(lim, fmax(-lim, (double)o));
    if (encoding == 1) return (t < 0 ? -uscale : uscale) * expm1(fabs(t));
    return uscale * t;
}

// x = the decoded velocity at owned active vertices, 0 elsewhere (halo images follow)
__global__ __launch_bounds__(256) void k_guess_decode(GridDev g, int W, int Wown, int Hown, int i0, int j0,
                                                      int encoding, double uscale,
                                                      const int32_t *__restrict__ av,
                                                      const float *__restrict__ out, double *__restrict__ x) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nn) return;
    const int k = (int)(q % g.nz);
    const long long col = q / g.nz;
    const int i = (int)(col % g.sx), j = (int)(col / g.sx);
    const int px = i - i0, py = j - j0;
    double u = 0.0, v = 0.0;
    if (px >= 0 && px < Wown && py >= 0 && py < Hown && av[col]) {
        const long long p = (long long)py * W + px;
        u = guess_decode(out[p * SG_OUT_PAD + k], encoding, uscale);
        v = guess_decode(out[p * SG_OUT_PAD + SG_V_CH + k], encoding, uscale);
    }
    x[2 * q] = u;
    x[2 * q + 1] = v;
}

void learned_guess_free(cism_ctx *c) {
    if (!c || !c->guess) return;
    cism_surrogate_destroy(c->guess->net);
    delete c->guess;
    c->guess = nullptr;
}

int learned_guess_device(cism_ctx *c) {
    cism_guess *gs = c->guess;
    if (!gs) {
        cism_set_error("learned guess: no network loaded (cism_learned_guess_load)");
        return CISM_ERR_ARG;
    }
    const GridDev &g = c->g;
    cism_surrogate *s = gs->net;
    const long long P = (long long)gs->W * gs->H;
    hipLaunchKernelGGL(k_guess_features, dim3(nblocks_for(P)), dim3(256), 0, c->stream, g, gs->W, gs->H, gs->Wown,
                       gs->Hown, gs->i0, gs->j0, c->av, c->sthck, c->dsdx, c->dsdy, c->beta, c->flwafact,
                       s->A[A_IN].y);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipEventRecord(s->ev0, s->stream));
    CISM_TRY(forward(s));
    HIP_TRY(hipEventRecord(s->ev1, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipEventElapsedTime(&s->last_forward_ms, s->ev0, s->ev1));
    hipLaunchKernelGGL(k_guess_decode, dim3(nblocks_for(g.nn)), dim3(256), 0, c->stream, g, gs->W, gs->Wown,
                       gs->Hown, gs->i0, gs->j0, gs->encoding, gs->uscale, c->av, s->A[A_OUT].y, c->x);
    HIP_TRY(hipGetLastError());
    return launch_halo_vec(c, c->x);
}

extern "C" int cism_learned_guess_load(cism_ctx *c, const float *params, long long nparams, int nz, int encoding,
                                       double uscale) {
    if (!c || !params || (encoding != 0 && encoding != 1) || !(uscale > 0)) {
        cism_set_error("learned_guess_load: invalid arguments");
        return CISM_ERR_ARG;
    }
    HIP_TRY(hipSetDevice(c->device));
    const GridDev &g = c->g;
    if (nz != g.nz || nz > SG_NZ_MAX) {
        // the output channels hold nz levels of u then of v from channel 11: another nz
        // would read the wrong channels
        cism_set_error("learned_guess_load: the weights predict %d levels, the grid has nz = %d (at most %d)", nz,
                       g.nz, SG_NZ_MAX);
        return CISM_ERR_ARG;
    }
    const int Wown = g.ihi - g.ilo + 1, Hown = g.jhi - g.jlo + 1;
    int W = Wown, H = Hown;
    if (g.periodic) {
        if (Wown % 8 || Hown % 8) {
            cism_set_error("learned_guess_load: a periodic grid needs ewn and nsn divisible by 8 (got %d x %d)",
                           Wown, Hown);
            return CISM_ERR_UNSUPPORTED;
        }
    } else {
        W = (Wown + 8 + 7) / 8 * 8;
        H = (Hown + 8 + 7) / 8 * 8;
    }
    learned_guess_free(c);
    cism_surrogate *s = nullptr;
    CISM_TRY(cism_surrogate_create(H, W, 1, c->device, &s));
    if (nparams != s->nparams) {
        cism_surrogate_destroy(s);
        cism_set_error("learned_guess_load: %lld parameters given, the network has %lld", nparams, s->nparams);
        return CISM_ERR_ARG;
    }
    const int st = cism_surrogate_set_params(s, params);
    if (st) {
        cism_surrogate_destroy(s);
        return st;
    }
    cism_guess *gs = new cism_guess();
    gs->net = s;
    gs->W = W; gs->H = H; gs->Wown = Wown; gs->Hown = Hown; gs->i0 = g.ilo; gs->j0 = g.jlo;
    gs->nz = nz; gs->encoding = encoding; gs->uscale = uscale;
    c->guess = gs;
    return CISM_OK;
}

// weight file: "CISMSG01", int32 nz, int32 encoding, float64 uscale, int64 nparams, then
// nparams float32 (little endian)
extern "C" int cism_learned_guess_load_file(cism_ctx *c, const char *path) {
    if (!c || !path) {
        cism_set_error("learned_guess_load_file: invalid arguments");
        return CISM_ERR_ARG;
    }
    FILE *f = fopen(path, "rb");
    if (!f) {
        cism_set_error("learned_guess_load_file: cannot open %s", path);
        return CISM_ERR_ARG;
    }
    char magic[8];
    int32_t nz = 0, enc = 0;
    double uscale = 0.0;
    long long n = 0;
    int ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "CISMSG01", 8) == 0 && fread(&nz, 4, 1, f) == 1 &&
             fread(&enc, 4, 1, f) == 1 && fread(&uscale, 8, 1, f) == 1 && fread(&n, 8, 1, f) == 1 && n > 0 &&
             n < (1LL << 28);
    std::vector<float> w;
    if (ok) {
        w.resize((size_t)n);
        ok = fread(w.data(), sizeof(float), (size_t)n, f) == (size_t)n;
    }
    fclose(f);
    if (!ok) {
        cism_set_error("learned_guess_load_file: %s is not a weight file", path);
        return CISM_ERR_ARG;
    }
    return cism_learned_guess_load(c, w.data(), n, nz, enc, uscale);
}

extern "C" int cism_learned_guess(cism_ctx *c, const double *stagthck, const double *dusrf_dx,
                                  const double *dusrf_dy, const double *beta_internal, const double *flwafact,
                                  const int32_t *active_vertex, double *uvel, double *vvel) {
    if (!c) {
        cism_set_error("learned_guess: null context");
        return CISM_ERR_ARG;
    }
    HIP_TRY(hipSetDevice(c->device));
    const GridDev &g = c->g;
    const size_t ncol = (size_t)g.sx * g.sy;
    struct { double *dst; const double *src; size_t n; } ins[] = {
        {c->sthck, stagthck, ncol}, {c->dsdx, dusrf_dx, ncol}, {c->dsdy, dusrf_dy, ncol},
        {c->beta, beta_internal, ncol}, {c->flwafact, flwafact, (size_t)g.ne}};
    for (auto &t : ins)
        if (t.src) HIP_TRY(hipMemcpyAsync(t.dst, t.src, sizeof(double) * t.n, hipMemcpyHostToDevice, c->stream));
    if (active_vertex)
        HIP_TRY(hipMemcpyAsync(c->av, active_vertex, sizeof(int32_t) * ncol, hipMemcpyHostToDevice, c->stream));
    CISM_TRY(learned_guess_device(c));
    if (uvel && vvel) CISM_TRY(download_vec(c, c->x, uvel, vvel));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CISM_OK;
}

extern "C" double cism_learned_guess_last_ms(const cism_ctx *c) {
    return c && c->guess ? (double)c->guess->net->last_forward_ms : 0.0;
}

// unet_kernels.hip -- kernels of the learned FWI preconditioner (3-D U-Net).
//
// Tensors are channels-last: element (b, z, y, x, c) of a [B][D][H][W][C]
// tensor sits at ((((b*D + z)*H + y)*W + x)*C + c).  Activations are BF16,
// gradients FP32; every reduction accumulates in FP32 on the matrix cores.
//
// 3x3x3 convolution ("same" zero padding) as an implicit GEMM on gfx950's
// v_mfma_f32_16x16x32_bf16:  M = 16 voxels along x, N = 16 output channels
// per accumulator, K = 27 taps x Cin channels, flattened as k = tap*Cin + ci
// (tap = (dz+1)*9 + (dy+1)*3 + (dx+1)).  Operand layout of the instruction
// (checked on the GPU): lane l holds A[m = l%16][k = (l/16)*8 .. +7],
// B[k = (l/16)*8 .. +7][n = l%16], and D[m = (l/16)*4 .. +3][n = l%16].
// Cin is a multiple of 8, so every 8-wide k group lies inside one tap and
// one input tensor: a lane's A fragment is one 16-byte channel vector.
//
//  conv_fwd     out = act(W * concat(inA, inB) + bias)       (forward)
//  conv_fwd     dIn = W' * dOut   with W' = taps flipped, channels swapped
//               (data gradient; same kernel, FP32 output, accumulating)
//  conv_wgrad   dW[co][tap*Cin+ci] += sum_v dOut[v][co] * in[v+tap][ci]
//               (M = 16 co, N = 16 ci, K = 32 voxels along x per MFMA)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "unet.hpp"

namespace fwi {
namespace unet {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

namespace {

__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int j = 0; j < 8; ++j) z[j] = (__bf16)0.0f;
    return z;
}

__device__ __forceinline__ bf16x8 load8(const bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ __forceinline__ bf16x8 load8(const float* p) {
    const float4 a = *reinterpret_cast<const float4*>(p);
    const float4 b = *reinterpret_cast<const float4*>(p + 4);
    bf16x8 r;
    r[0] = (__bf16)a.x; r[1] = (__bf16)a.y; r[2] = (__bf16)a.z; r[3] = (__bf16)a.w;
    r[4] = (__bf16)b.x; r[5] = (__bf16)b.y; r[6] = (__bf16)b.z; r[7] = (__bf16)b.w;
    return r;
}

// out mode: 0 = BF16 activation (bias, optional ReLU), 1 = FP32 store,
// 2 = FP32 accumulate (+=)
template <typename TA, int NB, int OUT>
__global__ void __launch_bounds__(256) conv_fwd_kernel(const TA* __restrict__ inA, int Ca,
                                                       const bf16* __restrict__ inB, int Cb,
                                                       const bf16* __restrict__ wt, int Kpad,
                                                       const float* __restrict__ bias, int Cout, int relu,
                                                       void* __restrict__ out, int B, int D, int H, int W) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = lane >> 4, m = lane & 15;
    const int x0 = blockIdx.x * 16;
    const int y = blockIdx.y * 4 + wave;
    const int bz = blockIdx.z;
    const int z = bz % D, b = bz / D;
    if (y >= H) return;  // uniform per wave; no barrier in this kernel
    const int Cin = Ca + Cb;
    const int K = 27 * Cin;
    floatx4 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[nb] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

    for (int k0 = 0; k0 < Kpad; k0 += 32) {
        const int kk = k0 + g * 8;
        bf16x8 a = zero8();
        if (kk < K) {
            const int tap = kk / Cin, ci = kk - tap * Cin;
            const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
            const int xx = x0 + m + dx, yy = y + dy, zz = z + dz;
            if (xx >= 0 && xx < W && yy >= 0 && yy < H && zz >= 0 && zz < D) {
                const long v = (((long)b * D + zz) * H + yy) * W + xx;
                if (ci < Ca) a = load8(inA + v * Ca + ci);
                else a = load8(inB + v * Cb + (ci - Ca));
            }
        }
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const bf16x8 w = load8(wt + (long)(nb * 16 + m) * Kpad + kk);
            acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w, acc[nb], 0, 0, 0);
        }
    }

#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        const int n = nb * 16 + m;
        if (n >= Cout) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int x = x0 + g * 4 + i;
            if (x >= W) continue;
            const long o = ((((long)b * D + z) * H + y) * W + x) * Cout + n;
            float v = acc[nb][i];
            if constexpr (OUT == 0) {
                v += bias[n];
                if (relu) v = fmaxf(v, 0.0f);
                static_cast<bf16*>(out)[o] = (bf16)v;
            } else if constexpr (OUT == 1) {
                static_cast<float*>(out)[o] = v;
            } else {
                static_cast<float*>(out)[o] += v;
            }
        }
    }
}

// dW[co][tap*Cin + ci] += sum over voxels of dOut[v][co] * concat(A,B)[v + tap][ci]
template <typename TI>
__global__ void __launch_bounds__(256) conv_wgrad_kernel(const float* __restrict__ dout, int Cout,
                                                         const TI* __restrict__ inA, int Ca,
                                                         const bf16* __restrict__ inB, int Cb,
                                                         float* __restrict__ part, int Kpad, int B, int D, int H,
                                                         int W, int rows_per_block) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = lane >> 4, m = lane & 15;
    const int Cin = Ca + Cb;
    const int ncib = (Cin + 15) / 16, ncob = Cout / 16;
    int t = blockIdx.x;
    const int tap = t % 27;
    t /= 27;
    const int cib = t % ncib;
    const int cob = t / ncib;
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const int co = cob * 16 + m;      // A row of this lane
    const int ci = cib * 16 + m;      // B column of this lane
    const int nrows = B * D * H;
    const int r0 = blockIdx.y * rows_per_block;
    const int r1 = min(r0 + rows_per_block, nrows);
    floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int r = r0 + wave; r < r1; r += 4) {
        const int y = r % H, zb = r / H, z = zb % D, b = zb / D;
        const int yy = y + dy, zz = z + dz;
        const bool rowok = yy >= 0 && yy < H && zz >= 0 && zz < D && ci < Cin;
        const long vrow = (((long)b * D + z) * H + y) * W;
        const long srow = (((long)b * D + zz) * H + yy) * W;
        for (int x0 = 0; x0 < W; x0 += 32) {
            bf16x8 a, bb;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int x = x0 + g * 8 + j;
                a[j] = (__bf16)(x < W ? dout[(vrow + x) * Cout + co] : 0.0f);
                const int xs = x + dx;
                float s = 0.0f;
                if (rowok && x < W && xs >= 0 && xs < W) {
                    if (ci < Ca) s = (float)inA[(srow + xs) * Ca + ci];
                    else s = (float)inB[(srow + xs) * Cb + (ci - Ca)];
                }
                bb[j] = (__bf16)s;
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc, 0, 0, 0);
        }
    }
    // D[m = co][n = ci]: lane holds rows g*4 + i, column m.  The four
    // waves' tiles, summed in wave order, are this block's partial sum for
    // its row chunk (blockIdx.y); wgrad_reduce_kernel adds the chunks in
    // order, so the result does not depend on scheduling.
    __shared__ float red[4][4][64];
#pragma unroll
    for (int i = 0; i < 4; ++i) red[wave][i][lane] = acc[i];
    __syncthreads();
    const int cw = cib * 16 + m;
    if (wave == 0 && cw < Cin) {
        float* pb = part + (long)blockIdx.y * Cout * Kpad;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int cr = cob * 16 + g * 4 + i;
            pb[(long)cr * Kpad + tap * Cin + cw] = ((red[0][i][lane] + red[1][i][lane]) + red[2][i][lane]) +
                                                   red[3][i][lane];
        }
    }
}

// dw[e] += sum over row chunks k (in order) of part[k][e], for the columns
// e % Kpad < kvalid that the chunks wrote
__global__ void __launch_bounds__(256) wgrad_reduce_kernel(const float* __restrict__ part, int nchunks, long n,
                                                           int Kpad, int kvalid, float* __restrict__ dw) {
    for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
        if ((int)(e % Kpad) >= kvalid) continue;
        float s = 0.0f;
        for (int k = 0; k < nchunks; ++k) s += part[(long)k * n + e];
        dw[e] += s;
    }
}

// ---- elementwise / small kernels ---------------------------------------------

__global__ void relu_bwd_kernel(const float* __restrict__ dact, const bf16* __restrict__ act, float* __restrict__ dpre,
                                long n) {
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        dpre[q] = (float)act[q] > 0.0f ? dact[q] : 0.0f;
}

__global__ void bias_grad_kernel(const float* __restrict__ d, long nvox, int C, float* __restrict__ db) {
    // one block per channel chunk of 64 threads x 4 voxel lanes
    __shared__ float sh[256];
    const int c = blockIdx.y;
    float s = 0.0f;
    for (long v = blockIdx.x * 256L + threadIdx.x; v < nvox; v += (long)gridDim.x * 256) s += d[v * C + c];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] += sh[0];  // one block per channel: no other writer
}

// 2x2x2 average pooling (out dims D/2 x H/2 x W/2)
__global__ void avgpool_fwd_kernel(const bf16* __restrict__ in, bf16* __restrict__ out, int B, int D, int H, int W,
                                   int C) {
    const int D2 = D / 2, H2 = H / 2, W2 = W / 2;
    const long n = (long)B * D2 * H2 * W2 * C;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = q % C;
        long v = q / C;
        const int x = v % W2; v /= W2;
        const int y = v % H2; v /= H2;
        const int z = v % D2;
        const int b = v / D2;
        float s = 0.0f;
        for (int k = 0; k < 8; ++k) {
            const int zz = 2 * z + (k >> 2), yy = 2 * y + ((k >> 1) & 1), xx = 2 * x + (k & 1);
            s += (float)in[((((long)b * D + zz) * H + yy) * W + xx) * C + c];
        }
        out[q] = (bf16)(s * 0.125f);
    }
}

// dIn (fine, FP32) += dOut(coarse)/8
__global__ void avgpool_bwd_kernel(const float* __restrict__ dout, float* __restrict__ din, int B, int D, int H, int W,
                                   int C) {
    const long n = (long)B * D * H * W * C;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = q % C;
        long v = q / C;
        const int x = v % W; v /= W;
        const int y = v % H; v /= H;
        const int z = v % D;
        const int b = v / D;
        din[q] += 0.125f * dout[((((long)b * (D / 2) + z / 2) * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
    }
}

// nearest 2x upsampling: out (D x H x W) from in (D/2 x H/2 x W/2)
__global__ void upsample_fwd_kernel(const bf16* __restrict__ in, bf16* __restrict__ out, int B, int D, int H, int W,
                                    int C) {
    const long n = (long)B * D * H * W * C;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = q % C;
        long v = q / C;
        const int x = v % W; v /= W;
        const int y = v % H; v /= H;
        const int z = v % D;
        const int b = v / D;
        out[q] = in[((((long)b * (D / 2) + z / 2) * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
    }
}

// dIn (coarse) = sum of the 8 children of dOut (fine); written, not accumulated
__global__ void upsample_bwd_kernel(const float* __restrict__ dout, float* __restrict__ din, int B, int D, int H, int W,
                                    int C) {
    const int D2 = D / 2, H2 = H / 2, W2 = W / 2;
    const long n = (long)B * D2 * H2 * W2 * C;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = q % C;
        long v = q / C;
        const int x = v % W2; v /= W2;
        const int y = v % H2; v /= H2;
        const int z = v % D2;
        const int b = v / D2;
        float s = 0.0f;
        for (int k = 0; k < 8; ++k) {
            const int zz = 2 * z + (k >> 2), yy = 2 * y + ((k >> 1) & 1), xx = 2 * x + (k & 1);
            s += dout[((((long)b * D + zz) * H + yy) * W + xx) * C + c];
        }
        din[q] = s;
    }
}

// 1x1x1 head: y[v] = sum_c act[v][c] * w[c] + b
__global__ void head_fwd_kernel(const bf16* __restrict__ act, int C, const float* __restrict__ w, float bias,
                                float* __restrict__ y, long nvox) {
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < nvox; v += (long)gridDim.x * blockDim.x) {
        float s = bias;
        for (int c = 0; c < C; ++c) s += (float)act[v * C + c] * w[c];
        y[v] = s;
    }
}

// dAct[v][c] = dy[v]*w[c];  dw[c] += sum_v dy[v]*act[v][c];  db += sum_v dy[v]
// (per-block partial sums of dw and db into bpart[block][0..C]; head_reduce_kernel
// adds the blocks in order)
__global__ void head_bwd_kernel(const float* __restrict__ dy, const bf16* __restrict__ act, int C,
                                const float* __restrict__ w, float* __restrict__ dact, float* __restrict__ bpart,
                                long nvox) {
    __shared__ float sh[17][256];
    float part[17];
    for (int c = 0; c <= C; ++c) part[c] = 0.0f;
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < nvox; v += (long)gridDim.x * blockDim.x) {
        const float d = dy[v];
        for (int c = 0; c < C; ++c) {
            dact[v * C + c] = d * w[c];
            part[c] += d * (float)act[v * C + c];
        }
        part[C] += d;
    }
    for (int c = 0; c <= C; ++c) sh[c][threadIdx.x] = part[c];
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k)
            for (int c = 0; c <= C; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c <= C; ++c) bpart[blockIdx.x * 17 + c] = sh[c][0];
}

__global__ void head_reduce_kernel(const float* __restrict__ part, int nblocks, int C, float* __restrict__ dw,
                                   float* __restrict__ db) {
    const int c = threadIdx.x;
    if (c > C) return;
    float s = 0.0f;
    for (int b = 0; b < nblocks; ++b) s += part[b * 17 + c];
    if (c < C)
        dw[c] += s;
    else
        *db += s;
}

// per-block partial sums of (y - t)^2 in FP64 and dy = 2 (y - t) / n
__global__ void mse_kernel(const float* __restrict__ y, const float* __restrict__ t, float* __restrict__ dy, long n,
                           double* __restrict__ partial) {
    __shared__ double sh[256];
    double s = 0.0;
    const float scale = 2.0f / (float)n;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        const float d = y[q] - t[q];
        dy[q] = scale * d;
        s += (double)d * (double)d;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void sum_partials_kernel(const double* __restrict__ partial, int n, double* __restrict__ out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int q = threadIdx.x; q < n; q += 256) s += partial[q];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ m1,
                            float* __restrict__ m2, long n, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const float gq = gr[q];
        const float a = b1 * m1[q] + (1.0f - b1) * gq;
        const float v = b2 * m2[q] + (1.0f - b2) * gq * gq;
        m1[q] = a;
        m2[q] = v;
        p[q] -= lr * (a * c1) / (sqrtf(v * c2) + eps);
    }
}

// Pack an FP32 master weight [Cout][27][Cin] into the BF16 B-operand layout
// [roundup16(Cout)][Kpad] (k = tap*Cin + ci, zero tail and zero pad rows),
// and the data-gradient weights [roundup16(Cin)][Kpad'] with
// k' = tap'*Cout + co, tap' the mirrored tap.
__global__ void pack_weights_kernel(const float* __restrict__ w, int Cout, int Cin, bf16* __restrict__ wt, int Kpad,
                                    bf16* __restrict__ wd, int Kpadd) {
    const int Co16 = (Cout + 15) / 16 * 16, Ci16 = (Cin + 15) / 16 * 16;
    const long n1 = (long)Co16 * Kpad, n2 = (long)Ci16 * Kpadd;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n1 + n2; q += (long)gridDim.x * blockDim.x) {
        if (q < n1) {
            const int co = q / Kpad, k = q % Kpad;
            wt[q] = (bf16)(co < Cout && k < 27 * Cin ? w[(long)co * 27 * Cin + k] : 0.0f);
        } else {
            const long r = q - n1;
            const int ci = r / Kpadd, k = r % Kpadd;
            float v = 0.0f;
            if (ci < Cin && k < 27 * Cout) {
                const int tapd = k / Cout, co = k % Cout;
                v = w[((long)co * 27 + (26 - tapd)) * Cin + ci];
            }
            wd[r] = (bf16)v;
        }
    }
}

// Network input: channels (gradient, illumination, model, 0 x 5) as BF16
// from three FP32 volumes already normalised.
__global__ void pack_input_kernel(const float* __restrict__ a, const float* __restrict__ b, const float* __restrict__ c,
                                  bf16* __restrict__ x, long nvox) {
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < nvox; v += (long)gridDim.x * blockDim.x) {
        bf16* o = x + v * 8;
        o[0] = (bf16)a[v];
        o[1] = (bf16)b[v];
        o[2] = (bf16)c[v];
        for (int k = 3; k < 8; ++k) o[k] = (bf16)0.0f;
    }
}

// dst[v][c] += src[v][c0 + c] for c < Cd (split of a concatenated gradient)
__global__ void channel_slice_kernel(const float* __restrict__ src, int Cs, int c0, float* __restrict__ dst, int Cd,
                                     long nvox) {
    const long n = nvox * Cd;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const long v = q / Cd;
        const int c = q % Cd;
        dst[q] += src[v * Cs + c0 + c];
    }
}

constexpr int kGrid = 2048;

}  // namespace

hipError_t channel_slice(const float* src, int Cs, int c0, float* dst, int Cd, long nvox, hipStream_t st) {
    channel_slice_kernel<<<kGrid, 256, 0, st>>>(src, Cs, c0, dst, Cd, nvox);
    return hipGetLastError();
}

// ---- launchers -----------------------------------------------------------------

#define FWI_CONV_DISPATCH(TA, OUT)                                                                                 \
    switch (nb) {                                                                                                  \
        case 1: conv_fwd_kernel<TA, 1, OUT><<<grid, 256, 0, st>>>(inA, Ca, inB, Cb, wt, Kpad, bias, Cout, relu, out, \
                                                                   B, D, H, W); break;                             \
        case 2: conv_fwd_kernel<TA, 2, OUT><<<grid, 256, 0, st>>>(inA, Ca, inB, Cb, wt, Kpad, bias, Cout, relu, out, \
                                                                   B, D, H, W); break;                             \
        case 3: conv_fwd_kernel<TA, 3, OUT><<<grid, 256, 0, st>>>(inA, Ca, inB, Cb, wt, Kpad, bias, Cout, relu, out, \
                                                                   B, D, H, W); break;                             \
        case 4: conv_fwd_kernel<TA, 4, OUT><<<grid, 256, 0, st>>>(inA, Ca, inB, Cb, wt, Kpad, bias, Cout, relu, out, \
                                                                   B, D, H, W); break;                             \
        case 6: conv_fwd_kernel<TA, 6, OUT><<<grid, 256, 0, st>>>(inA, Ca, inB, Cb, wt, Kpad, bias, Cout, relu, out, \
                                                                   B, D, H, W); break;                             \
        default: return hipErrorInvalidValue;                                                                      \
    }

hipError_t conv_fwd(const bf16* inA, int Ca, const bf16* inB, int Cb, const bf16* wt, int Kpad, const float* bias,
                    int Cout, bool relu, bf16* out, int B, int D, int H, int W, hipStream_t st) {
    const int nb = (Cout + 15) / 16;
    dim3 grid((W + 15) / 16, (H + 3) / 4, D * B);
    FWI_CONV_DISPATCH(bf16, 0)
    return hipGetLastError();
}

hipError_t conv_fwd_f32out(const bf16* inA, int Ca, const bf16* inB, int Cb, const bf16* wt, int Kpad, int Cout,
                           float* out, int B, int D, int H, int W, hipStream_t st) {
    const int nb = (Cout + 15) / 16;
    const float* bias = nullptr;
    const int relu = 0;
    dim3 grid((W + 15) / 16, (H + 3) / 4, D * B);
    FWI_CONV_DISPATCH(bf16, 1)
    return hipGetLastError();
}

hipError_t conv_dgrad(const float* dout, int Cout, const bf16* wd, int Kpadd, int Cin, float* din, int B, int D,
                      int H, int W, hipStream_t st) {
    const int nb = (Cin + 15) / 16;
    const float* inA = dout;
    const int Ca = Cout, Cb = 0;
    const bf16* inB = nullptr;
    const bf16* wt = wd;
    const int Kpad = Kpadd;
    const float* bias = nullptr;
    const int relu = 0;
    void* out = din;
    const int CoutD = Cin;
    dim3 grid((W + 15) / 16, (H + 3) / 4, D * B);
    {
        const int Cout = CoutD;
        FWI_CONV_DISPATCH(float, 2)
    }
    return hipGetLastError();
}

hipError_t conv_wgrad(const float* dout, int Cout, const bf16* inA, int Ca, const bf16* inB, int Cb, float* dw,
                      int ldw, int B, int D, int H, int W, float* part, long part_cap, hipStream_t st) {
    if (Cout % 16 != 0) return hipErrorInvalidValue;
    const int Cin = Ca + Cb;
    if (ldw < 27 * Cin) return hipErrorInvalidValue;
    const int rows = B * D * H;
    // row chunks of at least 16 rows, at most kWgradChunks of them
    const int nchunks = std::min(kWgradChunks, (rows + 15) / 16);
    const int rpb = (rows + nchunks - 1) / nchunks;
    const int nck = (rows + rpb - 1) / rpb;
    const long n = (long)Cout * ldw;
    if ((long)nck * n > part_cap) return hipErrorInvalidValue;
    dim3 grid(27 * ((Cin + 15) / 16) * (Cout / 16), nck);
    conv_wgrad_kernel<bf16><<<grid, 256, 0, st>>>(dout, Cout, inA, Ca, inB, Cb, part, ldw, B, D, H, W, rpb);
    wgrad_reduce_kernel<<<(unsigned)std::min<long>((n + 255) / 256, 1024), 256, 0, st>>>(part, nck, n, ldw, 27 * Cin,
                                                                                         dw);
    return hipGetLastError();
}

hipError_t relu_bwd(const float* dact, const bf16* act, float* dpre, long n, hipStream_t st) {
    relu_bwd_kernel<<<kGrid, 256, 0, st>>>(dact, act, dpre, n);
    return hipGetLastError();
}

hipError_t bias_grad(const float* d, long nvox, int C, float* db, hipStream_t st) {
    bias_grad_kernel<<<dim3(1, C), 256, 0, st>>>(d, nvox, C, db);
    return hipGetLastError();
}

hipError_t avgpool_fwd(const bf16* in, bf16* out, int B, int D, int H, int W, int C, hipStream_t st) {
    avgpool_fwd_kernel<<<kGrid, 256, 0, st>>>(in, out, B, D, H, W, C);
    return hipGetLastError();
}

hipError_t avgpool_bwd(const float* dout, float* din, int B, int D, int H, int W, int C, hipStream_t st) {
    avgpool_bwd_kernel<<<kGrid, 256, 0, st>>>(dout, din, B, D, H, W, C);
    return hipGetLastError();
}

hipError_t upsample_fwd(const bf16* in, bf16* out, int B, int D, int H, int W, int C, hipStream_t st) {
    upsample_fwd_kernel<<<kGrid, 256, 0, st>>>(in, out, B, D, H, W, C);
    return hipGetLastError();
}

hipError_t upsample_bwd(const float* dout, float* din, int B, int D, int H, int W, int C, hipStream_t st) {
    upsample_bwd_kernel<<<kGrid, 256, 0, st>>>(dout, din, B, D, H, W, C);
    return hipGetLastError();
}

hipError_t head_fwd(const bf16* act, int C, const float* w, float bias, float* y, long nvox, hipStream_t st) {
    head_fwd_kernel<<<kGrid, 256, 0, st>>>(act, C, w, bias, y, nvox);
    return hipGetLastError();
}

hipError_t head_bwd(const float* dy, const bf16* act, int C, const float* w, float* dact, float* dw, float* db,
                    long nvox, float* part, long part_cap, hipStream_t st) {
    if (C > 16 || part_cap < 256 * 17) return hipErrorInvalidValue;
    head_bwd_kernel<<<256, 256, 0, st>>>(dy, act, C, w, dact, part, nvox);
    head_reduce_kernel<<<1, 64, 0, st>>>(part, 256, C, dw, db);
    return hipGetLastError();
}

hipError_t mse(const float* y, const float* t, float* dy, long n, double* partial, double* out, hipStream_t st) {
    mse_kernel<<<1024, 256, 0, st>>>(y, t, dy, n, partial);
    sum_partials_kernel<<<1, 256, 0, st>>>(partial, 1024, out);
    return hipGetLastError();
}

hipError_t adam(float* p, const float* g, float* m1, float* m2, long n, float lr, float b1, float b2, float eps,
                int step, hipStream_t st) {
    const float c1 = 1.0f / (1.0f - powf(b1, (float)step));
    const float c2 = 1.0f / (1.0f - powf(b2, (float)step));
    adam_kernel<<<kGrid, 256, 0, st>>>(p, g, m1, m2, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t pack_weights(const float* w, int Cout, int Cin, bf16* wt, int Kpad, bf16* wd, int Kpadd, hipStream_t st) {
    pack_weights_kernel<<<kGrid, 256, 0, st>>>(w, Cout, Cin, wt, Kpad, wd, Kpadd);
    return hipGetLastError();
}

hipError_t pack_input(const float* a, const float* b, const float* c, bf16* x, long nvox, hipStream_t st) {
    pack_input_kernel<<<kGrid, 256, 0, st>>>(a, b, c, x, nvox);
    return hipGetLastError();
}

}  // namespace unet
}  // namespace fwi

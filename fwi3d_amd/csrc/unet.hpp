// unet.hpp -- launchers of the U-Net kernels (unet_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace fwi {
namespace unet {

typedef __bf16 bf16;

// 3x3x3 conv: out = relu?(W * concat(inA[Ca], inB[Cb]) + bias), BF16 out.
// wt: BF16 [roundup16(Cout)][Kpad], k = tap*Cin + ci, Kpad = roundup32(27*Cin).
hipError_t conv_fwd(const bf16* inA, int Ca, const bf16* inB, int Cb, const bf16* wt, int Kpad, const float* bias,
                    int Cout, bool relu, bf16* out, int B, int D, int H, int W, hipStream_t st);
// same without bias / activation, FP32 output (tests)
hipError_t conv_fwd_f32out(const bf16* inA, int Ca, const bf16* inB, int Cb, const bf16* wt, int Kpad, int Cout,
                           float* out, int B, int D, int H, int W, hipStream_t st);
// din[Cin] += W^T * dout[Cout]; wd: BF16 [roundup16(Cin)][Kpadd], k = tap'*Cout + co (mirrored taps)
hipError_t conv_dgrad(const float* dout, int Cout, const bf16* wd, int Kpadd, int Cin, float* din, int B, int D,
                      int H, int W, hipStream_t st);
// dw[co][tap*Cin + ci] += sum_v dout[v][co] * concat(inA,inB)[v+tap][ci];  dw row stride = ldw.
// Deterministic: partial sums of at most kWgradChunks row chunks go to part
// (part_cap floats, at least chunks * Cout * ldw) and are added in order.
constexpr int kWgradChunks = 64;
hipError_t conv_wgrad(const float* dout, int Cout, const bf16* inA, int Ca, const bf16* inB, int Cb, float* dw,
                      int ldw, int B, int D, int H, int W, float* part, long part_cap, hipStream_t st);
hipError_t relu_bwd(const float* dact, const bf16* act, float* dpre, long n, hipStream_t st);
hipError_t bias_grad(const float* d, long nvox, int C, float* db, hipStream_t st);
hipError_t avgpool_fwd(const bf16* in, bf16* out, int B, int D, int H, int W, int C, hipStream_t st);
hipError_t avgpool_bwd(const float* dout, float* din, int B, int D, int H, int W, int C, hipStream_t st);
hipError_t upsample_fwd(const bf16* in, bf16* out, int B, int D, int H, int W, int C, hipStream_t st);
hipError_t upsample_bwd(const float* dout, float* din, int B, int D, int H, int W, int C, hipStream_t st);
hipError_t head_fwd(const bf16* act, int C, const float* w, float bias, float* y, long nvox, hipStream_t st);
// deterministic like conv_wgrad: part holds at least 256 * 17 floats
hipError_t head_bwd(const float* dy, const bf16* act, int C, const float* w, float* dact, float* dw, float* db,
                    long nvox, float* part, long part_cap, hipStream_t st);
hipError_t mse(const float* y, const float* t, float* dy, long n, double* partial, double* out, hipStream_t st);
hipError_t adam(float* p, const float* g, float* m1, float* m2, long n, float lr, float b1, float b2, float eps,
                int step, hipStream_t st);
hipError_t pack_weights(const float* w, int Cout, int Cin, bf16* wt, int Kpad, bf16* wd, int Kpadd, hipStream_t st);
hipError_t pack_input(const float* a, const float* b, const float* c, bf16* x, long nvox, hipStream_t st);
hipError_t channel_slice(const float* src, int Cs, int c0, float* dst, int Cd, long nvox, hipStream_t st);

}  // namespace unet
}  // namespace fwi

// unet_net.hip -- the learned FWI preconditioner: a 3-level 3-D U-Net
// (16 / 32 / 64 channels, 3x3x3 convolutions, ReLU, 2x2x2 average pooling,
// nearest upsampling, skip concatenation, 1x1x1 head), trained with Adam on
// the GPU by the kernels of unet_kernels.hip.  BF16 weights and activations,
// FP32 accumulation, FP32 master weights and gradients.
//
//   x[3->8] -c1a-> e1a[16] -c1b-> e1b[16] -pool-> p1[16] -c2a-> e2a[32] -c2b-> e2b[32]
//   -pool-> p2[32] -c3a-> e3a[64] -c3b-> e3b[64] -up-> u2[64] ++ e2b -c4a-> d2a[32]
//   -c4b-> d2b[32] -up-> u1[32] ++ e1b -c5a-> d1a[16] -c5b-> d1b[16] -head-> y[1]
//
// Parameters are one flat FP32 vector: for each conv layer W[Cout][27][Cin]
// then b[Cout]; then the head w[16], b.  Taps are ordered
// tap = (dz+1)*9 + (dy+1)*3 + (dx+1).
#include <cmath>
#include <algorithm>
#include <cstdint>
#include <random>
#include <string>
#include <vector>

#include "../../include/fwi3d.h"
#include "unet.hpp"

using fwi::unet::bf16;

namespace {

struct ConvSpec {
    int ca, cb, cout;  // channels of the two concatenated inputs, output channels
};

// c1a c1b c2a c2b c3a c3b c4a c4b c5a c5b
const ConvSpec kConvs[10] = {{8, 0, 16},  {16, 0, 16}, {16, 0, 32}, {32, 0, 32}, {32, 0, 64},
                             {64, 0, 64}, {64, 32, 32}, {32, 0, 32}, {32, 16, 16}, {16, 0, 16}};
constexpr int kHeadC = 16;
constexpr int kInC = 3;  // gradient, illumination, model

int round_up(int v, int m) { return (v + m - 1) / m * m; }

}  // namespace

struct fwi_unet {
    int device = 0;
    hipStream_t st = nullptr;
    long nparams = 0;
    long woff[10] = {}, boff[10] = {}, hoff = 0;
    int kpad[10] = {}, kpadd[10] = {};
    float* p = nullptr;   // FP32 master parameters
    float* g = nullptr;   // FP32 gradients
    float* m1 = nullptr;  // Adam moments
    float* m2 = nullptr;
    bf16* wt[10] = {};    // packed forward weights
    bf16* wd[10] = {};    // packed data-gradient weights
    int adam_step = 0;
    double* partial = nullptr;
    double* result = nullptr;
    // per-chunk partial sums of the weight gradients (deterministic backward)
    float* wpart = nullptr;
    long wpart_cap = 0;
    // activations (BF16) and gradients (FP32) for the current shape
    int B = 0, D = 0, H = 0, W = 0;
    bool grads = false;
    bf16 *x = nullptr, *e1a = nullptr, *e1b = nullptr, *p1 = nullptr, *e2a = nullptr, *e2b = nullptr, *p2 = nullptr,
         *e3a = nullptr, *e3b = nullptr, *u2 = nullptr, *d2a = nullptr, *d2b = nullptr, *u1 = nullptr,
         *d1a = nullptr, *d1b = nullptr;
    float *in3 = nullptr, *y = nullptr, *tgt = nullptr, *dy = nullptr;
    float *gd1b = nullptr, *gd1a = nullptr, *gcat1 = nullptr, *gu1 = nullptr, *ge1b = nullptr, *gd2b = nullptr,
          *gd2a = nullptr, *gcat2 = nullptr, *gu2 = nullptr, *ge2b = nullptr, *ge3b = nullptr, *ge3a = nullptr,
          *gp2 = nullptr, *ge2a = nullptr, *gp1 = nullptr, *ge1a = nullptr;
    std::string err;
};

namespace {

int ufail(fwi_unet* n, int code, const std::string& m) {
    if (n) n->err = m;
    return code;
}

#define UCHK(n, expr)                                                                                          \
    do {                                                                                                       \
        hipError_t e_ = (expr);                                                                                \
        if (e_ != hipSuccess) return ufail(n, FWI_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define URC(expr)              \
    do {                       \
        int rc_ = (expr);      \
        if (rc_) return rc_;   \
    } while (0)

template <typename T>
int ualloc(fwi_unet* n, T** p, long count) {
    UCHK(n, hipMalloc((void**)p, count * sizeof(T)));
    UCHK(n, hipMemsetAsync(*p, 0, count * sizeof(T), n->st));
    return FWI_OK;
}

template <typename T>
void ufree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

void free_shape(fwi_unet* n) {
    for (bf16** q : {&n->x, &n->e1a, &n->e1b, &n->p1, &n->e2a, &n->e2b, &n->p2, &n->e3a, &n->e3b, &n->u2, &n->d2a,
                     &n->d2b, &n->u1, &n->d1a, &n->d1b})
        ufree(*q);
    for (float** q : {&n->in3, &n->y, &n->tgt, &n->dy, &n->gd1b, &n->gd1a, &n->gcat1, &n->gu1, &n->ge1b, &n->gd2b,
                      &n->gd2a, &n->gcat2, &n->gu2, &n->ge2b, &n->ge3b, &n->ge3a, &n->gp2, &n->ge2a, &n->gp1,
                      &n->ge1a})
        ufree(*q);
    n->B = n->D = n->H = n->W = 0;
    n->grads = false;
}

int ensure_shape(fwi_unet* n, int B, int D, int H, int W, bool grads) {
    if (B < 1 || D < 4 || H < 4 || W < 4 || D % 4 || H % 4 || W % 4)
        return ufail(n, FWI_ERR_ARG, "U-Net input dims must be >= 4 and divisible by 4");
    const bool same = n->B == B && n->D == D && n->H == H && n->W == W;
    if (same && (!grads || n->grads)) return FWI_OK;
    if (!same) free_shape(n);
    const long v1 = (long)B * D * H * W, v2 = v1 / 8, v3 = v1 / 64;
    if (!n->x) {
        URC(ualloc(n, &n->x, v1 * 8));
        URC(ualloc(n, &n->e1a, v1 * 16));
        URC(ualloc(n, &n->e1b, v1 * 16));
        URC(ualloc(n, &n->p1, v2 * 16));
        URC(ualloc(n, &n->e2a, v2 * 32));
        URC(ualloc(n, &n->e2b, v2 * 32));
        URC(ualloc(n, &n->p2, v3 * 32));
        URC(ualloc(n, &n->e3a, v3 * 64));
        URC(ualloc(n, &n->e3b, v3 * 64));
        URC(ualloc(n, &n->u2, v2 * 64));
        URC(ualloc(n, &n->d2a, v2 * 32));
        URC(ualloc(n, &n->d2b, v2 * 32));
        URC(ualloc(n, &n->u1, v1 * 32));
        URC(ualloc(n, &n->d1a, v1 * 16));
        URC(ualloc(n, &n->d1b, v1 * 16));
        URC(ualloc(n, &n->in3, v1 * kInC));
        URC(ualloc(n, &n->y, v1));
    }
    if (grads && !n->grads) {
        URC(ualloc(n, &n->tgt, v1));
        URC(ualloc(n, &n->dy, v1));
        URC(ualloc(n, &n->gd1b, v1 * 16));
        URC(ualloc(n, &n->gd1a, v1 * 16));
        URC(ualloc(n, &n->gcat1, v1 * 48));
        URC(ualloc(n, &n->gu1, v1 * 32));
        URC(ualloc(n, &n->ge1b, v1 * 16));
        URC(ualloc(n, &n->gd2b, v2 * 32));
        URC(ualloc(n, &n->gd2a, v2 * 32));
        URC(ualloc(n, &n->gcat2, v2 * 96));
        URC(ualloc(n, &n->gu2, v2 * 64));
        URC(ualloc(n, &n->ge2b, v2 * 32));
        URC(ualloc(n, &n->ge3b, v3 * 64));
        URC(ualloc(n, &n->ge3a, v3 * 64));
        URC(ualloc(n, &n->gp2, v3 * 32));
        URC(ualloc(n, &n->ge2a, v2 * 32));
        URC(ualloc(n, &n->gp1, v2 * 16));
        URC(ualloc(n, &n->ge1a, v1 * 16));
        n->grads = true;
    }
    n->B = B;
    n->D = D;
    n->H = H;
    n->W = W;
    return FWI_OK;
}

int pack_all(fwi_unet* n) {
    for (int l = 0; l < 10; ++l) {
        const ConvSpec& s = kConvs[l];
        UCHK(n, fwi::unet::pack_weights(n->p + n->woff[l], s.cout, s.ca + s.cb, n->wt[l], n->kpad[l], n->wd[l],
                                        n->kpadd[l], n->st));
    }
    return FWI_OK;
}

// forward pass on n->in3 ([3][B*D*H*W] planar FP32) -> n->y
int forward(fwi_unet* n) {
    using namespace fwi::unet;
    const int B = n->B, D = n->D, H = n->H, W = n->W;
    const int D2 = D / 2, H2 = H / 2, W2 = W / 2, D3 = D / 4, H3 = H / 4, W3 = W / 4;
    const long v1 = (long)B * D * H * W;
    hipStream_t st = n->st;
    UCHK(n, pack_input(n->in3, n->in3 + v1, n->in3 + 2 * v1, n->x, v1, st));
    auto conv = [&](int l, const bf16* a, const bf16* b, bf16* out, int d, int h, int w) {
        const ConvSpec& s = kConvs[l];
        return conv_fwd(a, s.ca, b, s.cb, n->wt[l], n->kpad[l], n->p + n->boff[l], s.cout, true, out, B, d, h, w, st);
    };
    UCHK(n, conv(0, n->x, nullptr, n->e1a, D, H, W));
    UCHK(n, conv(1, n->e1a, nullptr, n->e1b, D, H, W));
    UCHK(n, avgpool_fwd(n->e1b, n->p1, B, D, H, W, 16, st));
    UCHK(n, conv(2, n->p1, nullptr, n->e2a, D2, H2, W2));
    UCHK(n, conv(3, n->e2a, nullptr, n->e2b, D2, H2, W2));
    UCHK(n, avgpool_fwd(n->e2b, n->p2, B, D2, H2, W2, 32, st));
    UCHK(n, conv(4, n->p2, nullptr, n->e3a, D3, H3, W3));
    UCHK(n, conv(5, n->e3a, nullptr, n->e3b, D3, H3, W3));
    UCHK(n, upsample_fwd(n->e3b, n->u2, B, D2, H2, W2, 64, st));
    UCHK(n, conv(6, n->u2, n->e2b, n->d2a, D2, H2, W2));
    UCHK(n, conv(7, n->d2a, nullptr, n->d2b, D2, H2, W2));
    UCHK(n, upsample_fwd(n->d2b, n->u1, B, D, H, W, 32, st));
    UCHK(n, conv(8, n->u1, n->e1b, n->d1a, D, H, W));
    UCHK(n, conv(9, n->d1a, nullptr, n->d1b, D, H, W));
    float hb = 0.0f;
    UCHK(n, hipMemcpyAsync(&hb, n->p + n->hoff + kHeadC, sizeof(float), hipMemcpyDeviceToHost, st));
    UCHK(n, hipStreamSynchronize(st));
    UCHK(n, head_fwd(n->d1b, kHeadC, n->p + n->hoff, hb, n->y, v1, st));
    return FWI_OK;
}

// MSE loss against n->tgt and all parameter gradients into n->g
int backward(fwi_unet* n, double* loss) {
    using namespace fwi::unet;
    const int B = n->B, D = n->D, H = n->H, W = n->W;
    const int D2 = D / 2, H2 = H / 2, W2 = W / 2, D3 = D / 4, H3 = H / 4, W3 = W / 4;
    const long v1 = (long)B * D * H * W, v2 = v1 / 8, v3 = v1 / 64;
    hipStream_t st = n->st;
    UCHK(n, hipMemsetAsync(n->g, 0, n->nparams * sizeof(float), st));
    UCHK(n, mse(n->y, n->tgt, n->dy, v1, n->partial, n->result, st));
    // conv layer l: dact (in place -> dpre), bias and weight gradients, and
    // the data gradient into din (zeroed first) when din != nullptr
    auto layer = [&](int l, float* dact, const bf16* act, long nvox, const bf16* ina, const bf16* inb, float* din,
                     int d, int h, int w) -> int {
        const ConvSpec& s = kConvs[l];
        const int cin = s.ca + s.cb;
        UCHK(n, relu_bwd(dact, act, dact, nvox * s.cout, st));
        UCHK(n, bias_grad(dact, nvox, s.cout, n->g + n->boff[l], st));
        UCHK(n, conv_wgrad(dact, s.cout, ina, s.ca, inb, s.cb, n->g + n->woff[l], 27 * cin, B, d, h, w, n->wpart,
                           n->wpart_cap, st));
        if (!din) return FWI_OK;
        UCHK(n, hipMemsetAsync(din, 0, nvox * cin * sizeof(float), st));
        UCHK(n, conv_dgrad(dact, s.cout, n->wd[l], n->kpadd[l], cin, din, B, d, h, w, st));
        return FWI_OK;
    };
    auto slice = [&](const float* src, int cs, int c0, float* dst, int cd, long nvox) -> int {
        UCHK(n, hipMemsetAsync(dst, 0, nvox * cd * sizeof(float), st));
        UCHK(n, channel_slice(src, cs, c0, dst, cd, nvox, st));
        return FWI_OK;
    };
    UCHK(n, head_bwd(n->dy, n->d1b, kHeadC, n->p + n->hoff, n->gd1b, n->g + n->hoff, n->g + n->hoff + kHeadC, v1,
                     n->wpart, n->wpart_cap, st));
    URC(layer(9, n->gd1b, n->d1b, v1, n->d1a, nullptr, n->gd1a, D, H, W));
    URC(layer(8, n->gd1a, n->d1a, v1, n->u1, n->e1b, n->gcat1, D, H, W));
    URC(slice(n->gcat1, 48, 0, n->gu1, 32, v1));
    URC(slice(n->gcat1, 48, 32, n->ge1b, 16, v1));
    UCHK(n, upsample_bwd(n->gu1, n->gd2b, B, D, H, W, 32, st));
    URC(layer(7, n->gd2b, n->d2b, v2, n->d2a, nullptr, n->gd2a, D2, H2, W2));
    URC(layer(6, n->gd2a, n->d2a, v2, n->u2, n->e2b, n->gcat2, D2, H2, W2));
    URC(slice(n->gcat2, 96, 0, n->gu2, 64, v2));
    URC(slice(n->gcat2, 96, 64, n->ge2b, 32, v2));
    UCHK(n, upsample_bwd(n->gu2, n->ge3b, B, D2, H2, W2, 64, st));
    URC(layer(5, n->ge3b, n->e3b, v3, n->e3a, nullptr, n->ge3a, D3, H3, W3));
    URC(layer(4, n->ge3a, n->e3a, v3, n->p2, nullptr, n->gp2, D3, H3, W3));
    UCHK(n, avgpool_bwd(n->gp2, n->ge2b, B, D2, H2, W2, 32, st));
    URC(layer(3, n->ge2b, n->e2b, v2, n->e2a, nullptr, n->ge2a, D2, H2, W2));
    URC(layer(2, n->ge2a, n->e2a, v2, n->p1, nullptr, n->gp1, D2, H2, W2));
    UCHK(n, avgpool_bwd(n->gp1, n->ge1b, B, D, H, W, 16, st));
    URC(layer(1, n->ge1b, n->e1b, v1, n->e1a, nullptr, n->ge1a, D, H, W));
    URC(layer(0, n->ge1a, n->e1a, v1, n->x, nullptr, nullptr, D, H, W));
    double s = 0.0;
    UCHK(n, hipMemcpyAsync(&s, n->result, sizeof s, hipMemcpyDeviceToHost, st));
    UCHK(n, hipStreamSynchronize(st));
    if (loss) *loss = s / (double)v1;
    return FWI_OK;
}

int upload_inputs(fwi_unet* n, const float* x3, const float* target) {
    const long v1 = (long)n->B * n->D * n->H * n->W;
    UCHK(n, hipMemcpyAsync(n->in3, x3, v1 * kInC * sizeof(float), hipMemcpyHostToDevice, n->st));
    if (target) UCHK(n, hipMemcpyAsync(n->tgt, target, v1 * sizeof(float), hipMemcpyHostToDevice, n->st));
    return FWI_OK;
}

}  // namespace

extern "C" {

int fwi_unet_create(int device, uint64_t seed, fwi_unet** out) {
    if (!out) return FWI_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return FWI_ERR_NODEV;
    if (device < 0 || device >= ndev || hipSetDevice(device) != hipSuccess) return FWI_ERR_ARG;
    fwi_unet* n = new fwi_unet();
    n->device = device;
    auto bail = [&](int rc) {
        fwi_unet_destroy(n);
        return rc;
    };
    if (hipStreamCreateWithFlags(&n->st, hipStreamNonBlocking) != hipSuccess) return bail(FWI_ERR_HIP);
    long off = 0;
    for (int l = 0; l < 10; ++l) {
        const ConvSpec& s = kConvs[l];
        const int cin = s.ca + s.cb;
        n->woff[l] = off;
        off += (long)s.cout * 27 * cin;
        n->boff[l] = off;
        off += s.cout;
        n->kpad[l] = round_up(27 * cin, 32);
        n->kpadd[l] = round_up(27 * s.cout, 32);
    }
    n->hoff = off;
    off += kHeadC + 1;
    n->nparams = off;
    // He-normal initialisation of the convolutions (only the kInC real input
    // channels of the first layer are non-zero), small head
    std::vector<float> h(off, 0.0f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < 10; ++l) {
        const ConvSpec& s = kConvs[l];
        const int cin = s.ca + s.cb;
        const int fan_in = 27 * (l == 0 ? kInC : cin);
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / fan_in));
        for (int co = 0; co < s.cout; ++co)
            for (int t = 0; t < 27; ++t)
                for (int ci = 0; ci < cin; ++ci)
                    h[n->woff[l] + ((long)co * 27 + t) * cin + ci] = (l == 0 && ci >= kInC) ? 0.0f : nd(rng);
    }
    std::normal_distribution<float> hd(0.0f, 0.05f);
    for (int c = 0; c < kHeadC; ++c) h[n->hoff + c] = hd(rng);
    long wmax = 256 * 17;
    for (int l = 0; l < 10; ++l)
        wmax = std::max(wmax, (long)fwi::unet::kWgradChunks * kConvs[l].cout * 27 * (kConvs[l].ca + kConvs[l].cb));
    n->wpart_cap = wmax;
    if (ualloc(n, &n->p, off) || ualloc(n, &n->g, off) || ualloc(n, &n->m1, off) || ualloc(n, &n->m2, off) ||
        ualloc(n, &n->partial, 1024) || ualloc(n, &n->result, 4) || ualloc(n, &n->wpart, wmax))
        return bail(FWI_ERR_HIP);
    for (int l = 0; l < 10; ++l) {
        const ConvSpec& s = kConvs[l];
        if (ualloc(n, &n->wt[l], (long)round_up(s.cout, 16) * n->kpad[l]) ||
            ualloc(n, &n->wd[l], (long)round_up(s.ca + s.cb, 16) * n->kpadd[l]))
            return bail(FWI_ERR_HIP);
    }
    if (hipMemcpy(n->p, h.data(), off * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return bail(FWI_ERR_HIP);
    if (pack_all(n)) return bail(FWI_ERR_HIP);
    if (hipStreamSynchronize(n->st) != hipSuccess) return bail(FWI_ERR_HIP);
    *out = n;
    return FWI_OK;
}

int fwi_unet_destroy(fwi_unet* n) {
    if (!n) return FWI_OK;
    (void)hipSetDevice(n->device);
    if (n->st) (void)hipStreamSynchronize(n->st);
    free_shape(n);
    ufree(n->p);
    ufree(n->g);
    ufree(n->m1);
    ufree(n->m2);
    ufree(n->partial);
    ufree(n->result);
    ufree(n->wpart);
    for (int l = 0; l < 10; ++l) {
        ufree(n->wt[l]);
        ufree(n->wd[l]);
    }
    if (n->st) (void)hipStreamDestroy(n->st);
    delete n;
    return FWI_OK;
}

const char* fwi_unet_last_error(const fwi_unet* n) { return n ? n->err.c_str() : ""; }

long fwi_unet_num_params(const fwi_unet* n) { return n ? n->nparams : 0; }

int fwi_unet_get_params(fwi_unet* n, float* out) {
    if (!n || !out) return FWI_ERR_ARG;
    UCHK(n, hipSetDevice(n->device));
    UCHK(n, hipMemcpyAsync(out, n->p, n->nparams * sizeof(float), hipMemcpyDeviceToHost, n->st));
    UCHK(n, hipStreamSynchronize(n->st));
    return FWI_OK;
}

int fwi_unet_set_params(fwi_unet* n, const float* in) {
    if (!n || !in) return FWI_ERR_ARG;
    UCHK(n, hipSetDevice(n->device));
    UCHK(n, hipMemcpyAsync(n->p, in, n->nparams * sizeof(float), hipMemcpyHostToDevice, n->st));
    URC(pack_all(n));
    UCHK(n, hipStreamSynchronize(n->st));
    return FWI_OK;
}

int fwi_unet_forward(fwi_unet* n, const float* x3, int B, int D, int H, int W, float* y) {
    if (!n || !x3 || !y) return FWI_ERR_ARG;
    UCHK(n, hipSetDevice(n->device));
    URC(ensure_shape(n, B, D, H, W, false));
    URC(upload_inputs(n, x3, nullptr));
    URC(forward(n));
    UCHK(n, hipMemcpyAsync(y, n->y, (long)B * D * H * W * sizeof(float), hipMemcpyDeviceToHost, n->st));
    UCHK(n, hipStreamSynchronize(n->st));
    return FWI_OK;
}

int fwi_unet_gradients(fwi_unet* n, const float* x3, const float* target, int B, int D, int H, int W, double* loss,
                       float* grad) {
    if (!n || !x3 || !target) return FWI_ERR_ARG;
    UCHK(n, hipSetDevice(n->device));
    URC(ensure_shape(n, B, D, H, W, true));
    URC(upload_inputs(n, x3, target));
    URC(forward(n));
    URC(backward(n, loss));
    if (grad) {
        UCHK(n, hipMemcpyAsync(grad, n->g, n->nparams * sizeof(float), hipMemcpyDeviceToHost, n->st));
        UCHK(n, hipStreamSynchronize(n->st));
    }
    return FWI_OK;
}

int fwi_unet_train_step(fwi_unet* n, const float* x3, const float* target, int B, int D, int H, int W, float lr,
                        double* loss) {
    if (!n || !x3 || !target || !(lr > 0.0f)) return FWI_ERR_ARG;
    URC(fwi_unet_gradients(n, x3, target, B, D, H, W, loss, nullptr));
    n->adam_step++;
    UCHK(n, fwi::unet::adam(n->p, n->g, n->m1, n->m2, n->nparams, lr, 0.9f, 0.999f, 1e-8f, n->adam_step, n->st));
    URC(pack_all(n));
    UCHK(n, hipStreamSynchronize(n->st));
    return FWI_OK;
}

// Convolution layer probe for tests: out (FP32, no bias / activation) =
// conv(layer l) of a host channels-last BF16-representable input.
int fwi_unet_conv_probe(fwi_unet* n, int layer, const float* in, int B, int D, int H, int W, float* out) {
    if (!n || !in || !out || layer < 0 || layer > 9) return FWI_ERR_ARG;
    UCHK(n, hipSetDevice(n->device));
    const ConvSpec& s = kConvs[layer];
    const int cin = s.ca + s.cb;
    const long v = (long)B * D * H * W;
    std::vector<bf16> hb(v * cin);
    for (long q = 0; q < v * cin; ++q) hb[q] = (bf16)in[q];
    bf16* din = nullptr;
    float* dout = nullptr;
    URC(ualloc(n, &din, v * cin));
    URC(ualloc(n, &dout, v * s.cout));
    int rc = FWI_OK;
    hipError_t e = hipMemcpyAsync(din, hb.data(), v * cin * sizeof(bf16), hipMemcpyHostToDevice, n->st);
    if (e == hipSuccess)
        e = fwi::unet::conv_fwd_f32out(din, cin, nullptr, 0, n->wt[layer], n->kpad[layer], s.cout, dout, B, D, H, W,
                                       n->st);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, v * s.cout * sizeof(float), hipMemcpyDeviceToHost, n->st);
    if (e == hipSuccess) e = hipStreamSynchronize(n->st);
    if (e != hipSuccess) rc = ufail(n, FWI_ERR_HIP, hipGetErrorString(e));
    (void)hipFree(din);
    (void)hipFree(dout);
    return rc;
}

// Gradient probe of one convolution layer for tests: the data gradient
// dx = W^T dout ([v][Cin], FP32) and the weight gradient
// dw = sum_v dout x ([Cout][27][Cin], FP32) of a channels-last input `in`
// ([v][Cin], rounded to BF16 on upload; with two inputs the first Ca channels
// form input A and the others input B, as the network's skip concatenation)
// and an FP32 output gradient dout ([v][Cout]), with the backward kernels of
// the training step.
int fwi_unet_conv_grad_probe(fwi_unet* n, int layer, const float* in, const float* dout, int B, int D, int H, int W,
                             float* dx, float* dw) {
    if (!n || !in || !dout || !dx || !dw || layer < 0 || layer > 9 || B < 1 || D < 1 || H < 1 || W < 1)
        return FWI_ERR_ARG;
    UCHK(n, hipSetDevice(n->device));
    const ConvSpec& s = kConvs[layer];
    const int cin = s.ca + s.cb;
    const long v = (long)B * D * H * W;
    const long nw = (long)s.cout * 27 * cin;
    std::vector<bf16> ha(v * s.ca), hb(v * s.cb);
    for (long q = 0; q < v; ++q) {
        for (int c = 0; c < s.ca; ++c) ha[q * s.ca + c] = (bf16)in[q * cin + c];
        for (int c = 0; c < s.cb; ++c) hb[q * s.cb + c] = (bf16)in[q * cin + s.ca + c];
    }
    bf16 *da = nullptr, *db = nullptr;
    float *dgo = nullptr, *ddx = nullptr, *ddw = nullptr, *part = nullptr;
    const long part_cap = (long)fwi::unet::kWgradChunks * nw;
    int rc = ualloc(n, &da, v * s.ca);
    if (!rc && s.cb) rc = ualloc(n, &db, v * s.cb);
    if (!rc) rc = ualloc(n, &dgo, v * s.cout);
    if (!rc) rc = ualloc(n, &ddx, v * cin);
    if (!rc) rc = ualloc(n, &ddw, nw);
    if (!rc) rc = ualloc(n, &part, part_cap);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(da, ha.data(), v * s.ca * sizeof(bf16), hipMemcpyHostToDevice, n->st);
        if (e == hipSuccess && s.cb)
            e = hipMemcpyAsync(db, hb.data(), v * s.cb * sizeof(bf16), hipMemcpyHostToDevice, n->st);
        if (e == hipSuccess) e = hipMemcpyAsync(dgo, dout, v * s.cout * sizeof(float), hipMemcpyHostToDevice, n->st);
        if (e == hipSuccess) e = hipMemsetAsync(ddx, 0, v * cin * sizeof(float), n->st);
        if (e == hipSuccess) e = hipMemsetAsync(ddw, 0, nw * sizeof(float), n->st);
        if (e == hipSuccess)
            e = fwi::unet::conv_dgrad(dgo, s.cout, n->wd[layer], n->kpadd[layer], cin, ddx, B, D, H, W, n->st);
        if (e == hipSuccess)
            e = fwi::unet::conv_wgrad(dgo, s.cout, da, s.ca, db, s.cb, ddw, 27 * cin, B, D, H, W, part, part_cap,
                                      n->st);
        if (e == hipSuccess) e = hipMemcpyAsync(dx, ddx, v * cin * sizeof(float), hipMemcpyDeviceToHost, n->st);
        if (e == hipSuccess) e = hipMemcpyAsync(dw, ddw, nw * sizeof(float), hipMemcpyDeviceToHost, n->st);
        if (e == hipSuccess) e = hipStreamSynchronize(n->st);
        if (e != hipSuccess) rc = ufail(n, FWI_ERR_HIP, hipGetErrorString(e));
    }
    ufree(da);
    ufree(db);
    ufree(dgo);
    ufree(ddx);
    ufree(ddw);
    ufree(part);
    return rc;
}

}  // extern "C"

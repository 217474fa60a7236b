// noalloc.hip -- an entry point that allocates nothing.
//
// fwi_residual_noalloc runs the kernels of fwi_residual (the residual
// r = obs - cal and its FP64 sum of squares, kernels_aux.hip) on caller
// device buffers, with the per-block partial sums in a fixed device array
// of the library instead of a context's allocation.  It is for callers that
// cannot call hipMalloc: code under HIP graph capture, tight loops, hosts
// that forbid allocation.  Every call on a device shares that device's
// array, so calls that overlap on one device race; each call writes every
// partial it reads, so nothing needs clearing between calls.
#include <mutex>
#include <string>

#include "../../include/fwi3d.h"
#include "common.hpp"

namespace fwi {
void set_global_error(const std::string& msg);  // abi.hip: fwi_last_error(NULL)
}

namespace {

// kReduceBlocks partial sums (8 KiB); one copy per device, held until the
// process ends.  Host code never uses this name as an address: it takes the
// device address from hipGetSymbolAddress.
static __device__ __align__(256) double scratch[fwi::kReduceBlocks];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up once
// per device, with that device current).
int scratch_ptr(double** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) {
        fwi::set_global_error("fwi_residual_noalloc: no current HIP device");
        return FWI_ERR_NODEV;
    }
    if (dev >= kMaxDevices) {
        fwi::set_global_error("fwi_residual_noalloc: device index beyond the library's table");
        return FWI_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fwi::set_global_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return FWI_ERR_HIP;
        }
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return FWI_OK;
}

}  // namespace

extern "C" int fwi_residual_noalloc(const float* obs, float* cal, long n, double* sum_sq, void* stream) {
    if (!obs || !cal || !sum_sq) {
        fwi::set_global_error("fwi_residual_noalloc: NULL buffer");
        return FWI_ERR_ARG;
    }
    if (n < 1 || n > FWI_NOALLOC_MAX_SAMPLES) {
        fwi::set_global_error("fwi_residual_noalloc: n must be in [1, FWI_NOALLOC_MAX_SAMPLES = 2^34]");
        return FWI_ERR_RANGE;
    }
    double* partial = nullptr;
    const int rc = scratch_ptr(&partial);
    if (rc != FWI_OK) return rc;
    hipError_t e = fwi::launch_residual(obs, cal, n, partial, sum_sq, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
        fwi::set_global_error(std::string("fwi_residual_noalloc: ") + hipGetErrorString(e));
        return FWI_ERR_HIP;
    }
    return FWI_OK;
}

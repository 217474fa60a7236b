// capi.cpp -- C entry points into the host library for tests: the host-side
// set-up arithmetic (wavelet, dampers, parameter parsing) that feeds the
// device library, and the data-preparation steps of dataset.hpp.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "admin.hpp"
#include "dataset.hpp"

extern "C" {

// Modeling::set_wavelet (modeling.cu:36-59); returns tlag.
int fwi3d_host_wavelet(int nt, float dt, float fmax, float* w) {
    const float t0 = 2.0f * sqrtf((float)M_PI) / fmax;
    const float fc = fmax / (3.0f * sqrtf((float)M_PI));
    for (int n = 0; n < nt; n++) {
        const float td = n * dt - t0;
        const float arg = (float)(M_PI * M_PI * M_PI * fc * fc * td * td);
        w[n] = 1e5f * (1.0f - 2.0f * arg) * expf(-arg);
    }
    return (int)(t0 / dt) + 1;
}

// Modeling::set_abc_dampers 1-D table (modeling.cu:96-99).
void fwi3d_host_b1d(int nb, float abc_factor, float* b1d) {
    for (int i = 0; i < nb; i++) b1d[i] = expf(-powf(abc_factor * (nb - i), 2.0f));
}

// catch_parameter (admin.cpp:53-88); returns the value length or -1.
int fwi3d_host_catch_parameter(const char* target, const char* file, char* out, int cap) {
    try {
        const std::string v = catch_parameter(target, file);
        if ((int)v.size() + 1 > cap) return -1;
        std::memcpy(out, v.c_str(), v.size() + 1);
        return (int)v.size();
    } catch (...) {
        return -1;
    }
}

}  // extern "C"

extern "C" {

// dataset.hpp entry points (generate_setup.py's smoothing,
// generate_input_data.py's normalisation and mute)
void fwi3d_host_gaussian_filter3d(float* v, int n0, int n1, int n2, double sigma) {
    gaussian_filter3d(v, n0, n1, n2, sigma);
}

void fwi3d_host_rtm_input(float* data, int nt, int nr, float dt, float sx, float sy, const float* rx, const float* ry) {
    rtm_input(data, nt, nr, dt, sx, sy, rx, ry);
}

}  // extern "C"

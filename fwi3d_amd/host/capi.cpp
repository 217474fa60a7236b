// capi.cpp -- C entry points into the host library for tests: the host-side
// set-up arithmetic (wavelet, dampers, parameter parsing) that feeds the
// device library, run through the Modeling class's own methods, the
// inversion driver's line-search step, and the data-preparation steps of
// dataset.hpp.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "admin.hpp"
#include "dataset.hpp"
#include "inversion.hpp"
#include "modeling.hpp"

namespace {
// Runs the drivers' own set-up methods (Modeling::set_wavelet,
// Modeling::set_b1d) on given sizes, without a parameters file or device.
struct SetupProbe : Modeling {
    int wavelet_of(int n, float step, float f, float* out) {
        nt = n;
        dt = step;
        fmax = f;
        set_wavelet();
        std::copy(wavelet.begin(), wavelet.end(), out);
        return tlag;
    }
    void b1d_of(int cells, float factor, float* out) {
        nb = cells;
        abc_factor = factor;
        set_b1d();
        std::copy(b1d.begin(), b1d.end(), out);
    }
};
}  // namespace

extern "C" {

// Modeling::set_wavelet (modeling.cu:36-59) on nt samples; returns tlag.
int fwi3d_host_wavelet(int nt, float dt, float fmax, float* w) {
    if (nt < 1 || !w) return -1;
    return SetupProbe().wavelet_of(nt, dt, fmax, w);
}

// Modeling::set_b1d, the 1-D Cerjan table of set_abc_dampers (modeling.cu:96-99).
void fwi3d_host_b1d(int nb, float abc_factor, float* b1d) {
    if (nb > 0 && b1d) SetupProbe().b1d_of(nb, abc_factor, b1d);
}

// parabola_vertex / usable_step: Inversion::optimization's step
// (inversion.cu:291-312 and the guard around it).
float fwi3d_host_parabola_vertex(float a0, float a1, float a2, float f0, float f1, float f2) {
    return parabola_vertex(a0, a1, a2, f0, f1, f2);
}

float fwi3d_host_usable_step(float vertex, float a1, float a2, float f1, float f2, const float* m, const float* g,
                             long n, float vcfl, int* fell_back) {
    bool fb = false;
    const float s = usable_step(vertex, a1, a2, f1, f2, m, g, n, vcfl, &fb);
    if (fell_back) *fell_back = fb ? 1 : 0;
    return s;
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

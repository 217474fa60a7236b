// fwi3d_oracle.cpp -- CPU oracle for the FWI3D wave-propagation path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain C++17/OpenMP, the
// arithmetic of the reference kernels (phbastosa/FWI3D, src/modeling,
// src/migration, src/inversion) so that the HIP library can be checked
// against it.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg load it; the product never links, imports or calls it.
//
// Layout: every padded volume uses the reference host layout
//     c = i + j*nzz + k*nxx*nzz      (i = z fastest, j = x, k = y slowest)
// (modeling.cu:190, modeling.cu:334-336).
//
// Floating-point contract: compiled with -ffp-contract=off, so every
// expression below is evaluated exactly in the order written, with one
// IEEE rounding per operation.  For the float instantiation the
// per-axis second derivative is summed in double, as the reference does
// because its FD coefficients are double literals (modeling.cuh:9-13);
// the pair sums inside it are float additions and the result is rounded
// back to float when stored, exactly as `float d2P_dx2 = (...)` does.
// The double instantiation ("FP64 build") carries every value in double
// and is used only to report FP32 error against an FP64 solution.
//
// Function-by-function pinning (which test checks what) is listed in the
// header of tests/test_oracle.py.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

// modeling.cuh:9-13 -- the reference's 8th-order coefficients, as the same
// truncated double literals.
constexpr double FDM1 = 0.001785714285;
constexpr double FDM2 = 0.025396825396;
constexpr double FDM3 = 0.200000000000;
constexpr double FDM4 = 1.600000000000;
constexpr double FDM5 = 2.847222222222;

inline long lidx(int i, int j, int k, int nxx, int nzz) {
    return (long)i + (long)j * nzz + (long)k * (long)nxx * nzz;
}

// One 9-point second derivative along an axis with stride `s`, following
// modeling.cu:354-358: float pair sums, double coefficients, double
// accumulation left to right, times idh2, rounded to the storage type.
template <typename T> struct Deriv;

template <> struct Deriv<float> {
    static inline float d2(const float* p, long c, long s, float idh2) {
        double d = -FDM1 * (double)(p[c - 4 * s] + p[c + 4 * s]);
        d = d + FDM2 * (double)(p[c - 3 * s] + p[c + 3 * s]);
        d = d - FDM3 * (double)(p[c - 2 * s] + p[c + 2 * s]);
        d = d + FDM4 * (double)(p[c - 1 * s] + p[c + 1 * s]);
        d = d - FDM5 * (double)(p[c]);
        d = d * (double)idh2;
        return (float)d;
    }
};

template <> struct Deriv<double> {
    static inline double d2(const double* p, long c, long s, double idh2) {
        double d = -FDM1 * (p[c - 4 * s] + p[c + 4 * s]);
        d = d + FDM2 * (p[c - 3 * s] + p[c + 3 * s]);
        d = d - FDM3 * (p[c - 2 * s] + p[c + 2 * s]);
        d = d + FDM4 * (p[c - 1 * s] + p[c + 1 * s]);
        d = d - FDM5 * (p[c]);
        return d * idh2;
    }
};

struct Dims {
    int nxx, nyy, nzz;
    long vol() const { return (long)nxx * nyy * nzz; }
};

// get_boundary_damper (modeling.cu:395-520) evaluated from the tables built
// by oracle_dampers.  Returns 1 for cells at least nb away from every face.
inline float damper_at(const float* b1d, const float* b2d, const float* b3d,
                       int i, int j, int k, const Dims& d, int nb) {
    auto dist = [nb](int q, int n) { return q < nb ? q : (q >= n - nb ? n - 1 - q : -1); };
    int di = dist(i, d.nzz), dj = dist(j, d.nxx), dk = dist(k, d.nyy);
    if (di < 0 && dj < 0 && dk < 0) return 1.0f;
    if (dj < 0 && dk < 0) return b1d[di];
    if (di < 0 && dk < 0) return b1d[dj];
    if (di < 0 && dj < 0) return b1d[dk];
    if (di < 0) return b2d[dj + dk * nb];          // modeling.cu:434-449
    if (dj < 0) return b2d[di + dk * nb];          // modeling.cu:451-466
    if (dk < 0) return b2d[di + dj * nb];          // modeling.cu:468-483
    return b3d[di + dj * nb + (long)dk * nb * nb]; // modeling.cu:486-517
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// Setup arithmetic
// ---------------------------------------------------------------------------

// Modeling::set_wavelet, modeling.cu:36-59.  Returns tlag.
int oracle_wavelet(int nt, float dt, float fmax, float* w) {
    float t0 = 2.0f * sqrtf((float)M_PI) / fmax;
    float fc = fmax / (3.0f * sqrtf((float)M_PI));
    int tlag = (int)(t0 / dt) + 1;
    for (int n = 0; n < nt; n++) {
        float td = n * dt - t0;
        float arg = (float)(M_PI * M_PI * M_PI * fc * fc * td * td);
        w[n] = 1e5f * (1.0f - 2.0f * arg) * expf(-arg);
    }
    return tlag;
}

// Modeling::set_abc_dampers, modeling.cu:83-140 (tables after the final
// subtractions).  b1d: nb, b2d: nb*nb, b3d: nb*nb*nb.
void oracle_dampers(int nb, float abc_factor, float* b1d, float* b2d, float* b3d) {
    std::vector<float> b2(nb * nb, 0.0f);
    for (int i = 0; i < nb; i++) b1d[i] = expf(-powf(abc_factor * (nb - i), 2.0f));
    for (int i = 0; i < nb; i++)
        for (int j = 0; j < nb; j++) {
            b2[j + i * nb] += b1d[i];
            b2[i + j * nb] += b1d[i];
        }
    long n3 = (long)nb * nb * nb;
    for (long q = 0; q < n3; q++) b3d[q] = 0.0f;
    for (int i = 0; i < nb; i++)
        for (int j = 0; j < nb; j++)
            for (int k = 0; k < nb; k++) {
                long q = i + j * nb + (long)k * nb * nb;
                b3d[q] += b2[i + j * nb];
                b3d[q] += b2[j + k * nb];
                b3d[q] += b2[i + k * nb];
            }
    for (int q = 0; q < nb * nb; q++) b2d[q] = b2[q] - 1.0f;
    for (long q = 0; q < n3; q++) b3d[q] -= 5.0f;
}

// Damper value of one padded cell (modeling.cu:395-520), from the tables.
float oracle_damper_at(const float* b1d, const float* b2d, const float* b3d, int i, int j, int k,
                       int nxx, int nyy, int nzz, int nb) {
    Dims d{nxx, nyy, nzz};
    return damper_at(b1d, b2d, b3d, i, j, k, d, nb);
}

// Modeling::expand_boundary, modeling.cu:181-228: edge-replicating padding.
void oracle_expand_boundary(const float* in, float* out, int nx, int ny, int nz, int nb) {
    int nxx = nx + 2 * nb, nyy = ny + 2 * nb, nzz = nz + 2 * nb;
    for (long index = 0; index < (long)nx * ny * nz; index++) {
        int k = (int)(index / ((long)nx * nz));
        int j = (int)((index - (long)k * nx * nz) / nz);
        int i = (int)(index - (long)j * nz - (long)k * nx * nz);
        out[lidx(i + nb, j + nb, k + nb, nxx, nzz)] = in[index];
    }
    for (int k = nb; k < nyy - nb; k++)
        for (int j = nb; j < nxx - nb; j++)
            for (int i = 0; i < nb; i++) {
                out[lidx(i, j, k, nxx, nzz)] = in[0 + (long)(j - nb) * nz + (long)(k - nb) * nx * nz];
                out[lidx(nzz - i - 1, j, k, nxx, nzz)] =
                    in[(nz - 1) + (long)(j - nb) * nz + (long)(k - nb) * nx * nz];
            }
    for (int k = 0; k < nyy; k++)
        for (int j = 0; j < nb; j++)
            for (int i = 0; i < nzz; i++) {
                out[lidx(i, j, k, nxx, nzz)] = out[lidx(i, nb, k, nxx, nzz)];
                out[lidx(i, nxx - j - 1, k, nxx, nzz)] = out[lidx(i, nxx - nb - 1, k, nxx, nzz)];
            }
    for (int k = 0; k < nb; k++)
        for (int j = 0; j < nxx; j++)
            for (int i = 0; i < nzz; i++) {
                out[lidx(i, j, k, nxx, nzz)] = out[lidx(i, j, nb, nxx, nzz)];
                out[lidx(i, j, nyy - k - 1, nxx, nzz)] = out[lidx(i, j, nyy - nb - 1, nxx, nzz)];
            }
}

// Modeling::reduce_boundary, modeling.cu:230-241: crop to the interior.
void oracle_reduce_boundary(const float* in, float* out, int nx, int ny, int nz, int nb) {
    int nxx = nx + 2 * nb, nzz = nz + 2 * nb;
    for (long index = 0; index < (long)nx * ny * nz; index++) {
        int k = (int)(index / ((long)nx * nz));
        int j = (int)((index - (long)k * nx * nz) / nz);
        int i = (int)(index - (long)j * nz - (long)k * nx * nz);
        out[index] = in[lidx(i + nb, j + nb, k + nb, nxx, nzz)];
    }
}

// Modeling::initialization, modeling.cu:270-272, 286-288: metres -> padded index.
int oracle_position_index(float x, float dh, int nb) { return (int)((x + 0.5f * dh) / dh) + nb; }

// ---------------------------------------------------------------------------
// Seeded random boundary (this build's generator; the reference's curand /
// Poisson-disk generator at modeling.cu:558-673 is not reproducible).  The
// same counter-based hash and the same float expression order are used by
// the HIP kernel, so the two must agree bit for bit.
// ---------------------------------------------------------------------------

static inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

float oracle_rbc_uniform(uint64_t seed, int shot, long cell) {
    uint64_t key = (seed * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)(uint32_t)shot * 0xD1B54A32D192ED03ull) ^
                   (uint64_t)cell;
    uint64_t h = splitmix64(key);
    int u24 = (int)(h >> 40);
    return (float)(u24 - 8388608) * (1.0f / 8388608.0f);
}

// Random boundary over an already edge-padded model (random_boundary_bg,
// modeling.cu:617-666, restated with uniform noise and one ramp):
//   v = v_face * (1 + amp * ramp * u),  ramp = max over boundary axes of
//   1 - d/nb (d = cells from the outer edge), u ~ U[-1, 1) from the hash,
//   then v = min(v, vclamp) when vclamp > 0.
void oracle_random_boundary(float* vp, int nxx, int nyy, int nzz, int nb, uint64_t seed, int shot,
                            float amp, float vclamp) {
    std::vector<float> src(vp, vp + (long)nxx * nyy * nzz);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nyy; k++)
        for (int j = 0; j < nxx; j++)
            for (int i = 0; i < nzz; i++) {
                bool in_i = i >= nb && i < nzz - nb, in_j = j >= nb && j < nxx - nb, in_k = k >= nb && k < nyy - nb;
                if (in_i && in_j && in_k) continue;
                int ii = std::min(std::max(i, nb), nzz - nb - 1);
                int jj = std::min(std::max(j, nb), nxx - nb - 1);
                int kk = std::min(std::max(k, nb), nyy - nb - 1);
                float v = src[lidx(ii, jj, kk, nxx, nzz)];
                float ramp = 0.0f;
                if (!in_j) { int d = j < nb ? j : nxx - j - 1; ramp = std::max(ramp, 1.0f - (float)d / (float)nb); }
                if (!in_k) { int d = k < nb ? k : nyy - k - 1; ramp = std::max(ramp, 1.0f - (float)d / (float)nb); }
                if (!in_i) { int d = i < nb ? i : nzz - i - 1; ramp = std::max(ramp, 1.0f - (float)d / (float)nb); }
                long c = lidx(i, j, k, nxx, nzz);
                float u = oracle_rbc_uniform(seed, shot, c);
                float t = amp * ramp;
                t = t * u;
                float f = 1.0f + t;
                float val = v * f;
                if (vclamp > 0.0f && val > vclamp) val = vclamp;
                vp[c] = val;
            }
}

// Reference-style random boundary (random_boundary_bg + random_boundary_gp,
// modeling.cu:558-666) made deterministic: each boundary cell starts from
// the nearest interior cell and gets per-axis background noise
// (1 - d/nb) * var_vp * n with n the Irwin-Hall sum of 12 hashed uniforms
// minus 6 (x, then y, then z); then the blobs (xc, yc, zc, r, amp rows) in
// order, each added as amp * e^(-q/2) with q = |X - c|^2 / r^2 when q <= 16,
// every blob followed by the clamp to [vlo, vhi]; finally min(v, vclamp).
// e^x is evaluated by exp_det below.
static double exp_det(double x) {
    // 2^k e^r with r = x - k ln2, e^r by its degree-14 Taylor polynomial
    const double c[15] = {1.0,          1.0,           1.0 / 2.0,       1.0 / 6.0,        1.0 / 24.0,
                          1.0 / 120.0,  1.0 / 720.0,   1.0 / 5040.0,    1.0 / 40320.0,    1.0 / 362880.0,
                          1.0 / 3628800.0, 1.0 / 39916800.0, 1.0 / 479001600.0, 1.0 / 6227020800.0,
                          1.0 / 87178291200.0};
    const double k = std::floor(x * 1.4426950408889634 + 0.5);
    const double r = x - k * 0.6931471805599453;
    double p = c[14];
    for (int n = 13; n >= 0; --n) p = c[n] + r * p;
    return std::ldexp(p, (int)k);
}

double oracle_exp_det(double x) { return exp_det(x); }

static float irwin_hall(uint64_t base, long cell, int axis) {
    float s = 0.0f;
    for (int m = 0; m < 12; ++m) {
        const uint64_t h = splitmix64(base ^ ((uint64_t)cell * 36u + (uint64_t)(axis * 12 + m)));
        s = s + (float)(uint32_t)(h >> 40) * (1.0f / 16777216.0f);
    }
    return s - 6.0f;
}

float oracle_irwin_hall(uint64_t seed, int shot, long cell, int axis) {
    const uint64_t base = (seed * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)(uint32_t)shot * 0xD1B54A32D192ED03ull) ^
                          0x5851F42D4C957F2Dull;
    return irwin_hall(base, cell, axis);
}

void oracle_random_boundary_blobs(float* vp, int nxx, int nyy, int nzz, int nb, float dh, uint64_t seed, int shot,
                                  float var_vp, const float* blobs, int nblobs, float vlo, float vhi, float vclamp) {
    const uint64_t base = (seed * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)(uint32_t)shot * 0xD1B54A32D192ED03ull) ^
                          0x5851F42D4C957F2Dull;
#pragma omp parallel for schedule(dynamic)
    for (int k = 0; k < nyy; k++)
        for (int j = 0; j < nxx; j++)
            for (int i = 0; i < nzz; i++) {
                const bool in_x = j >= nb && j < nxx - nb, in_y = k >= nb && k < nyy - nb,
                           in_z = i >= nb && i < nzz - nb;
                if (in_x && in_y && in_z) continue;
                const long c = lidx(i, j, k, nxx, nzz);
                // start from the nearest interior cell (random_boundary_bg's index_ref)
                const int ii = i < nb ? nb : (i > nzz - nb - 1 ? nzz - nb - 1 : i);
                const int jj = j < nb ? nb : (j > nxx - nb - 1 ? nxx - nb - 1 : j);
                const int kk = k < nb ? nb : (k > nyy - nb - 1 ? nyy - nb - 1 : k);
                float v = vp[lidx(ii, jj, kk, nxx, nzz)];
                if (!in_x) {
                    const int d = j < nb ? j : nxx - j - 1;
                    const float f = 1.0f - (float)d / (float)nb;
                    v = v + (f * var_vp) * irwin_hall(base, c, 0);
                }
                if (!in_y) {
                    const int d = k < nb ? k : nyy - k - 1;
                    const float f = 1.0f - (float)d / (float)nb;
                    v = v + (f * var_vp) * irwin_hall(base, c, 1);
                }
                if (!in_z) {
                    const int d = i < nb ? i : nzz - i - 1;
                    const float f = 1.0f - (float)d / (float)nb;
                    v = v + (f * var_vp) * irwin_hall(base, c, 2);
                }
                const float X = (float)j * dh, Y = (float)k * dh, Z = (float)i * dh;
                for (int q = 0; q < nblobs; ++q) {
                    const float* b = blobs + 5 * q;
                    const double r = (double)b[3];
                    const double ax = ((double)X - (double)b[0]) / r, ay = ((double)Y - (double)b[1]) / r,
                                 az = ((double)Z - (double)b[2]) / r;
                    const double qq = (ax * ax + ay * ay) + az * az;
                    if (qq <= 16.0) v = v + (float)((double)b[4] * exp_det(-0.5 * qq));
                    // random_boundary_gp clamps after every blob, near or far
                    v = v > vhi ? vhi : v;
                    v = v < vlo ? vlo : v;
                }
                if (vclamp > 0.0f && v > vclamp) v = vclamp;
                vp[c] = v;
            }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Propagation
// ---------------------------------------------------------------------------

namespace {

// One interior update B = dt^2 vp^2 L(A) + 2A - B over all cells with
// 4 <= idx < n-4 (compute_pressure, modeling.cu:350-372).
template <typename T>
void stencil_update(const float* vp, const T* A, T* B, const Dims& d, T idh2, T dt) {
    const long sx = d.nzz, sy = (long)d.nxx * d.nzz;
#pragma omp parallel for schedule(static)
    for (int k = 4; k < d.nyy - 4; k++)
        for (int j = 4; j < d.nxx - 4; j++) {
            long base = (long)j * sx + (long)k * sy;
            for (int i = 4; i < d.nzz - 4; i++) {
                long c = base + i;
                T v = (T)vp[c];
                T vp2 = v * v;
                T dxx = Deriv<T>::d2(A, c, sx, idh2);
                T dyy = Deriv<T>::d2(A, c, sy, idh2);
                T dzz = Deriv<T>::d2(A, c, 1, idh2);
                B[c] = dt * dt * vp2 * (dxx + dyy + dzz) + (T)2 * A[c] - B[c];
            }
        }
}

// ABC damping applied after every stencil read of the step
// (modeling.cu:374-380; the in-place race of the reference is defined as
// damp-after-read).
template <typename T>
void damp_interior(T* A, T* B, const float* b1d, const float* b2d, const float* b3d, const Dims& d, int nb) {
#pragma omp parallel for schedule(static)
    for (int k = 4; k < d.nyy - 4; k++)
        for (int j = 4; j < d.nxx - 4; j++)
            for (int i = 4; i < d.nzz - 4; i++) {
                long c = lidx(i, j, k, d.nxx, d.nzz);
                T damper = (T)damper_at(b1d, b2d, b3d, i, j, k, d, nb);
                A[c] *= damper;
                B[c] *= damper;
            }
}

template <typename T>
void forward_impl(const float* vp, int nxx, int nyy, int nzz, int nb, float dh, float dt, int nt, int tlag,
                  const float* wavelet, int sIdx, int sIdy, int sIdz, int nrec, const int* rIdx, const int* rIdy,
                  const int* rIdz, int abc, const float* b1d, const float* b2d, const float* b3d, int n_begin,
                  int n_end, T* P, T* Pold, T* seis) {
    Dims d{nxx, nyy, nzz};
    const float idh2f = 1.0f / (dh * dh);
    const float idh3f = 1.0f / (dh * dh * dh);
    const T idh2 = (T)idh2f, idh3 = (T)idh3f, tdt = (T)dt;
    T* A = P;
    T* B = Pold;
    const long src = lidx(sIdz, sIdx, sIdy, nxx, nzz);
    for (int n = n_begin; n < n_end; n++) {
        if (n < nt) A[src] += idh3 * (T)wavelet[n];            // modeling.cu:347-348
        stencil_update<T>(vp, A, B, d, idh2, tdt);             // modeling.cu:350-372
        if (abc) damp_interior<T>(A, B, b1d, b2d, b3d, d, nb); // modeling.cu:374-380
        if (seis && n > tlag)                                  // modeling.cu:391-392
            for (int r = 0; r < nrec; r++) seis[(long)(n - tlag) + (long)r * nt] = A[lidx(rIdz[r], rIdx[r], rIdy[r], nxx, nzz)];
        std::swap(A, B);                                       // modeling.cu:320
    }
    if (A != P) {  // odd number of steps: hand back (current, previous) in (P, Pold)
        long n = d.vol();
        std::vector<T> tmp(A, A + n);
        std::memcpy(Pold, P, n * sizeof(T));
        std::memcpy(P, tmp.data(), n * sizeof(T));
    }
}

// Backward pass: time-reversed source reconstruction + receiver-side
// propagation + imaging (cross_correlation, migration.cu:134-202) or
// gradient (build_gradient, inversion.cu:400-468), with the receiver
// injection of inject_seismogram / inject_residuo (migration.cu:119-132,
// inversion.cu:385-398) applied in receiver order before the step.
//
// phys (0 = the reference) switches on corrections the reference lacks:
//   1: FWI gradient with d2/dy2 (inversion.cu:466 omits it);
//   2: source subtraction -- the newer source-side field of backward step b
//      is the forward field of step nsteps - b, which carries the source term
//      idh3 * wavelet[nsteps - b] (for nsteps - b < nt); it is added back to
//      the reconstructed field at the source cell, which makes the
//      reconstruction the exact reverse of the forward recurrence;
//   4: aligned imaging -- the image and illumination use the current source
//      field So (the time of the receiver field R) instead of the
//      reconstructed one.
// One backward step b: receiver injection into Pr, then the reconstruction
// Ps <- step(Psold, Ps), the receiver update Prold <- step(Pr, Prold) and the
// imaging terms added to image / illum (the caller swaps Ps/Psold and
// Pr/Prold afterwards).
template <typename T>
void backward_step(const float* vp, int nxx, int nyy, int nzz, float dh, float dt, int nt, const float* data,
                   int nrec, const int* rIdx, const int* rIdy, const int* rIdz, int fwi, int b, T* Ps, T* Psold,
                   T* Pr, T* Prold, T* image, T* illum, int phys, const float* wavelet, long src, int nsteps) {
    const float idh2f = 1.0f / (dh * dh);
    const float idh3f = 1.0f / (dh * dh * dh);
    const T idh2 = (T)idh2f, idh3 = (T)idh3f, tdt = (T)dt;
    const long sx = nzz, sy = (long)nxx * nzz;
    if (b < nt)
        for (int r = 0; r < nrec; r++)
            Pr[lidx(rIdz[r], rIdx[r], rIdy[r], nxx, nzz)] += idh3 * (T)data[(long)(nt - b - 1) + (long)r * nt];
    const int ns = nsteps - b;
    const bool sub = (phys & 2) && src >= 0 && ns >= 0 && ns < nt;
    const T src_term = sub ? idh3 * (T)wavelet[ns] : (T)0;
#pragma omp parallel for schedule(static)
    for (int k = 4; k < nyy - 4; k++)
        for (int j = 4; j < nxx - 4; j++) {
            long base = (long)j * sx + (long)k * sy;
            for (int i = 4; i < nzz - 4; i++) {
                long c = base + i;
                T v = (T)vp[c];
                T vp2 = v * v;
                T sxx = Deriv<T>::d2(Psold, c, sx, idh2);
                T syy = Deriv<T>::d2(Psold, c, sy, idh2);
                T szz = Deriv<T>::d2(Psold, c, 1, idh2);
                T rxx = Deriv<T>::d2(Pr, c, sx, idh2);
                T ryy = Deriv<T>::d2(Pr, c, sy, idh2);
                T rzz = Deriv<T>::d2(Pr, c, 1, idh2);
                T sn = tdt * tdt * vp2 * (sxx + syy + szz) + (T)2 * Psold[c] - Ps[c];
                if (sub && c == src) sn = sn + src_term;
                Ps[c] = sn;
                Prold[c] = tdt * tdt * vp2 * (rxx + ryy + rzz) + (T)2 * Pr[c] - Prold[c];
                const T si = (phys & 4) ? Psold[c] : Ps[c];
                illum[c] += si * si;
                if (fwi)
                    image[c] += tdt * Pr[c] * ((phys & 1) ? sxx + syy + szz : sxx + szz) * vp2;  // inversion.cu:466
                else
                    image[c] += si * Pr[c];  // migration.cu:200
            }
        }
}

// Backward pass: time-reversed source reconstruction + receiver-side
// propagation + imaging (cross_correlation, migration.cu:134-202) or
// gradient (build_gradient, inversion.cu:400-468), with the receiver
// injection of inject_seismogram / inject_residuo (migration.cu:119-132,
// inversion.cu:385-398) applied in receiver order before the step.
//
// phys (0 = the reference) switches on corrections the reference lacks:
//   1: FWI gradient with d2/dy2 (inversion.cu:466 omits it);
//   2: source subtraction -- the newer source-side field of backward step b
//      is the forward field of step nsteps - b, which carries the source term
//      idh3 * wavelet[nsteps - b] (for nsteps - b < nt); it is added back to
//      the reconstructed field at the source cell, which makes the
//      reconstruction the exact reverse of the forward recurrence;
//   4: aligned imaging -- the image and illumination use the current source
//      field So (the time of the receiver field R) instead of the
//      reconstructed one.
template <typename T>
void backward_impl(const float* vp, int nxx, int nyy, int nzz, float dh, float dt, int nt, const float* data,
                   int nrec, const int* rIdx, const int* rIdy, const int* rIdz, int fwi, int b_begin, int b_end,
                   T* S, T* So, T* R, T* Ro, T* image, T* illum, int phys = 0, const float* wavelet = nullptr,
                   long src = -1, int nsteps = 0) {
    Dims d{nxx, nyy, nzz};
    T* Ps = S;
    T* Psold = So;
    T* Pr = R;
    T* Prold = Ro;
    for (int b = b_begin; b < b_end; b++) {
        backward_step<T>(vp, nxx, nyy, nzz, dh, dt, nt, data, nrec, rIdx, rIdy, rIdz, fwi, b, Ps, Psold, Pr, Prold,
                         image, illum, phys, wavelet, src, nsteps);
        std::swap(Ps, Psold);
        std::swap(Pr, Prold);
    }
    long n = d.vol();
    if (Ps != S) {
        std::vector<T> tmp(Ps, Ps + n);
        std::memcpy(So, S, n * sizeof(T));
        std::memcpy(S, tmp.data(), n * sizeof(T));
    }
    if (Pr != R) {
        std::vector<T> tmp(Pr, Pr + n);
        std::memcpy(Ro, R, n * sizeof(T));
        std::memcpy(R, tmp.data(), n * sizeof(T));
    }
}

}  // namespace

extern "C" {

// Runs steps [n_begin, n_end) of Modeling::forward_solver
// (modeling.cu:309-322) from the state (P = current, Pold = previous) and
// leaves (current, previous) in (P, Pold).  seis (nt*nrec, layout t + r*nt)
// may be NULL; samples are written only for n > tlag.  abc != 0 applies
// the Cerjan damper from (b1d, b2d, b3d).
void oracle_forward_f32(const float* vp, int nxx, int nyy, int nzz, int nb, float dh, float dt, int nt, int tlag,
                        const float* wavelet, int sIdx, int sIdy, int sIdz, int nrec, const int* rIdx,
                        const int* rIdy, const int* rIdz, int abc, const float* b1d, const float* b2d,
                        const float* b3d, int n_begin, int n_end, float* P, float* Pold, float* seis) {
    forward_impl<float>(vp, nxx, nyy, nzz, nb, dh, dt, nt, tlag, wavelet, sIdx, sIdy, sIdz, nrec, rIdx, rIdy, rIdz,
                        abc, b1d, b2d, b3d, n_begin, n_end, P, Pold, seis);
}

void oracle_forward_f64(const float* vp, int nxx, int nyy, int nzz, int nb, float dh, float dt, int nt, int tlag,
                        const float* wavelet, int sIdx, int sIdy, int sIdz, int nrec, const int* rIdx,
                        const int* rIdy, const int* rIdz, int abc, const float* b1d, const float* b2d,
                        const float* b3d, int n_begin, int n_end, double* P, double* Pold, double* seis) {
    forward_impl<double>(vp, nxx, nyy, nzz, nb, dh, dt, nt, tlag, wavelet, sIdx, sIdy, sIdz, nrec, rIdx, rIdy, rIdz,
                         abc, b1d, b2d, b3d, n_begin, n_end, P, Pold, seis);
}

// Runs backward steps [b_begin, b_end) of Migration::backward_propagation
// (migration.cu:61-81; fwi = 0) or Inversion::backward_propagation
// (inversion.cu:235-253; fwi = 1).  S/So are the source-side fields
// (d_P/d_Pold at the end of the forward pass), R/Ro the receiver-side
// fields; `image` receives the image (RTM) or the gradient (FWI).
void oracle_backward_f32(const float* vp, int nxx, int nyy, int nzz, float dh, float dt, int nt,
                         const float* data, int nrec, const int* rIdx, const int* rIdy, const int* rIdz, int fwi,
                         int b_begin, int b_end, float* S, float* So, float* R, float* Ro, float* image,
                         float* illum) {
    backward_impl<float>(vp, nxx, nyy, nzz, dh, dt, nt, data, nrec, rIdx, rIdy, rIdz, fwi, b_begin, b_end, S, So, R,
                         Ro, image, illum);
}

void oracle_backward_f64(const float* vp, int nxx, int nyy, int nzz, float dh, float dt, int nt,
                         const float* data, int nrec, const int* rIdx, const int* rIdy, const int* rIdz, int fwi,
                         int b_begin, int b_end, double* S, double* So, double* R, double* Ro, double* image,
                         double* illum) {
    backward_impl<double>(vp, nxx, nyy, nzz, dh, dt, nt, data, nrec, rIdx, rIdy, rIdz, fwi, b_begin, b_end, S, So,
                          R, Ro, image, illum);
}

// As oracle_backward_f32/f64 with the corrections `phys` (see backward_impl);
// wavelet (nt samples), the source cell (sIdx, sIdy, sIdz) and the forward
// pass length nsteps = nt + tlag serve the source subtraction.
void oracle_backward_phys_f32(const float* vp, int nxx, int nyy, int nzz, float dh, float dt, int nt,
                              const float* data, int nrec, const int* rIdx, const int* rIdy, const int* rIdz, int fwi,
                              int b_begin, int b_end, float* S, float* So, float* R, float* Ro, float* image,
                              float* illum, int phys, const float* wavelet, int sIdx, int sIdy, int sIdz, int nsteps) {
    backward_impl<float>(vp, nxx, nyy, nzz, dh, dt, nt, data, nrec, rIdx, rIdy, rIdz, fwi, b_begin, b_end, S, So, R,
                         Ro, image, illum, phys, wavelet, lidx(sIdz, sIdx, sIdy, nxx, nzz), nsteps);
}

void oracle_backward_phys_f64(const float* vp, int nxx, int nyy, int nzz, float dh, float dt, int nt,
                              const float* data, int nrec, const int* rIdx, const int* rIdy, const int* rIdz, int fwi,
                              int b_begin, int b_end, double* S, double* So, double* R, double* Ro, double* image,
                              double* illum, int phys, const float* wavelet, int sIdx, int sIdy, int sIdz,
                              int nsteps) {
    backward_impl<double>(vp, nxx, nyy, nzz, dh, dt, nt, data, nrec, rIdx, rIdy, rIdz, fwi, b_begin, b_end, S, So,
                          R, Ro, image, illum, phys, wavelet, lidx(sIdz, sIdx, sIdy, nxx, nzz), nsteps);
}

// A batch of nshots shots advanced step by step, the shots in slot order
// within every step, shot s adding its terms into the image / illumination
// of group s / sb (the summation order of a batched backward pass with sb
// shots per block; the caller folds the groups in group order afterwards).
// Per shot: vps[s], datas[s] (nrecs[s] x nt), receivers rx/ry/rz[s], source
// src[3s..3s+2] (x, y, z), fields S/So/R/Ro[s], handed back as (current,
// previous) like oracle_backward_f32.  Steps [b_begin, b_end); phys as in
// backward_impl.
void oracle_backward_batched_f32(const float* const* vps, int nshots, int sb, int nxx, int nyy, int nzz, float dh,
                                 float dt, int nt, const float* const* datas, const int* nrecs,
                                 const int* const* rIdx, const int* const* rIdy, const int* const* rIdz, int fwi,
                                 int b_begin, int b_end, float* const* S, float* const* So, float* const* R,
                                 float* const* Ro, float* const* image, float* const* illum, int phys,
                                 const float* wavelet, const int* src, int nsteps) {
    std::vector<float*> ps(S, S + nshots), pso(So, So + nshots), pr(R, R + nshots), pro(Ro, Ro + nshots);
    for (int b = b_begin; b < b_end; b++)
        for (int s = 0; s < nshots; s++) {
            const long sc = src[3 * s] >= 0 ? lidx(src[3 * s + 2], src[3 * s], src[3 * s + 1], nxx, nzz) : -1;
            backward_step<float>(vps[s], nxx, nyy, nzz, dh, dt, nt, datas[s], nrecs[s], rIdx[s], rIdy[s], rIdz[s],
                                 fwi, b, ps[s], pso[s], pr[s], pro[s], image[s / sb], illum[s / sb], phys, wavelet,
                                 sc, nsteps);
            std::swap(ps[s], pso[s]);
            std::swap(pr[s], pro[s]);
        }
    const long n = (long)nxx * nyy * nzz;
    if ((b_end - b_begin) % 2 == 1 && b_end > b_begin)
        for (int s = 0; s < nshots; s++) {
            std::vector<float> tmp(S[s], S[s] + n);
            std::memcpy(S[s], So[s], n * sizeof(float));
            std::memcpy(So[s], tmp.data(), n * sizeof(float));
            tmp.assign(R[s], R[s] + n);
            std::memcpy(R[s], Ro[s], n * sizeof(float));
            std::memcpy(Ro[s], tmp.data(), n * sizeof(float));
        }
}

// ---------------------------------------------------------------------------
// Post-processing and optimisation arithmetic
// ---------------------------------------------------------------------------

// Migration::export_seismic, migration.cu:90-117: crop image and
// illumination, divide, then minus the z second difference.
void oracle_export_seismic(const float* image_pad, const float* illum_pad, int nx, int ny, int nz, int nb, float dh,
                           float* out) {
    long n = (long)nx * ny * nz;
    std::vector<float> image(n), sumPs(n);
    oracle_reduce_boundary(image_pad, image.data(), nx, ny, nz, nb);
    oracle_reduce_boundary(illum_pad, sumPs.data(), nx, ny, nz, nb);
    const float idh2 = 1.0f / (dh * dh);
    for (long q = 0; q < n; q++) sumPs[q] = image[q] / sumPs[q];
    for (long q = 0; q < n; q++) {
        int k = (int)(q / ((long)nx * nz));
        int j = (int)((q - (long)k * nx * nz) / nz);
        int i = (int)(q - (long)j * nz - (long)k * nx * nz);
        out[q] = 0.0f;
        if (i > 0 && i < nz - 1) out[q] = -1.0f * (sumPs[q - 1] - 2.0f * sumPs[q] + sumPs[q + 1]) * idh2;
    }
}

// Inversion::compute_gradient post-processing, inversion.cu:142-169:
// gradient/illumination, masked minus z second difference, scaled by
// 1/max|g|.  Returns max|g| (computed as a true reduction).
float oracle_gradient_post(const float* grad_pad, const float* illum_pad, int nx, int ny, int nz, int nb, float dh,
                           float zmask, float* out) {
    long n = (long)nx * ny * nz;
    std::vector<float> g(n), p(n);
    oracle_reduce_boundary(grad_pad, g.data(), nx, ny, nz, nb);
    oracle_reduce_boundary(illum_pad, p.data(), nx, ny, nz, nb);
    const float idh2 = 1.0f / (dh * dh);
    for (long q = 0; q < n; q++) p[q] = g[q] / p[q];
    float gmax = -1e9f;
    for (long q = 0; q < n; q++) {
        int i = (int)(q % nz);
        out[q] = 0.0f;
        if (i > (int)(zmask / dh) && i < nz - 1) out[q] = -1.0f * (p[q - 1] - 2.0f * p[q] + p[q + 1]) * idh2;
        gmax = gmax < fabsf(out[q]) ? fabsf(out[q]) : gmax;
    }
    for (long q = 0; q < n; q++) out[q] *= 1.0f / gmax;
    return gmax;
}

// Inversion::set_adjoint_source, inversion.cu:190-201.  residual = obs - cal
// (float), returns sum r^2 in long double; *sum_f32 receives the
// reference's own sequential float accumulation.
double oracle_misfit(const float* obs, const float* cal, long n, float* residual, float* sum_f32) {
    long double acc = 0.0L;
    float facc = 0.0f;
    for (long q = 0; q < n; q++) {
        float r = obs[q] - cal[q];
        if (residual) residual[q] = r;
        facc += r * r;
        acc += (long double)r * (long double)r;
    }
    if (sum_f32) *sum_f32 = facc;
    return (double)acc;
}

// Inversion::optimization, inversion.cu:291-312: vertex of the parabola
// through (a0,f0), (a1,f1), (a2,f2), in float as the reference.
float oracle_parabola_step(float a0, float a1, float a2, float f0, float f1, float f2) {
    float num = (a1 * a1 - a2 * a2) * f0 + (a2 * a2 - a0 * a0) * f1 + (a0 * a0 - a1 * a1) * f2;
    float den = (a1 - a2) * f0 + (a2 - a0) * f1 + (a0 - a1) * f2;
    return 0.5f * (num / den);
}

// Inversion::linesearch / update_model model arithmetic
// (inversion.cu:317-321, 351-356): m' = m - a*g, v = 1/sqrt(m').
void oracle_model_step(const float* m, const float* g, float alpha, long n, float* m_out, float* v_out) {
    for (long q = 0; q < n; q++) {
        float mm = m[q] - alpha * g[q];
        if (m_out) m_out[q] = mm;
        v_out[q] = 1.0f / sqrtf(mm);
    }
}

// Discrete Laplacian Dxx + Dyy + Dzz of a padded field over the updated
// interior (the bracket of modeling.cu:372), zero elsewhere.
void oracle_laplacian_f32(const float* p, int nxx, int nyy, int nzz, float dh, float* out) {
    const float idh2 = 1.0f / (dh * dh);
    const long sx = nzz, sy = (long)nxx * nzz;
    std::memset(out, 0, sizeof(float) * (size_t)nxx * nyy * nzz);
    for (int k = 4; k < nyy - 4; k++)
        for (int j = 4; j < nxx - 4; j++)
            for (int i = 4; i < nzz - 4; i++) {
                long c = lidx(i, j, k, nxx, nzz);
                out[c] = Deriv<float>::d2(p, c, sx, idh2) + Deriv<float>::d2(p, c, sy, idh2) +
                         Deriv<float>::d2(p, c, 1, idh2);
            }
}

void oracle_laplacian_f64(const double* p, int nxx, int nyy, int nzz, double dh, double* out) {
    const double idh2 = 1.0 / (dh * dh);
    const long sx = nzz, sy = (long)nxx * nzz;
    std::memset(out, 0, sizeof(double) * (size_t)nxx * nyy * nzz);
    for (int k = 4; k < nyy - 4; k++)
        for (int j = 4; j < nxx - 4; j++)
            for (int i = 4; i < nzz - 4; i++) {
                long c = lidx(i, j, k, nxx, nzz);
                out[c] = Deriv<double>::d2(p, c, sx, idh2) + Deriv<double>::d2(p, c, sy, idh2) +
                         Deriv<double>::d2(p, c, 1, idh2);
            }
}

int oracle_num_threads() { return omp_get_max_threads(); }

}  // extern "C"

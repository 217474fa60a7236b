// dataset.cpp -- see dataset.hpp.
#include "dataset.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>

std::vector<double> linspace(double start, double stop, int num) {
    std::vector<double> v(num > 0 ? num : 0);
    if (num == 1) v[0] = start;
    if (num < 2) return v;
    const double step = (stop - start) / (double)(num - 1);
    for (int i = 0; i < num; i++) v[i] = (double)i * step + start;
    v[num - 1] = stop;
    return v;
}

namespace {

void write_grid(const std::string& path, const std::vector<double>& xs, const std::vector<double>& ys, double z) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::invalid_argument("Error: " + path + " could not be opened!");
    // meshgrid(x, y) flattened in Fortran order: y varies fastest
    for (double x : xs)
        for (double y : ys) std::fprintf(f, "%.2f,%.2f,%.2f\n", x, y, z);
    std::fclose(f);
}

// Sum in numpy's pairwise order for arrays of at most 128 elements (eight
// interleaved partial sums, combined pairwise, then the remainder).
double numpy_sum(const std::vector<double>& a) {
    const size_t n = a.size();
    if (n < 8) {
        double r = 0.0;
        for (double x : a) r += x;
        return r;
    }
    double r[8];
    for (int j = 0; j < 8; j++) r[j] = a[j];
    size_t i = 8;
    for (; i < n - (n % 8); i += 8)
        for (int j = 0; j < 8; j++) r[j] += a[i + j];
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; i++) res += a[i];
    return res;
}

// Normalised Gaussian weights w[-radius..radius], radius = (int)(4 sigma + 0.5).
std::vector<double> gaussian_weights(double sigma, int& radius) {
    radius = (int)(4.0 * sigma + 0.5);
    std::vector<double> phi(2 * radius + 1);
    const double c = -0.5 / (sigma * sigma);
    for (int i = 0; i < 2 * radius + 1; i++) {
        const long x = i - radius;
        phi[i] = std::exp(c * (double)(x * x));
    }
    const double s = numpy_sum(phi);
    for (double& w : phi) w = w / s;
    return phi;
}

inline long reflect(long q, long n) {
    const long p = 2 * n;
    q %= p;
    if (q < 0) q += p;
    return q < n ? q : p - 1 - q;
}

// One 1-D pass along the axis with stride `s` and length n for every line;
// `lines` lists the first element of each line.  Lines are independent, so
// they are shared among OpenMP threads (each with its own line buffer); the
// result does not depend on the thread count.
void pass1d(float* v, long n, long s, const std::vector<long>& lines, const std::vector<double>& w, int r) {
    const long nlines = (long)lines.size();
#pragma omp parallel
    {
        std::vector<double> buf(n + 2 * r);
#pragma omp for schedule(static)
        for (long l = 0; l < nlines; l++) {
            const long base = lines[l];
            for (long q = -r; q < n + r; q++) buf[q + r] = (double)v[base + reflect(q, n) * s];
            for (long i = 0; i < n; i++) {
                const double* x = buf.data() + i + r;
                double t = x[0] * w[r];
                for (int j = r; j >= 1; j--) t += (x[-j] + x[j]) * w[r - j];
                v[base + i * s] = (float)t;
            }
        }
    }
}

}  // namespace

void write_acquisition(const std::string& sps_path, const std::string& rps_path, const Acquisition& a) {
    write_grid(sps_path, linspace(a.sx0, a.sx1, a.nsx), linspace(a.sy0, a.sy1, a.nsy), a.sz);
    write_grid(rps_path, linspace(a.rx0, a.rx1, a.nrx), linspace(a.ry0, a.ry1, a.nry), a.rz);
}

void gaussian_filter3d(float* v, int n0, int n1, int n2, double sigma) {
    int r = 0;
    const std::vector<double> w = gaussian_weights(sigma, r);
    std::vector<long> lines;
    // axis 0 (stride 1)
    for (long k = 0; k < n2; k++)
        for (long j = 0; j < n1; j++) lines.push_back(j * n0 + k * (long)n0 * n1);
    pass1d(v, n0, 1, lines, w, r);
    // axis 1 (stride n0)
    lines.clear();
    for (long k = 0; k < n2; k++)
        for (long i = 0; i < n0; i++) lines.push_back(i + k * (long)n0 * n1);
    pass1d(v, n1, n0, lines, w, r);
    // axis 2 (stride n0*n1)
    lines.clear();
    for (long j = 0; j < n1; j++)
        for (long i = 0; i < n0; i++) lines.push_back(i + j * (long)n0);
    pass1d(v, n2, (long)n0 * n1, lines, w, r);
}

void initial_model(const float* m_true, float* m_init, int nz, int nx, int ny, float dh, float zmask, double sigma) {
    const long n = (long)nz * nx * ny;
    for (long q = 0; q < n; q++) m_init[q] = 1.0f / m_true[q];
    gaussian_filter3d(m_init, nz, nx, ny, sigma);
    for (long q = 0; q < n; q++) m_init[q] = 1.0f / m_init[q];
    const int zid = (int)(zmask / dh);
    for (long k = 0; k < ny; k++)
        for (long j = 0; j < nx; j++)
            for (int i = 0; i < zid && i < nz; i++) {
                const long q = i + j * (long)nz + k * (long)nz * nx;
                m_init[q] = m_true[q];
            }
}

void rtm_input(float* data, int nt, int nr, float dt, float sx, float sy, const float* rx, const float* ry) {
    const long n = (long)nt * nr;
    float amax = 0.0f;
    for (long q = 0; q < n; q++) amax = std::fabs(data[q]) > amax ? std::fabs(data[q]) : amax;
    const float inv = 1.0f / amax;
    for (long q = 0; q < n; q++) data[q] = data[q] * inv;
    const float velocity = 3600.0f, delay = 0.18f;
    const double gstd = 30.0;
    for (int r = 0; r < nr; r++) {
        const float dx = rx[r] - sx, dy = ry[r] - sy;
        const float distance = std::sqrt(dx * dx + dy * dy);
        const int tid = (int)((distance / velocity + delay) / dt);
        for (int t = 0; t < tid && t < nt; t++) {
            const double u = (double)(t - tid) / gstd;
            data[t + (long)r * nt] = (float)((double)data[t + (long)r * nt] * std::exp(-0.5 * (u * u)));
        }
    }
}

void read_positions(const std::string& path, std::vector<float>& x, std::vector<float>& y, std::vector<float>& z) {
    std::ifstream f(path);
    if (!f.is_open()) throw std::invalid_argument("Error: " + path + " could not be opened!");
    std::string line;
    x.clear();
    y.clear();
    z.clear();
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::stringstream ss(line);
        std::string a, b, c;
        if (!std::getline(ss, a, ',') || !std::getline(ss, b, ',') || !std::getline(ss, c, ','))
            throw std::invalid_argument("Error: bad position line in " + path + ": " + line);
        x.push_back(std::stof(a));
        y.push_back(std::stof(b));
        z.push_back(std::stof(c));
    }
}

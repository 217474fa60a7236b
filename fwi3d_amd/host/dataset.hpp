// dataset.hpp -- the reference workflow's data preparation in C++.
//
// Restates the two Python steps of the reference workflow so that a run
// needs no Python:
//   * run/generate_setup.py:24-68: the SPS/RPS acquisition files (a regular
//     shot grid and a regular receiver grid, "%.2f" comma-separated rows,
//     y varying fastest) and the initial FWI model 1 / G_sigma(1 / v_true)
//     with the layers above the depth mask kept from the true model;
//   * run/generate_input_data.py:28-46: the RTM input data, each shot
//     normalised by its largest |sample| and tapered with a Gaussian
//     (sigma = 30 samples) before the estimated direct-wave time
//     offset / 3600 m/s + 0.18 s.
// The Gaussian filter is scipy.ndimage.gaussian_filter's (mode "reflect",
// truncate 4): one 1-D pass per axis (z, then x, then y of the z-fastest
// model), each accumulated in double and stored in float.
#pragma once

#include <string>
#include <vector>

struct Acquisition {
    int nsx = 6, nsy = 6;                 // shots per axis
    double sx0 = 1000, sx1 = 4000;        // shot x range [m] (also used for y)
    double sy0 = 1000, sy1 = 4000;
    double sz = 100;                      // shot depth [m]
    int nrx = 51, nry = 51;               // receivers per axis
    double rx0 = 0, rx1 = 5000, ry0 = 0, ry1 = 5000;
    double rz = 0;                        // receiver depth [m]
};

// numpy.linspace(start, stop, num) (float64, last point exactly stop).
std::vector<double> linspace(double start, double stop, int num);

// Writes the SPS and RPS files of generate_setup.py:24-51.
void write_acquisition(const std::string& sps_path, const std::string& rps_path, const Acquisition& a);

// In-place 3-D Gaussian filter of a float volume v[i + j*n0 + k*n0*n1].
void gaussian_filter3d(float* v, int n0, int n1, int n2, double sigma);

// generate_setup.py:53-68: m_init = 1 / G_sigma(1 / m_true), rows above
// depth index (int)(zmask / dh) taken from m_true.  Volumes are nz*nx*ny,
// z fastest.
void initial_model(const float* m_true, float* m_init, int nz, int nx, int ny, float dh, float zmask, double sigma);

// generate_input_data.py:28-46 for one shot: data[t + r*nt] normalised and
// muted before the direct wave of source (sx, sy) at receivers (rx, ry).
void rtm_input(float* data, int nt, int nr, float dt, float sx, float sy, const float* rx, const float* ry);

// Reads the "x,y,z" rows of an SPS/RPS file (float32, '#' lines skipped).
void read_positions(const std::string& path, std::vector<float>& x, std::vector<float>& y, std::vector<float>& z);

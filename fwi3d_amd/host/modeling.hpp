// modeling.hpp -- acoustic modeling on the MI355X library.
//
// Same public surface as the reference's class Modeling
// (src/modeling/modeling.cuh:18-90): srcId, geometry, parameters,
// set_parameters(), initialization(), forward_solver(), get_seismogram(),
// show_information().  Propagation runs in libfwi3d_hip.so through the C ABI
// of include/fwi3d.h; this class keeps the host-side set-up (parameter
// parsing, wavelet, dampers, geometry, file output).
//
// Optional keys added by this build (ignored by the reference's parser):
//   batch_shots   = N   shots advanced together by one launch per time step
//                       (default 1: one shot after the other, as the
//                       reference's shot loop)
//   rbc_style     = blobs | ramp   random boundary of migration and
//                       inversion (default blobs: set_random_boundary's
//                       background noise and Gaussian blobs, sized by the
//                       {mig,inv}_rbc_ratio and _varVp keys; ramp: uniform
//                       noise of +-A*v growing linearly to the outer edge)
//   rbc_seed      = S   seed of the random boundary (default 20261203)
//   rbc_amplitude = A   ramp amplitude (default 0.6)
//   arithmetic    = reference | fp32   (default reference)
//   backward_corrections = comma-separated subset of
//                   gradient_dyy, subtract_source, aligned_imaging
//                       (default none: the reference's backward passes)
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fwi3d.h"
#include "geometry.hpp"

class Modeling {
protected:
    bool ABC = false;

    float abc_length = 0.0f;
    float abc_factor = 0.0f;
    float vmax = 0.0f, vmin = 0.0f, idh2 = 0.0f;
    float dh = 0.0f, dt = 0.0f, fmax = 0.0f, idh3 = 0.0f;

    std::string title;

    int nxx = 0, nyy = 0, nzz = 0;
    long volsize = 0;
    int nt = 0, nx = 0, ny = 0, nz = 0, nb = 0;
    long nPoints = 0;
    int sIdx = 0, sIdy = 0, sIdz = 0, tlag = 0;

    std::vector<float> model;       // interior vp, z fastest (nz*nx*ny)
    std::vector<float> seismogram;  // nt*nrec, t + r*nt
    std::vector<float> wavelet;
    std::vector<float> b1d;

    std::string data_folder;

    fwi_ctx* ctx = nullptr;
    int batch = 1;                  // shot slots of ctx
    int first_shot = 0, shot_count = 1;
    uint64_t rbc_seed = 20261203ull;
    float rbc_amplitude = 0.6f;
    bool rbc_blobs = true;  // rbc_style
    int arith = FWI_ARITH_REFERENCE;
    int physics = 0;  // FWI_PHYS_* bits for the backward passes
    // autotune key: 0 off (the library's default launch shapes), 1 (the
    // default) forward passes timed, backward passes on fixed one-shot
    // 64 x 16 tiles (the summation order of the default shape: results
    // equal those of 0), 2 every pass kind timed, shots per block or
    // thread included
    int autotune = 1;
    std::string autotune_cache;     // optional cache file of the choices
    unsigned tuned_kinds = 0;       // bit (kind + 4 * context): done

    void set_wavelet();
    void set_geometry();
    void set_properties();
    void set_seismograms();
    void set_abc_dampers();
    void set_b1d();  // b1d from nb and abc_factor
    void set_main_parameters();
    void set_build_parameters();

    void create_context(int boundary_cells, fwi_ctx** out);
    void check(int rc, const std::string& what) const;
    // once per context and pass kind, before its first pass (autotune key)
    void tune(fwi_ctx* c, int context, int kind);
    void grid_index(float x, float y, float z, int* ix, int* iy, int* iz) const;
    float cfl_velocity() const;
    // the random boundary of rbc_style on slot `slot` of c for shot `shot`
    void random_boundary(fwi_ctx* c, int slot, int shot, float ratio, float varVp);

public:
    int srcId = 0;
    Geometry* geometry = nullptr;
    std::string parameters;

    Modeling() = default;
    Modeling(const Modeling&) = delete;
    Modeling& operator=(const Modeling&) = delete;
    virtual ~Modeling();

    void set_parameters();
    void show_information();

    // Shots [srcId, srcId + shots_this_launch()) go to the slots of the
    // context; with batch_shots = 1 this is the reference's per-shot call.
    void initialization();
    void forward_solver();
    void get_seismogram();

    int shots_per_launch() const { return batch; }
    int shots_this_launch() const { return shot_count; }
};

// common.hpp -- shared definitions of the gfx950 wave-propagation library.
//
// Device volume layout ("x-fast"):   e(x, y, z) = x + y*px + z*pp
//   x = reference j (model_x), y = reference k (model_y), z = reference i
//   (depth); px = nxx rounded up to 64 floats (256-byte rows), pp = px*nyy.
// The step kernels tile the x-y plane and march along z, so every HBM row
// access is a 256-byte coalesced wave load.  Host volumes keep the
// reference layout (i fastest) and are transposed on upload/download.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace fwi {

// src/modeling/modeling.cuh:9-13 -- the reference's 8th-order second
// derivative coefficients, the same truncated double literals.
constexpr double FDM1 = 0.001785714285;
constexpr double FDM2 = 0.025396825396;
constexpr double FDM3 = 0.200000000000;
constexpr double FDM4 = 1.600000000000;
constexpr double FDM5 = 2.847222222222;

// Largest stable Courant number v*dt/dh of the 3-D 8th-order leapfrog
// scheme: sqrt(4 / (3 * 6.501587...)), with 6.501587 = |symbol at Nyquist|.
constexpr double CFL_MAX = 0.452856;

enum Mode { FWD_RBC = 0, FWD_ABC = 1, BWD_RTM = 2, BWD_FWI = 3 };
enum Arith { ARITH_REF = 0, ARITH_FP32 = 1 };

// Optional corrections of the backward pass (bits of StepArgs::phys; the
// default 0 is the reference's behaviour).
enum Physics {
    PHYS_GRAD_DYY = 1,         // FWI gradient with d2/dy2 (inversion.cu:466 omits it)
    PHYS_SUBTRACT_SOURCE = 2,  // remove the injected source during reconstruction
    PHYS_ALIGNED_IMAGING = 4,  // image/illumination from the source field at the receiver field's time
};

struct Layout {
    int nx, ny, nz;  // padded sizes (x = reference nxx, y = nyy, z = nzz)
    int px;          // row pitch in floats
    long pp;         // plane pitch in floats
    long vol;        // allocation in floats
    __host__ __device__ long at(int x, int y, int z) const { return (long)x + (long)y * px + (long)z * pp; }
};

inline Layout make_layout(int nxx, int nyy, int nzz) {
    Layout L;
    L.nx = nxx;
    L.ny = nyy;
    L.nz = nzz;
    L.px = (nxx + 63) / 64 * 64;
    L.pp = (long)L.px * nyy;
    L.vol = L.pp * nzz;
    return L;
}

// Launch configuration of a step kernel.
struct Tuning {
    int by = 4;       // waves (rows of 64 lanes) per block
    int rpt = 2;      // tile rows per thread (register blocking in y)
    int zchunks = 1;  // blocks along z per (x, y) tile column
    int sb = 1;       // shots per block (batched launches)
    int pd = 1;       // prefetch depth: planes of loads in flight (1 or 2)
    int big = 0;      // per-plane addressing: set for volumes of 2 GiB and more
    int tb = 1;       // time steps per launch (2: two-step launches where available)
    int spt = 1;      // shots per thread: 2 = one thread advances a cell of two shots (sb = 2, rpt = 1)
    // tile and z chunks of the two-step launches (tb = 2)
    int tby = 16, trpt = 2, tzc = 1;
};

// Arguments of one step launch.  The fields of shot s live at base + s*vol.
struct StepArgs {
    Layout L;
    long sstride;          // floats between consecutive shots' volumes
    int nshots;            // shots in this launch
    int ntx, nty, zlen, nzc;
    int nitems;            // ntx * nty * nzc * shot groups
    const float* vp;       // per shot
    // cells x, y, z in [share_nb, n - share_nb) hold the same velocity in
    // every shot of the launch (the random boundaries write only outside);
    // -1: no such guarantee
    int share_nb;
    const float* cur;      // field read by the stencil: A (forward) / So (backward)
    float* old;            // in place: previous -> next
    const float* rcur;     // backward: receiver field R
    float* rold;           // backward: in place Ro -> next
    float* image;          // RTM image or FWI gradient (one per shot group)
    float* illum;          // illumination (one per shot group)
    long istride;          // floats between shot groups' image/illum
    float idh2, dt2, dt;
    // source term: forward, added to the next step's field by the thread
    // that writes the source cell; backward with PHYS_SUBTRACT_SOURCE, added
    // back to the reconstructed field (the term the newer field carries)
    const long* src_off;   // per shot, device offset of the source cell
    float inj;             // forward: idh3 * wavelet[n+1] (0 when n+1 >= nt)
    int phys;              // backward: Physics bits
    // backward imaging of this launch: 1 = this step's terms; 0 = deferred
    // to a later launch; 2 = the previous step's terms, then this step's;
    // not with aligned imaging: 3 = image of the previous step and this one,
    // 4 = illumination of the two previous steps and this one, 5 = both
    // (other terms deferred)
    int img;
    // ABC
    const float* b1d;
    int nb;
    // forward receiver gather into a time-major row: seis[r] = cur[rec] * damp
    const long* rec_off;
    const float* rec_damp;
    float* seis_row;       // NULL: no recording this step
    const int* rec_start;  // per shot: receivers [rec_start[s], rec_start[s+1])
};

// Arguments of a two-step launch (step2_kernel.hpp): steps n, n+1 of the
// random-boundary forward pass or b, b+1 of the RTM backward pass, one shot
// per block.  Inputs (cur, old) and outputs (out_old = step-1 field,
// out_cur = step-2 field) are distinct buffers.
constexpr int kMaxRecPlanes = 4;  // receiver planes per shot a two-step launch looks up
struct Step2Args {
    Layout L;
    long sstride;          // floats between consecutive shots' volumes
    int nshots;
    int ntx, nty, zlen, nzc, nitems;
    const float* vp;
    const float* cur;      // forward: field of step n; backward: So
    const float* old;      // forward: field of step n-1; backward: S
    float* out_old;
    float* out_cur;
    const float* rcur;     // backward: R, Ro and their outputs
    const float* rold;
    float* rout_old;
    float* rout_cur;
    float* image;          // per-shot partials (istride apart)
    float* illum;
    long istride;
    float idh2, dt2, idh3;
    const long* src_off;   // forward: per shot
    float inj1, inj2;      // forward: source terms of steps n+1, n+2 (0: none)
    const long* rec_off;   // forward: gather of step n from cur
    const int* rec_start;
    float* seis_row0;      // seismogram row of step n (NULL: none)
    float* seis_row1;      // seismogram row of step n+1 (NULL: none)
    const int* rp_z;       // [nshots][kMaxRecPlanes] receiver planes (-1: none)
    const int* rp_map;     // [nshots][kMaxRecPlanes][ny*px] receiver cell index or -1
    const int* cell_start; // receivers of cell c: cell_rec[cell_start[c] .. cell_start[c+1])
    const int* cell_rec;
    const float* data_row1;  // backward: data injected before step b+1 (NULL: none)
};

// Host-side launchers (kernels_step.hip).
hipError_t launch_step(int mode, int arith, const Tuning& t, const StepArgs& a, hipStream_t st);
// two-step launches (step_inst_two.hip): available for this mode / arithmetic / tile?
bool step2_supported(int mode, int arith, const Tuning& t);
hipError_t launch_step2(int mode, int arith, const Tuning& t, const Step2Args& a, hipStream_t st);
bool step_config_supported(int mode, int arith, const Tuning& t);
// deferred imaging (StepArgs::img 0/2) available for this backward shape,
// with or without the corrections of StepArgs::phys
bool step_defer_supported(int mode, int arith, const Tuning& t, bool corrections);
// the image every other launch and the illumination every third (StepArgs::img
// 0, 3, 4, 5) available for this backward shape and these Physics bits (not
// with aligned imaging)
bool step_defer3_supported(int mode, int arith, const Tuning& t, int phys);
void step_grid(const Layout& L, const Tuning& t, int nshots, int sb, StepArgs& a, dim3& grid, dim3& block);

// Host-side launchers (kernels_aux.hip).  "host layout" buffers here are
// device buffers holding data in the reference layout (staging).
hipError_t launch_to_device(const float* host_layout, float* dev, const Layout& L, hipStream_t st);
hipError_t launch_to_host(const float* dev, float* host_layout, const Layout& L, int crop, hipStream_t st);
hipError_t launch_expand(const float* interior_host_layout, float* dev, const Layout& L, int nb, hipStream_t st);
hipError_t launch_random_boundary(float* dev, const Layout& L, int nb, uint64_t seed, int shot, float amp,
                                  float vclamp, hipStream_t st);
hipError_t launch_blob_boundary(float* dev, const Layout& L, int nb, float dh, uint64_t seed, int shot,
                                float var_vp, const float* blobs, int nblobs, float vlo, float vhi, float vclamp,
                                hipStream_t st);
hipError_t launch_add_points(float* dev, const long* off, int n, long sstride, float value, hipStream_t st);
hipError_t launch_fold_groups(float* v, long vol, int groups, hipStream_t st);
hipError_t launch_copy(const float* src, float* dst, long nfloats, hipStream_t st);
hipError_t launch_fill_noise(float* v, long n, uint64_t seed, hipStream_t st);
hipError_t launch_flush_imaging(const float* cur, const float* old, const float* rold, float* image, float* illum,
                                const Layout& L, int nshots, int sb, long istride, int ill_terms, int img_term,
                                hipStream_t st);
// every entry of map[map_n] -1, then map[pos[j]] = j for j < ncells
hipError_t launch_cell_map(int* map, long map_n, const long* pos, int ncells, hipStream_t st);
hipError_t launch_inject(float* field, const long* cell_off, const int* cell_start, const int* cell_rec,
                         int ncells, const float* data_row, float idh3, hipStream_t st);
hipError_t launch_export_image(const float* image, const float* illum, float* out, const Layout& L, int nb,
                               float idh2, int zfirst, hipStream_t st);
hipError_t launch_scale_by_inverse(float* v, long n, const double* gmax, hipStream_t st);
hipError_t launch_absmax(const float* v, long n, double* partial, double* result, hipStream_t st);
hipError_t launch_residual(const float* obs, float* cal, long n, double* partial, double* result, hipStream_t st);
hipError_t launch_stats(const float* v, const Layout& L, double* partial, double* result, hipStream_t st);
hipError_t launch_transpose2d(const float* src, float* dst, int rows, int cols, hipStream_t st);

constexpr int kReduceBlocks = 1024;

}  // namespace fwi

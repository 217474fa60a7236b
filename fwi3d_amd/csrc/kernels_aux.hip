// kernels_aux.hip -- layout transposes, model set-up, injection and the
// post-processing / FP64 reduction kernels around the time-step kernels.
#include <algorithm>

#include "common.hpp"

namespace fwi {

namespace {

constexpr int TT = 32;  // transpose tile

// Reference layout (i fastest) -> device layout (x fastest), full volume.
// One k-slab per blockIdx.z is a (nxx x nzz) matrix transposed via LDS.
__global__ void __launch_bounds__(256) to_device_kernel(const float* __restrict__ src, float* __restrict__ dev,
                                                        Layout L) {
    __shared__ float tile[TT][TT + 1];
    const int i0 = blockIdx.x * TT, j0 = blockIdx.y * TT, k = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const long slab = (long)k * L.nx * L.nz;
#pragma unroll
    for (int m = 0; m < TT; m += 8) {
        const int j = j0 + ty + m, i = i0 + tx;
        if (j < L.nx && i < L.nz) tile[ty + m][tx] = src[slab + (long)j * L.nz + i];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < TT; m += 8) {
        const int i = i0 + ty + m, j = j0 + tx;
        if (j < L.nx && i < L.nz) dev[L.at(j, k, i)] = tile[tx][ty + m];
    }
}

// Device layout -> reference layout, optionally cropping c cells per side.
__global__ void __launch_bounds__(256) to_host_kernel(const float* __restrict__ dev, float* __restrict__ dst,
                                                      Layout L, int c) {
    __shared__ float tile[TT][TT + 1];
    const int nzo = L.nz - 2 * c, nxo = L.nx - 2 * c;
    const int i0 = blockIdx.x * TT, j0 = blockIdx.y * TT, k = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
#pragma unroll
    for (int m = 0; m < TT; m += 8) {
        const int i = i0 + ty + m, j = j0 + tx;
        if (j < nxo && i < nzo) tile[ty + m][tx] = dev[L.at(j + c, k + c, i + c)];
    }
    __syncthreads();
    const long slab = (long)k * nxo * nzo;
#pragma unroll
    for (int m = 0; m < TT; m += 8) {
        const int j = j0 + ty + m, i = i0 + tx;
        if (j < nxo && i < nzo) dst[slab + (long)j * nzo + i] = tile[tx][ty + m];
    }
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Modeling::expand_boundary (modeling.cu:181-228): the reference's three
// face passes amount to replicating the nearest interior cell.
__global__ void __launch_bounds__(256) expand_kernel(const float* __restrict__ interior, float* __restrict__ dev,
                                                     Layout L, int nb) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= L.nx || y >= L.ny) return;
    const int nx = L.nx - 2 * nb, ny = L.ny - 2 * nb, nz = L.nz - 2 * nb;
    const int i = clampi(z - nb, 0, nz - 1), j = clampi(x - nb, 0, nx - 1), k = clampi(y - nb, 0, ny - 1);
    dev[L.at(x, y, z)] = interior[(long)i + (long)j * nz + (long)k * nx * nz];
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Seeded random boundary: v = v_face*(1 + amp*ramp*u), u ~ U[-1,1) from a
// counter hash of (seed, shot, reference cell index), ramp = 1 - d/nb
// (random_boundary_bg's ramp, modeling.cu:637-662), max over boundary axes.
__global__ void __launch_bounds__(256) random_boundary_kernel(float* __restrict__ dev, Layout L, int nb, uint64_t seed,
                                                              int shot, float amp, float vclamp) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= L.nx || y >= L.ny) return;
    const bool in_z = z >= nb && z < L.nz - nb, in_x = x >= nb && x < L.nx - nb, in_y = y >= nb && y < L.ny - nb;
    if (in_z && in_x && in_y) return;
    const float v = dev[L.at(clampi(x, nb, L.nx - nb - 1), clampi(y, nb, L.ny - nb - 1), clampi(z, nb, L.nz - nb - 1))];
    float ramp = 0.0f;
    if (!in_x) { const int d = x < nb ? x : L.nx - x - 1; ramp = fmaxf(ramp, 1.0f - (float)d / (float)nb); }
    if (!in_y) { const int d = y < nb ? y : L.ny - y - 1; ramp = fmaxf(ramp, 1.0f - (float)d / (float)nb); }
    if (!in_z) { const int d = z < nb ? z : L.nz - z - 1; ramp = fmaxf(ramp, 1.0f - (float)d / (float)nb); }
    const long cell = (long)z + (long)x * L.nz + (long)y * L.nx * L.nz;
    const uint64_t key = (seed * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)(uint32_t)shot * 0xD1B54A32D192ED03ull) ^
                         (uint64_t)cell;
    const int u24 = (int)(splitmix64(key) >> 40);
    const float u = (float)(u24 - 8388608) * (1.0f / 8388608.0f);
    float t = amp * ramp;
    t = t * u;
    const float f = 1.0f + t;
    float val = v * f;
    if (vclamp > 0.0f && val > vclamp) val = vclamp;
    dev[L.at(x, y, z)] = val;
}

// Reference-style random boundary (random_boundary_bg + random_boundary_gp,
// modeling.cu:558-666), deterministic and in one launch.
//  * background: each boundary cell starts from the nearest interior cell
//    (so a repeated call starts afresh, as random_boundary_bg does); for
//    each axis on which it lies in the boundary (x, then y, then z) it
//    gets v += (f * var_vp) * n with f = 1 - d/nb and n ~ N(0, 1)
//    approximated by the Irwin-Hall sum of 12 hashed uniforms minus 6;
//  * blobs: in list order, v += amp * exp(-q/2), q = |X - c|^2 / r^2 (in
//    double), for every blob with q <= 16 (4 radii; the reference sums every
//    blob everywhere, and a blob beyond 4 radii adds under amp * e^-8), and
//    v is clamped to [vlo, vhi] after each blob, in reach or not;
//  * finally v = min(v, vclamp).
// Only the nb-cell boundary is written.  random_boundary_gp tests the cell
// coordinate with X > nb*dh && X < x_max - nb*dh, which also admits the
// outermost interior layer; that layer belongs to the imaged model and
// would accumulate blobs over repeated calls, so it is left alone here.
// A block (64 x 4 cells of one z-plane) first gathers, in list order, the
// blobs whose 4-radius sphere meets its cells (wave ballots compact the
// indices into LDS), then every cell applies them in that order.
__device__ __forceinline__ double exp_det(double x) {
    // e^x = 2^k e^r, k = floor(x / ln2 + 1/2), r = x - k ln2 in
    // [-ln2/2, ln2/2]; e^r by its Taylor polynomial of degree 14 (Horner;
    // truncation below 1e-17), exact scaling by 2^k.  Only basic double
    // operations, so host and device results are identical.
    constexpr double c[15] = {1.0,
                              1.0,
                              1.0 / 2.0,
                              1.0 / 6.0,
                              1.0 / 24.0,
                              1.0 / 120.0,
                              1.0 / 720.0,
                              1.0 / 5040.0,
                              1.0 / 40320.0,
                              1.0 / 362880.0,
                              1.0 / 3628800.0,
                              1.0 / 39916800.0,
                              1.0 / 479001600.0,
                              1.0 / 6227020800.0,
                              1.0 / 87178291200.0};
    const double k = floor(x * 1.4426950408889634 + 0.5);
    const double r = x - k * 0.6931471805599453;
    double p = c[14];
#pragma unroll
    for (int n = 13; n >= 0; --n) p = c[n] + r * p;
    return ldexp(p, (int)k);
}

__device__ __forceinline__ float irwin_hall_normal(uint64_t base, long cell, int axis) {
    float s = 0.0f;
    for (int m = 0; m < 12; ++m) {
        const uint64_t h = splitmix64(base ^ ((uint64_t)cell * 36u + (uint64_t)(axis * 12 + m)));
        s = s + (float)(uint32_t)(h >> 40) * (1.0f / 16777216.0f);
    }
    return s - 6.0f;
}

__global__ void __launch_bounds__(256) blob_boundary_kernel(float* __restrict__ dev, Layout L, int nb, float dh,
                                                            uint64_t base, float var_vp,
                                                            const float* __restrict__ blobs, int nblobs, float vlo,
                                                            float vhi, float vclamp) {
    constexpr int CAP = 1024;
    __shared__ int list[CAP];
    __shared__ int wave_count[4];
    __shared__ int count;
    const int tid = threadIdx.y * 64 + threadIdx.x;
    const int wave = threadIdx.y, lane = threadIdx.x;
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 4, z = blockIdx.z;
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    const bool zb = z < nb || z >= L.nz - nb;
    // block-uniform: does any cell of the block lie in the boundary?
    const bool any = zb || x0 < nb || x0 + 63 >= L.nx - nb || y0 < nb || y0 + 3 >= L.ny - nb;
    if (!any) return;
    const bool in_x = x >= nb && x < L.nx - nb, in_y = y >= nb && y < L.ny - nb, in_z = !zb;
    const bool mine = x < L.nx && y < L.ny && !(in_x && in_y && in_z);

    float v = 0.0f;
    if (mine) {
        v = dev[L.at(clampi(x, nb, L.nx - nb - 1), clampi(y, nb, L.ny - nb - 1), clampi(z, nb, L.nz - nb - 1))];
        const long cell = (long)z + (long)x * L.nz + (long)y * L.nx * L.nz;
        if (!in_x) {
            const int d = x < nb ? x : L.nx - x - 1;
            const float f = 1.0f - (float)d / (float)nb;
            v = v + (f * var_vp) * irwin_hall_normal(base, cell, 0);
        }
        if (!in_y) {
            const int d = y < nb ? y : L.ny - y - 1;
            const float f = 1.0f - (float)d / (float)nb;
            v = v + (f * var_vp) * irwin_hall_normal(base, cell, 1);
        }
        if (!in_z) {
            const int d = z < nb ? z : L.nz - z - 1;
            const float f = 1.0f - (float)d / (float)nb;
            v = v + (f * var_vp) * irwin_hall_normal(base, cell, 2);
        }
    }
    const float X = (float)x * dh, Y = (float)y * dh, Z = (float)z * dh;
    // tile box in metres
    const double bx0 = (double)((float)x0 * dh), bx1 = (double)((float)min(x0 + 63, L.nx - 1) * dh);
    const double by0 = (double)((float)y0 * dh), by1 = (double)((float)min(y0 + 3, L.ny - 1) * dh);
    const double bz = (double)Z;

    // The clamp follows every blob, in reach or not.  Once blob 0 has passed
    // v stays inside [vlo, vhi], so a clamp where no blob is in reach is a
    // no-op except at blob 0: the first gathered list applies it up front
    // when blob 0 is not in it (also for a block that gathered nothing).
    bool started = false;
    auto apply = [&](int n) {
        if (!mine) return;
        if (!started) {
            started = true;
            if (nblobs > 0 && (n == 0 || list[0] != 0)) {
                v = v > vhi ? vhi : v;
                v = v < vlo ? vlo : v;
            }
        }
        for (int q = 0; q < n; ++q) {
            const float* b = blobs + 5 * list[q];
            const double r = (double)b[3];
            const double ax = ((double)X - (double)b[0]) / r, ay = ((double)Y - (double)b[1]) / r,
                         az = ((double)Z - (double)b[2]) / r;
            const double qq = (ax * ax + ay * ay) + az * az;
            if (qq <= 16.0) v = v + (float)((double)b[4] * exp_det(-0.5 * qq));
            if (qq <= 16.0 || list[q] == 0) {
                v = v > vhi ? vhi : v;
                v = v < vlo ? vlo : v;
            }
        }
    };

    if (tid == 0) count = 0;
    __syncthreads();
    for (int c0 = 0; c0 < nblobs; c0 += 256) {
        bool hit = false;
        const int bi = c0 + tid;
        if (bi < nblobs) {
            const float* b = blobs + 5 * bi;
            const double cx = b[0], cy = b[1], cz = b[2], reach = 4.0 * (double)b[3] * 1.0001;
            const double dx = cx < bx0 ? bx0 - cx : (cx > bx1 ? cx - bx1 : 0.0);
            const double dy = cy < by0 ? by0 - cy : (cy > by1 ? cy - by1 : 0.0);
            const double dz = cz - bz;
            hit = dx * dx + dy * dy + dz * dz <= reach * reach;
        }
        const uint64_t mask = __ballot(hit);
        if (lane == 0) wave_count[wave] = __popcll(mask);
        __syncthreads();
        int off = count;
        for (int w = 0; w < wave; ++w) off += wave_count[w];
        if (hit) list[off + __popcll(mask & ((1ull << lane) - 1ull))] = bi;
        __syncthreads();
        if (tid == 0) count += wave_count[0] + wave_count[1] + wave_count[2] + wave_count[3];
        __syncthreads();
        if (count > CAP - 256) {  // flush before the next chunk could overflow
            apply(count);
            __syncthreads();
            if (tid == 0) count = 0;
            __syncthreads();
        }
    }
    apply(count);
    if (mine) {
        if (vclamp > 0.0f && v > vclamp) v = vclamp;
        dev[L.at(x, y, z)] = v;
    }
}

// Source injection of the first step for every shot of a launch
// (compute_pressure's atomicAdd, modeling.cu:347-348): shot s's field starts
// at s*sstride and its source cell is off[s].
__global__ void add_points_kernel(float* dev, const long* __restrict__ off, int n, long sstride, float value) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n && off[s] >= 0) {  // a negative offset: the slot has no source
        float* p = dev + (long)s * sstride + off[s];
        *p = *p + value;
    }
}

// Fold shot-group partial accumulators into group 0 (in group order) and
// clear the others: image/illumination of a batched backward pass.
__global__ void __launch_bounds__(256) fold_groups_kernel(float* __restrict__ v, long vol, int groups) {
    for (long q = blockIdx.x * 256L + threadIdx.x; q < vol; q += (long)gridDim.x * 256) {
        float acc = v[q];
        for (int g = 1; g < groups; ++g) {
            acc = acc + v[(long)g * vol + q];
            v[(long)g * vol + q] = 0.0f;
        }
        v[q] = acc;
    }
}

// Deferred backward imaging terms still pending when a pass or a segment
// ends (the step kernels update the illumination every third launch and the
// image every other one).  cur, old and rold are the fields the next launch
// would read: the newest reconstructed source field (c), the one before it
// (o) and the receiver field the last launch read (ro).  Pending terms are
// added in step order, each step's in shot order within the shot group:
// illum += o*o (two pending), illum += c*c (one or two), image += c*ro
// (RTM).  Updated cells only (4 <= x, y, z < n-4); shot s's fields at
// s*vol; group grp (shots grp*sb .. grp*sb+sb-1) accumulates at grp*istride.
__global__ void __launch_bounds__(256) flush_imaging_kernel(const float* __restrict__ cur,
                                                            const float* __restrict__ old,
                                                            const float* __restrict__ rold, float* __restrict__ image,
                                                            float* __restrict__ illum, Layout L, int sb, long istride,
                                                            int ill_terms, int img_term) {
    const int n = (L.ny - 8) * L.px;
    for (int z = 4 + blockIdx.y; z < L.nz - 4; z += gridDim.y) {
        const long fb = (long)blockIdx.z * sb * L.vol + (long)z * L.pp;
        const long ib = (long)blockIdx.z * istride + (long)z * L.pp;
        for (int q = blockIdx.x * 256 + threadIdx.x; q < n; q += gridDim.x * 256) {
            const int y = 4 + q / L.px;
            const int x = q - (y - 4) * L.px;
            if (x < 4 || x >= L.nx - 4) continue;
            const long e = (long)y * L.px + x;
            if (ill_terms > 0) {
                float il = illum[ib + e];
                if (ill_terms > 1)
                    for (int g = 0; g < sb; ++g) {
                        const float o = old[fb + g * L.vol + e];
                        il = il + o * o;
                    }
                for (int g = 0; g < sb; ++g) {
                    const float c = cur[fb + g * L.vol + e];
                    il = il + c * c;
                }
                illum[ib + e] = il;
            }
            if (img_term) {
                float im = image[ib + e];
                for (int g = 0; g < sb; ++g) im = im + cur[fb + g * L.vol + e] * rold[fb + g * L.vol + e];
                image[ib + e] = im;
            }
        }
    }
}

// inject_seismogram / inject_residuo (migration.cu:119-132,
// inversion.cu:385-398): field[cell] += idh3 * data[r] for every receiver r
// on the cell, summed in receiver order (deterministic, unlike atomics).
__global__ void __launch_bounds__(256) inject_kernel(float* __restrict__ field, const long* __restrict__ cell_off,
                                                     const int* __restrict__ cell_start,
                                                     const int* __restrict__ cell_rec, int ncells,
                                                     const float* __restrict__ data_row, float idh3) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const long off = cell_off[c];
    float v = field[off];
    for (int q = cell_start[c]; q < cell_start[c + 1]; ++q) v = v + idh3 * data_row[cell_rec[q]];
    field[off] = v;
}

// Receiver-plane maps of the two-step launches: every entry -1, then entry
// pos[j] = j for each receiver cell j.
__global__ void __launch_bounds__(256) fill_i32_kernel(int* __restrict__ v, long n, int value) {
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) v[q] = value;
}

__global__ void __launch_bounds__(256) cell_map_kernel(int* __restrict__ map, const long* __restrict__ pos,
                                                       int ncells) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < ncells) map[pos[j]] = j;
}

// Migration::export_seismic (migration.cu:90-117) into the reference
// interior layout: I = image/illum, out = -(I[i-1] - 2I[i] + I[i+1])*idh2.
__global__ void __launch_bounds__(256) export_image_kernel(const float* __restrict__ image,
                                                           const float* __restrict__ illum, float* __restrict__ out,
                                                           Layout L, int nb, float idh2, int zfirst) {
    const int nz = L.nz - 2 * nb, nx = L.nx - 2 * nb, ny = L.ny - 2 * nb;
    const long n = (long)nx * ny * nz;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int k = (int)(q / ((long)nx * nz));
        const int j = (int)((q - (long)k * nx * nz) / nz);
        const int i = (int)(q - (long)j * nz - (long)k * nx * nz);
        float v = 0.0f;
        if (i > zfirst && i < nz - 1) {
            const long c = L.at(j + nb, k + nb, i + nb);
            const float a = image[c - L.pp] / illum[c - L.pp];
            const float b = image[c] / illum[c];
            const float d = image[c + L.pp] / illum[c + L.pp];
            v = -1.0f * (a - 2.0f * b + d) * idh2;
        }
        out[q] = v;
    }
}

__global__ void __launch_bounds__(256) scale_kernel(float* __restrict__ v, long n, const double* __restrict__ gmax) {
    const float s = 1.0f / (float)gmax[0];
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        v[q] = v[q] * s;
}

// ---- FP64 two-pass reductions (fixed order: deterministic) ---------------

template <int OP>
__device__ __forceinline__ double combine(double a, double b) {
    if constexpr (OP == 0) return a + b;        // sum
    else if constexpr (OP == 1) return fmax(a, b);  // max
    else return fmin(a, b);                      // min
}

template <int OP>
__device__ double block_reduce(double v, double* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = combine<OP>(sh[t], sh[t + s]);
        __syncthreads();
    }
    return sh[0];
}

// |v| max over a contiguous array (gradient normalisation, inversion.cu:152-160).
__global__ void __launch_bounds__(256) absmax_partial(const float* __restrict__ v, long n, double* partial) {
    __shared__ double sh[256];
    double m = 0.0;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) m = fmax(m, (double)fabsf(v[q]));
    const double r = block_reduce<1>(m, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// residual r = obs - cal in place over cal, and per-block sum of (double)r^2
// (Inversion::set_adjoint_source, inversion.cu:190-201).
__global__ void __launch_bounds__(256) residual_partial(const float* __restrict__ obs, float* __restrict__ cal, long n,
                                                        double* partial) {
    __shared__ double sh[256];
    double s = 0.0;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        const float r = obs[q] - cal[q];
        cal[q] = r;
        s += (double)r * (double)r;
    }
    const double r = block_reduce<0>(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// max, min and non-finite count over the cells of a padded device volume.
__global__ void __launch_bounds__(256) stats_partial(const float* __restrict__ v, Layout L, double* partial) {
    __shared__ double sh[256];
    double mx = -INFINITY, mn = INFINITY, bad = 0.0;
    const long n = (long)L.nx * L.ny * L.nz;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        const int x = (int)(q % L.nx);
        const long r = q / L.nx;
        const int y = (int)(r % L.ny);
        const int z = (int)(r / L.ny);
        const float f = v[L.at(x, y, z)];
        if (!isfinite(f)) {
            bad += 1.0;
        } else {
            mx = fmax(mx, (double)f);
            mn = fmin(mn, (double)f);
        }
    }
    const double a = block_reduce<1>(mx, sh);
    __syncthreads();
    const double b = block_reduce<2>(mn, sh);
    __syncthreads();
    const double c = block_reduce<0>(bad, sh);
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x + 0] = a;
        partial[3 * blockIdx.x + 1] = b;
        partial[3 * blockIdx.x + 2] = c;
    }
}

// Second pass: one block folds `nparts` partials (`nval` values each, op
// codes 0 sum / 1 max / 2 min) in a fixed order.
__global__ void __launch_bounds__(256) finish_kernel(const double* __restrict__ partial, int nparts, int nval,
                                                     int op0, int op1, int op2, double* result) {
    __shared__ double sh[256];
    for (int v = 0; v < nval; ++v) {
        const int op = v == 0 ? op0 : (v == 1 ? op1 : op2);
        double acc = op == 0 ? 0.0 : (op == 1 ? -INFINITY : INFINITY);
        for (int q = threadIdx.x; q < nparts; q += 256) {
            const double p = partial[(long)q * nval + v];
            acc = op == 0 ? acc + p : (op == 1 ? fmax(acc, p) : fmin(acc, p));
        }
        double r;
        if (op == 0) r = block_reduce<0>(acc, sh);
        else if (op == 1) r = block_reduce<1>(acc, sh);
        else r = block_reduce<2>(acc, sh);
        if (threadIdx.x == 0) result[v] = r;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) transpose2d_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                          int rows, int cols) {
    __shared__ float tile[TT][TT + 1];
    const int c0 = blockIdx.x * TT, r0 = blockIdx.y * TT;
    const int tx = threadIdx.x, ty = threadIdx.y;
#pragma unroll
    for (int m = 0; m < TT; m += 8) {
        const int r = r0 + ty + m, c = c0 + tx;
        if (r < rows && c < cols) tile[ty + m][tx] = src[(long)r * cols + c];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < TT; m += 8) {
        const int c = c0 + ty + m, r = r0 + tx;
        if (r < rows && c < cols) dst[(long)c * rows + r] = tile[tx][ty + m];
    }
}

}  // namespace

hipError_t launch_to_device(const float* src, float* dev, const Layout& L, hipStream_t st) {
    dim3 grid((L.nz + TT - 1) / TT, (L.nx + TT - 1) / TT, L.ny);
    to_device_kernel<<<grid, dim3(TT, 8), 0, st>>>(src, dev, L);
    return hipGetLastError();
}

hipError_t launch_to_host(const float* dev, float* dst, const Layout& L, int crop, hipStream_t st) {
    dim3 grid((L.nz - 2 * crop + TT - 1) / TT, (L.nx - 2 * crop + TT - 1) / TT, L.ny - 2 * crop);
    to_host_kernel<<<grid, dim3(TT, 8), 0, st>>>(dev, dst, L, crop);
    return hipGetLastError();
}

hipError_t launch_expand(const float* interior, float* dev, const Layout& L, int nb, hipStream_t st) {
    dim3 grid((L.nx + 63) / 64, (L.ny + 3) / 4, L.nz);
    expand_kernel<<<grid, dim3(64, 4), 0, st>>>(interior, dev, L, nb);
    return hipGetLastError();
}

hipError_t launch_random_boundary(float* dev, const Layout& L, int nb, uint64_t seed, int shot, float amp,
                                  float vclamp, hipStream_t st) {
    dim3 grid((L.nx + 63) / 64, (L.ny + 3) / 4, L.nz);
    random_boundary_kernel<<<grid, dim3(64, 4), 0, st>>>(dev, L, nb, seed, shot, amp, vclamp);
    return hipGetLastError();
}

hipError_t launch_blob_boundary(float* dev, const Layout& L, int nb, float dh, uint64_t seed, int shot,
                                float var_vp, const float* blobs, int nblobs, float vlo, float vhi, float vclamp,
                                hipStream_t st) {
    const uint64_t base = (seed * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)(uint32_t)shot * 0xD1B54A32D192ED03ull) ^
                          0x5851F42D4C957F2Dull;
    dim3 grid((L.nx + 63) / 64, (L.ny + 3) / 4, L.nz);
    blob_boundary_kernel<<<grid, dim3(64, 4), 0, st>>>(dev, L, nb, dh, base, var_vp, blobs, nblobs, vlo, vhi, vclamp);
    return hipGetLastError();
}

hipError_t launch_add_points(float* dev, const long* off, int n, long sstride, float value, hipStream_t st) {
    add_points_kernel<<<(n + 63) / 64, 64, 0, st>>>(dev, off, n, sstride, value);
    return hipGetLastError();
}

// Plain streaming copy: the calibration kernel whose bandwidth and HBM
// counters the stencil kernels are compared with.  Each thread moves four
// 16-byte vectors (all four loads issued before the stores), a block moves
// 16 KiB, and the grid covers the buffer once.
typedef float copy_vec __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) copy_kernel(const copy_vec* __restrict__ src, copy_vec* __restrict__ dst,
                                                   long n) {
    const long base = blockIdx.x * 1024L + threadIdx.x;
    copy_vec v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (base + u * 256 < n) v[u] = __builtin_nontemporal_load(src + base + u * 256);
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (base + u * 256 < n) __builtin_nontemporal_store(v[u], dst + base + u * 256);
}

// Trial data of the autotuner: hashed values uniform in [-1, 1), so that
// the trials run on fields whose bits change the way a wavefield's do (the
// step kernels draw less power, and run faster, on zero fields).
__global__ void __launch_bounds__(256) fill_noise_kernel(float* __restrict__ v, long n, uint64_t seed) {
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const uint64_t h = splitmix64(seed ^ ((uint64_t)i * 0x9E3779B97F4A7C15ull));
        v[i] = (float)(int)(h >> 40) * (1.0f / 8388608.0f) - 1.0f;
    }
}

hipError_t launch_fill_noise(float* v, long n, uint64_t seed, hipStream_t st) {
    fill_noise_kernel<<<4096, 256, 0, st>>>(v, n, seed);
    return hipGetLastError();
}

hipError_t launch_copy(const float* src, float* dst, long nfloats, hipStream_t st) {
    const long n = nfloats / 4;
    copy_kernel<<<(unsigned)((n + 1023) / 1024), 256, 0, st>>>((const copy_vec*)src, (copy_vec*)dst, n);
    return hipGetLastError();
}

hipError_t launch_fold_groups(float* v, long vol, int groups, hipStream_t st) {
    if (groups <= 1) return hipSuccess;
    fold_groups_kernel<<<4096, 256, 0, st>>>(v, vol, groups);
    return hipGetLastError();
}

hipError_t launch_flush_imaging(const float* cur, const float* old, const float* rold, float* image, float* illum,
                                const Layout& L, int nshots, int sb, long istride, int ill_terms, int img_term,
                                hipStream_t st) {
    if (ill_terms < 0 || ill_terms > 2 || nshots <= 0 || sb <= 0 || nshots % sb != 0 || L.nz <= 8 || L.ny <= 8)
        return hipErrorInvalidValue;
    if (ill_terms == 0 && !img_term) return hipSuccess;
    const int n = (L.ny - 8) * L.px;
    const dim3 grid((unsigned)std::min((n + 255) / 256, 64), (unsigned)std::min(L.nz - 8, 4096),
                    (unsigned)(nshots / sb));
    flush_imaging_kernel<<<grid, 256, 0, st>>>(cur, old, rold, image, illum, L, sb, istride, ill_terms, img_term);
    return hipGetLastError();
}

hipError_t launch_cell_map(int* map, long map_n, const long* pos, int ncells, hipStream_t st) {
    fill_i32_kernel<<<2048, 256, 0, st>>>(map, map_n, -1);
    if (ncells > 0) cell_map_kernel<<<(ncells + 255) / 256, 256, 0, st>>>(map, pos, ncells);
    return hipGetLastError();
}

hipError_t launch_inject(float* field, const long* cell_off, const int* cell_start, const int* cell_rec, int ncells,
                         const float* data_row, float idh3, hipStream_t st) {
    if (ncells <= 0) return hipSuccess;
    inject_kernel<<<(ncells + 255) / 256, 256, 0, st>>>(field, cell_off, cell_start, cell_rec, ncells, data_row,
                                                          idh3);
    return hipGetLastError();
}

hipError_t launch_export_image(const float* image, const float* illum, float* out, const Layout& L, int nb,
                               float idh2, int zfirst, hipStream_t st) {
    export_image_kernel<<<2048, 256, 0, st>>>(image, illum, out, L, nb, idh2, zfirst);
    return hipGetLastError();
}

hipError_t launch_scale_by_inverse(float* v, long n, const double* gmax, hipStream_t st) {
    scale_kernel<<<2048, 256, 0, st>>>(v, n, gmax);
    return hipGetLastError();
}

hipError_t launch_absmax(const float* v, long n, double* partial, double* result, hipStream_t st) {
    absmax_partial<<<kReduceBlocks, 256, 0, st>>>(v, n, partial);
    finish_kernel<<<1, 256, 0, st>>>(partial, kReduceBlocks, 1, 1, 0, 0, result);
    return hipGetLastError();
}

hipError_t launch_residual(const float* obs, float* cal, long n, double* partial, double* result, hipStream_t st) {
    residual_partial<<<kReduceBlocks, 256, 0, st>>>(obs, cal, n, partial);
    finish_kernel<<<1, 256, 0, st>>>(partial, kReduceBlocks, 1, 0, 0, 0, result);
    return hipGetLastError();
}

hipError_t launch_stats(const float* v, const Layout& L, double* partial, double* result, hipStream_t st) {
    stats_partial<<<kReduceBlocks, 256, 0, st>>>(v, L, partial);
    finish_kernel<<<1, 256, 0, st>>>(partial, kReduceBlocks, 3, 1, 2, 0, result);
    return hipGetLastError();
}

hipError_t launch_transpose2d(const float* src, float* dst, int rows, int cols, hipStream_t st) {
    if (rows <= 0 || cols <= 0) return hipSuccess;
    dim3 grid((cols + TT - 1) / TT, (rows + TT - 1) / TT);
    transpose2d_kernel<<<grid, dim3(TT, 8), 0, st>>>(src, dst, rows, cols);
    return hipGetLastError();
}

}  // namespace fwi

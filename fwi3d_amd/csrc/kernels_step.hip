// kernels_step.hip -- the time-step kernels of the wave-propagation path.
//
// One launch advances every cell 4 <= idx < n-4 of one or more shots by one
// leapfrog step of the 8th-order acoustic scheme:
//   forward  (compute_pressure, src/modeling/modeling.cu:324-382)
//       new = dt^2 vp^2 L(cur) + 2 cur - old            [+ Cerjan damping]
//   backward (cross_correlation, src/migration/migration.cu:181-202, and
//             build_gradient, src/inversion/inversion.cu:491-558)
//       S   = dt^2 vp^2 L(So) + 2 So - S   (time-reversed source field)
//       Ro  = dt^2 vp^2 L(R)  + 2 R  - Ro  (receiver field)
//       illum += S^2;  image += S R   |  grad += dt R (Dxx + Dzz)(So) vp^2
//
// Work decomposition (gfx950, wave64):
//  * a block owns a 64 x TY tile of the x-y plane (TY = BY*RPT rows; each of
//    the BY waves covers one 64-float row, 256 B, per register row) and
//    marches along z over one z-chunk;
//  * each thread keeps the 9 z-samples z-4..z+4 of its column in a register
//    queue (Dzz needs no LDS and no re-read);
//  * the centre plane z is staged with a 4-cell halo in LDS, double
//    buffered, so one barrier per plane suffices; Dxx and Dyy read LDS;
//  * HBM traffic per cell and step: read cur, old, vp; write new (16 B).
//    Halo rows are re-read by neighbouring tiles, which the XCD-aware item
//    order keeps on the same XCD so they hit its L2.
//  * old and new share one buffer: each cell reads its own old value and
//    overwrites it, so no second field write is needed.  The ABC damping of
//    the current field (the reference's P *= damper, which races with the
//    neighbours' reads) is applied instead when the cell is next read as
//    "old", giving the same values in a single pass.
//  * SB shots per block (batched mode) share the item's image/illumination
//    read-modify-write.
//
// Arithmetic.  ARITH_REF reproduces the reference kernel literally: float
// pair sums, double coefficients and accumulation per axis, rounding to
// float per axis, then the float update in source order; the library is
// built with -ffp-contract=off so each operation rounds once, which makes
// the result bit-identical to the CPU oracle.  ARITH_FP32 evaluates the
// derivative with FP32 FMAs.
#include "common.hpp"

namespace fwi {

namespace {

__device__ __forceinline__ float d2_ref(float c, float m1, float p1, float m2, float p2, float m3, float p3, float m4,
                                        float p4, float idh2) {
    double d = -FDM1 * (double)(m4 + p4);
    d = d + FDM2 * (double)(m3 + p3);
    d = d - FDM3 * (double)(m2 + p2);
    d = d + FDM4 * (double)(m1 + p1);
    d = d - FDM5 * (double)c;
    d = d * (double)idh2;
    return (float)d;
}

__device__ __forceinline__ float d2_f32(float c, float m1, float p1, float m2, float p2, float m3, float p3, float m4,
                                        float p4, float idh2) {
    const float c1 = (float)FDM1, c2 = (float)FDM2, c3 = (float)FDM3, c4 = (float)FDM4, c5 = (float)FDM5;
    float d = -c5 * c;
    d = __builtin_fmaf(c4, m1 + p1, d);
    d = __builtin_fmaf(-c3, m2 + p2, d);
    d = __builtin_fmaf(c2, m3 + p3, d);
    d = __builtin_fmaf(-c1, m4 + p4, d);
    return d * idh2;
}

template <int ARITH>
__device__ __forceinline__ float d2(float c, float m1, float p1, float m2, float p2, float m3, float p3, float m4,
                                    float p4, float idh2) {
    if constexpr (ARITH == ARITH_REF)
        return d2_ref(c, m1, p1, m2, p2, m3, p3, m4, p4, idh2);
    else
        return d2_f32(c, m1, p1, m2, p2, m3, p3, m4, p4, idh2);
}

// get_boundary_damper (modeling.cu:395-520) from the 1-D table only: the
// reference's b2d/b3d entries are fl(b1[a]+b1[b]) - 1 and
// ((fl(b1[i]+b1[j]) + fl(b1[j]+b1[k])) + fl(b1[i]+b1[k])) - 5 (modeling.cu:
// 101-127), recomputed here with the same float operations.  The lookup is
// split per axis: the x and y parts are fixed while a thread marches along
// z, and the z part is uniform over a block.
struct DampAxis {
    float b;  // b1d[d] when the cell is within nb of a face on this axis
    bool on;
};

__device__ __forceinline__ DampAxis damp_axis(const float* __restrict__ b1d, int nb, int q, int n) {
    const int d = q < nb ? q : (q >= n - nb ? n - 1 - q : -1);
    return {d >= 0 ? b1d[d] : 0.0f, d >= 0};
}

__device__ __forceinline__ float damper(DampAxis i, DampAxis j, DampAxis k) {
    const int cnt = (int)i.on + (int)j.on + (int)k.on;
    if (cnt == 0) return 1.0f;
    if (cnt == 1) return i.on ? i.b : (j.on ? j.b : k.b);
    if (cnt == 2) {
        const float a = i.on ? i.b : j.b;
        const float b = k.on ? k.b : j.b;
        return (a + b) - 1.0f;
    }
    const float t1 = i.b + j.b, t2 = j.b + k.b, t3 = i.b + k.b;
    return ((t1 + t2) + t3) - 5.0f;
}

template <int MODE, int ARITH, int BY, int RPT, int SB>
__global__ void __launch_bounds__(64 * BY) step_kernel(const StepArgs a) {
    constexpr int NF = (MODE == BWD_RTM || MODE == BWD_FWI) ? 2 : 1;
    constexpr bool BWD = NF == 2;
    constexpr bool ABC = MODE == FWD_ABC;
    constexpr int TY = BY * RPT;
    constexpr int LW = 64 + 8;
    constexpr int LH = TY + 8;
    constexpr int HY = (8 + BY - 1) / BY;  // y-halo rows loaded per wave
    __shared__ float lds[2][SB * NF][LH][LW];

    const Layout& L = a.L;
    const int tx = threadIdx.x, ty = threadIdx.y;

    // Forward receiver gather of the current field (compute_seismogram,
    // modeling.cu:384-393).  The current field is read-only in this launch
    // (new values go to the old buffer), so any block may read it.
    if constexpr (!BWD) {
        if (a.seis_row) {
            const int nrec = a.rec_start[a.nshots];
            const int per = (nrec + gridDim.x - 1) / gridDim.x;
            const int beg = blockIdx.x * per;
            const int end = min(beg + per, nrec);
            for (int q = beg + ty * 64 + tx; q < end; q += 64 * BY) {
                float v = a.cur[a.rec_off[q]];
                if (a.rec_damp) v = v * a.rec_damp[q];
                a.seis_row[q] = v;
            }
        }
    }

    // XCD-aware item order: workgroups are dispatched round-robin over the
    // 8 XCDs, so blocks b, b+8, b+16, ... share an L2.  Give each XCD a
    // contiguous range of items (adjacent tiles share halo rows).
    const int nblk = gridDim.x;
    const int xcd = blockIdx.x & 7;
    const int slot = blockIdx.x >> 3;
    const int per_xcd_blocks = nblk >> 3;
    const int range = (a.nitems + 7) >> 3;
    const int item_end = min((xcd + 1) * range, a.nitems);

    for (int item = xcd * range + slot; item < item_end; item += per_xcd_blocks) {
        const int tile_x = item % a.ntx;
        const int r1 = item / a.ntx;
        const int tile_y = r1 % a.nty;
        const int r2 = r1 / a.nty;
        const int cz = r2 % a.nzc;
        const int grp = r2 / a.nzc;

        const int x = tile_x * 64 + tx;
        const int y0 = 4 + tile_y * TY;
        const int zb = 4 + cz * a.zlen;
        const int ze = min(zb + a.zlen, L.nz - 4);
        if (zb >= ze) continue;  // uniform per block
        const bool xin = x >= 4 && x < L.nx - 4;
        const bool xld = x < L.nx;
        // x-halo lanes: 0..3 load x-4 (left), 60..63 load x+4 (right)
        const int hxx = tx < 4 ? x - 4 : x + 4;
        const int hxc = tx < 4 ? tx : tx + 8;
        const bool hx_lane = tx < 4 || tx >= 60;
        const bool hxok = hx_lane && hxx >= 0 && hxx < L.nx;

        int yr[RPT];
        bool act[RPT], yld[RPT];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            yr[r] = y0 + ty + r * BY;
            act[r] = xin && yr[r] < L.ny - 4;
            yld[r] = xld && yr[r] < L.ny;
        }
        int hy_y[HY], hy_row[HY];
        bool hy_ok[HY];
#pragma unroll
        for (int m = 0; m < HY; ++m) {
            const int h = ty + m * BY;
            hy_y[m] = h < 4 ? y0 - 4 + h : y0 + TY + (h - 4);
            hy_row[m] = h < 4 ? h : TY + h;
            hy_ok[m] = h < 8 && xld && hy_y[m] < L.ny;
        }

        DampAxis dmp_x{0.0f, false}, dmp_y[RPT];
        if constexpr (ABC) {
            dmp_x = damp_axis(a.b1d, a.nb, x, L.nx);
#pragma unroll
            for (int r = 0; r < RPT; ++r) dmp_y[r] = damp_axis(a.b1d, a.nb, yr[r], L.ny);
        }

        const float* fld[SB][NF];
        long sbase[SB];
        long src_off[SB];
#pragma unroll
        for (int s = 0; s < SB; ++s) {
            const int shot = grp * SB + s;
            sbase[s] = (long)shot * a.sstride;
            fld[s][0] = a.cur + sbase[s];
            if constexpr (BWD) fld[s][1] = a.rcur + sbase[s];
            src_off[s] = BWD ? -1 : a.src_off[shot];
        }
        const long gbase = (long)grp * a.istride;

        // Pipeline registers: q = z-samples z-4..z+4 of the stencil fields;
        // hx/hy = halo values of the next plane; o/v/ro/im/il = per-cell
        // streams of the plane being computed, *_n = of the next plane.
        float q[SB][NF][RPT][9];
        float qn[SB][NF][RPT];
        float hx[SB][NF][RPT];
        float hy[SB][NF][HY];
        float o[SB][RPT], v[SB][RPT], ro[SB][RPT], im[RPT], il[RPT];
        float o_n[SB][RPT], v_n[SB][RPT], ro_n[SB][RPT], im_n[RPT], il_n[RPT];

        auto load_plane = [&](int zz, float (&oo)[SB][RPT], float (&vv)[SB][RPT], float (&rr)[SB][RPT],
                              float (&ii)[RPT], float (&ll)[RPT]) {
            // halos of plane zz and the streamed per-cell values of plane zz
#pragma unroll
            for (int s = 0; s < SB; ++s)
#pragma unroll
                for (int f = 0; f < NF; ++f) {
#pragma unroll
                    for (int r = 0; r < RPT; ++r)
                        hx[s][f][r] = (hxok && yr[r] < L.ny) ? fld[s][f][L.at(hxx, yr[r], zz)] : 0.0f;
#pragma unroll
                    for (int m = 0; m < HY; ++m) hy[s][f][m] = hy_ok[m] ? fld[s][f][L.at(x, hy_y[m], zz)] : 0.0f;
                }
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const long off = L.at(x, yr[r], zz);
#pragma unroll
                for (int s = 0; s < SB; ++s) {
                    oo[s][r] = act[r] ? __builtin_nontemporal_load(a.old + sbase[s] + off) : 0.0f;
                    vv[s][r] = act[r] ? __builtin_nontemporal_load(a.vp + sbase[s] + off) : 0.0f;
                    if constexpr (BWD) rr[s][r] = act[r] ? __builtin_nontemporal_load(a.rold + sbase[s] + off) : 0.0f;
                }
                if constexpr (BWD) {
                    ii[r] = act[r] ? __builtin_nontemporal_load(a.image + gbase + off) : 0.0f;
                    ll[r] = act[r] ? __builtin_nontemporal_load(a.illum + gbase + off) : 0.0f;
                }
            }
        };
        auto stage_plane = [&](int buf, int qi) {
            // centre values (queue slot qi) and the halos into LDS buffer buf
#pragma unroll
            for (int s = 0; s < SB; ++s)
#pragma unroll
                for (int f = 0; f < NF; ++f) {
#pragma unroll
                    for (int r = 0; r < RPT; ++r) {
                        lds[buf][s * NF + f][ty + r * BY + 4][tx + 4] = q[s][f][r][qi];
                        if (hx_lane) lds[buf][s * NF + f][ty + r * BY + 4][hxc] = hx[s][f][r];
                    }
#pragma unroll
                    for (int m = 0; m < HY; ++m)
                        if (ty + m * BY < 8) lds[buf][s * NF + f][hy_row[m]][tx + 4] = hy[s][f][m];
                }
        };

        // prologue: planes zb-4 .. zb+4 into the queue, plane zb into LDS
#pragma unroll
        for (int s = 0; s < SB; ++s)
#pragma unroll
            for (int f = 0; f < NF; ++f)
#pragma unroll
                for (int r = 0; r < RPT; ++r)
#pragma unroll
                    for (int d = 0; d < 9; ++d)
                        q[s][f][r][d] = yld[r] ? fld[s][f][L.at(x, yr[r], zb - 4 + d)] : 0.0f;
        load_plane(zb, o, v, ro, im, il);
        stage_plane(0, 4);
        __syncthreads();

        for (int z = zb; z < ze; ++z) {
            const int buf = (z - zb) & 1;
            const bool more = z + 1 < ze;  // uniform per block
            // 1. prefetch plane z+1 (queue head z+5, halos, streams)
            if (more) {
#pragma unroll
                for (int s = 0; s < SB; ++s)
#pragma unroll
                    for (int f = 0; f < NF; ++f)
#pragma unroll
                        for (int r = 0; r < RPT; ++r)
                            qn[s][f][r] = yld[r] ? fld[s][f][L.at(x, yr[r], z + 5)] : 0.0f;
                load_plane(z + 1, o_n, v_n, ro_n, im_n, il_n);
            }

            // 2. compute plane z
            DampAxis dmp_z{0.0f, false};
            if constexpr (ABC) dmp_z = damp_axis(a.b1d, a.nb, z, L.nz);
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                if (!act[r]) continue;
                const long off = L.at(x, yr[r], z);
                const int lr = ty + r * BY + 4;
                const int lc = tx + 4;
                float ill = il[r], img = im[r];
#pragma unroll
                for (int s = 0; s < SB; ++s) {
                    const float(*P)[LW] = lds[buf][s * NF];
                    const float c = q[s][0][r][4];
                    const float dxx = d2<ARITH>(c, P[lr][lc - 1], P[lr][lc + 1], P[lr][lc - 2], P[lr][lc + 2],
                                                P[lr][lc - 3], P[lr][lc + 3], P[lr][lc - 4], P[lr][lc + 4], a.idh2);
                    const float dyy = d2<ARITH>(c, P[lr - 1][lc], P[lr + 1][lc], P[lr - 2][lc], P[lr + 2][lc],
                                                P[lr - 3][lc], P[lr + 3][lc], P[lr - 4][lc], P[lr + 4][lc], a.idh2);
                    const float* qz = q[s][0][r];
                    const float dzz = d2<ARITH>(c, qz[3], qz[5], qz[2], qz[6], qz[1], qz[7], qz[0], qz[8], a.idh2);
                    const float vel = v[s][r];
                    const float vp2 = vel * vel;
                    const float K = a.dt2 * vp2;
                    if constexpr (!BWD) {
                        float oo = o[s][r];
                        float dmp = 1.0f;
                        if constexpr (ABC) {
                            dmp = damper(dmp_z, dmp_x, dmp_y[r]);
                            oo = oo * dmp;
                        }
                        float nv = K * ((dxx + dyy) + dzz);
                        nv = nv + 2.0f * c;
                        nv = nv - oo;
                        if constexpr (ABC) nv = nv * dmp;
                        if (off == src_off[s]) nv = nv + a.inj;
                        __builtin_nontemporal_store(nv, a.old + sbase[s] + off);
                    } else {
                        const float(*R)[LW] = lds[buf][s * NF + 1];
                        const float cr = q[s][1][r][4];
                        const float rxx = d2<ARITH>(cr, R[lr][lc - 1], R[lr][lc + 1], R[lr][lc - 2], R[lr][lc + 2],
                                                    R[lr][lc - 3], R[lr][lc + 3], R[lr][lc - 4], R[lr][lc + 4],
                                                    a.idh2);
                        const float ryy = d2<ARITH>(cr, R[lr - 1][lc], R[lr + 1][lc], R[lr - 2][lc], R[lr + 2][lc],
                                                    R[lr - 3][lc], R[lr + 3][lc], R[lr - 4][lc], R[lr + 4][lc],
                                                    a.idh2);
                        const float* rz = q[s][1][r];
                        const float rzz =
                            d2<ARITH>(cr, rz[3], rz[5], rz[2], rz[6], rz[1], rz[7], rz[0], rz[8], a.idh2);
                        float sn = K * ((dxx + dyy) + dzz);
                        sn = sn + 2.0f * c;
                        sn = sn - o[s][r];
                        float rn = K * ((rxx + ryy) + rzz);
                        rn = rn + 2.0f * cr;
                        rn = rn - ro[s][r];
                        __builtin_nontemporal_store(sn, a.old + sbase[s] + off);
                        __builtin_nontemporal_store(rn, a.rold + sbase[s] + off);
                        ill = ill + sn * sn;
                        if constexpr (MODE == BWD_RTM) {
                            img = img + sn * cr;
                        } else {
                            float gterm = a.dt * cr;
                            gterm = gterm * (dxx + dzz);
                            gterm = gterm * vp2;
                            img = img + gterm;
                        }
                    }
                }
                if constexpr (BWD) {
                    __builtin_nontemporal_store(ill, a.illum + gbase + off);
                    __builtin_nontemporal_store(img, a.image + gbase + off);
                }
            }

            // 3. stage plane z+1 into the other LDS buffer, advance the queue
            if (more) {
                stage_plane(buf ^ 1, 5);
#pragma unroll
                for (int s = 0; s < SB; ++s) {
#pragma unroll
                    for (int f = 0; f < NF; ++f)
#pragma unroll
                        for (int r = 0; r < RPT; ++r) {
#pragma unroll
                            for (int d = 0; d < 8; ++d) q[s][f][r][d] = q[s][f][r][d + 1];
                            q[s][f][r][8] = qn[s][f][r];
                        }
#pragma unroll
                    for (int r = 0; r < RPT; ++r) {
                        o[s][r] = o_n[s][r];
                        v[s][r] = v_n[s][r];
                        if constexpr (BWD) ro[s][r] = ro_n[s][r];
                    }
                }
                if constexpr (BWD) {
#pragma unroll
                    for (int r = 0; r < RPT; ++r) {
                        im[r] = im_n[r];
                        il[r] = il_n[r];
                    }
                }
            }
            __syncthreads();
        }
    }
}

template <int MODE, int ARITH, int BY, int RPT, int SB>
hipError_t launch_one(const StepArgs& a, dim3 grid, hipStream_t st) {
    step_kernel<MODE, ARITH, BY, RPT, SB><<<grid, dim3(64, BY), 0, st>>>(a);
    return hipGetLastError();
}

using LaunchFn = hipError_t (*)(const StepArgs&, dim3, hipStream_t);

struct Entry {
    int mode, arith, by, rpt, sb;
    LaunchFn fn;
};

#define FWI_E(M, A, BY, RPT, SB) {M, A, BY, RPT, SB, &launch_one<M, A, BY, RPT, SB>}
#define FWI_TILES(M, A, SB)                                                                              \
    FWI_E(M, A, 4, 1, SB), FWI_E(M, A, 4, 2, SB), FWI_E(M, A, 8, 1, SB), FWI_E(M, A, 8, 2, SB),         \
        FWI_E(M, A, 4, 4, SB), FWI_E(M, A, 16, 1, SB)

#define FWI_BATCH(M)                                                                                     \
    FWI_E(M, ARITH_REF, 4, 1, 2), FWI_E(M, ARITH_REF, 4, 2, 2), FWI_E(M, ARITH_REF, 8, 1, 2),            \
        FWI_E(M, ARITH_REF, 16, 1, 2), FWI_E(M, ARITH_REF, 4, 1, 4), FWI_E(M, ARITH_REF, 8, 1, 4),        \
        FWI_E(M, ARITH_REF, 16, 1, 4)

const Entry kTable[] = {
    FWI_TILES(FWD_RBC, ARITH_REF, 1), FWI_TILES(FWD_ABC, ARITH_REF, 1),
    FWI_TILES(BWD_RTM, ARITH_REF, 1), FWI_TILES(BWD_FWI, ARITH_REF, 1),
    FWI_TILES(FWD_RBC, ARITH_FP32, 1), FWI_TILES(FWD_ABC, ARITH_FP32, 1),
    FWI_TILES(BWD_RTM, ARITH_FP32, 1), FWI_TILES(BWD_FWI, ARITH_FP32, 1),
    // batched: shots per block 2 and 4
    FWI_BATCH(FWD_RBC), FWI_BATCH(FWD_ABC), FWI_BATCH(BWD_RTM), FWI_BATCH(BWD_FWI),
};

const Entry* find(int mode, int arith, const Tuning& t) {
    for (const Entry& e : kTable)
        if (e.mode == mode && e.arith == arith && e.by == t.by && e.rpt == t.rpt && e.sb == t.sb) return &e;
    return nullptr;
}

}  // namespace

bool step_config_supported(int mode, int arith, const Tuning& t) {
    return find(mode, arith, t) != nullptr && t.zchunks >= 1;
}

void step_grid(const Layout& L, const Tuning& t, int nshots, int sb, StepArgs& a, dim3& grid, dim3& block) {
    const int TY = t.by * t.rpt;
    a.ntx = (L.nx + 63) / 64;
    a.nty = (L.ny - 8 + TY - 1) / TY;
    // z-chunks of zlen planes; recount so that no chunk is empty (every
    // chunk then satisfies 4 <= zb < ze <= nz-4 and its priming loads of
    // planes zb-4 .. zb+3 stay inside the volume)
    const int M = L.nz - 8;
    a.zlen = (M + t.zchunks - 1) / t.zchunks;
    a.nzc = (M + a.zlen - 1) / a.zlen;
    a.nitems = a.ntx * a.nty * a.nzc * (nshots / sb);
    const int blocks = (a.nitems + 7) / 8 * 8;
    grid = dim3(blocks);
    block = dim3(64, t.by);
}

hipError_t launch_step(int mode, int arith, const Tuning& t, const StepArgs& a, hipStream_t st) {
    const Entry* e = find(mode, arith, t);
    if (!e || a.nshots % t.sb != 0) return hipErrorInvalidConfiguration;
    StepArgs args = a;
    dim3 grid, block;
    step_grid(a.L, t, a.nshots, t.sb, args, grid, block);
    return e->fn(args, grid, st);
}

}  // namespace fwi

// pair_kernel.hpp -- the batched time-step kernel with two shots per thread.
//
// Same leapfrog steps as step_kernel (step_kernel.hpp; forward
// compute_pressure, src/modeling/modeling.cu:324-382; backward
// cross_correlation, src/migration/migration.cu:181-202, and build_gradient,
// src/inversion/inversion.cu:400-468), with each thread advancing one cell
// of two consecutive shots of a batch (a shot pair) instead of one:
//
//  * a block of 64 x BY threads owns a 64 x BY tile of the x-y plane of both
//    shots and marches along z; each thread keeps the 9 z-samples of its
//    column of both shots' stencil fields in registers, and both shots'
//    centre planes are staged (with their 4-cell halos) in LDS, double
//    buffered, one barrier per plane;
//  * the shots' velocity models differ only in the random boundary (the
//    model is set for the whole batch, and the boundary generators write
//    only cells within nb of a face), so in the interior the second shot's
//    velocity is the first shot's: the load is skipped there (an
//    out-of-range offset), 4 B per cell of the second shot;
//  * the pair shares one image/illumination (or gradient/illumination)
//    partial: both shots' terms are added in registers, in shot order, one
//    read-modify-write per cell for two shots, with the deferred cadence of
//    step_kernel (IMG 0, 2, 3, 4, 5: image every other launch, illumination
//    every third, each step's terms in shot order) -- the summation order
//    of step_kernel with two shots per block, so the partials and the
//    folded image are the same bits.
//
// HBM bytes per shot, cell and step (compulsory): forward 16 B with the
// velocity read for both shots, 15.1 B at the 401 x 401 x 201 (+50) grid
// where 43 % of the cells are interior; RTM backward 28 B of fields and
// velocity (27.1 B with the shared velocity) plus the image/illumination
// read-modify-writes, 6.7 B per step on average, halved by the pair: 30.5 B
// against 34.7 B for one shot per thread.
//
// Instances: reference arithmetic, no backward corrections (a batched pass
// with corrections runs the two-groups-per-block instances of step_kernel,
// which sum in the same order), volumes below 2 GiB.
#pragma once

#include "step_kernel.hpp"

namespace fwi {

namespace {

// at least four waves per SIMD (at most 128 VGPRs): one block of 16 waves,
// two of 8 or four of 4 per CU
template <int MODE, int ARITH, int BY, int IMG>
__global__ void __launch_bounds__(64 * BY, 4) pair_kernel(const StepArgs a) {
    constexpr int NF = (MODE == BWD_RTM || MODE == BWD_FWI) ? 2 : 1;
    constexpr bool BWD = NF == 2;
    static_assert(IMG == 1 || BWD, "deferred imaging: backward launches only");
    static_assert(IMG != 3 && IMG != 5 ? true : MODE == BWD_RTM, "image every other launch: RTM only");
    constexpr bool RMW_IMG = BWD && (IMG == 1 || IMG == 2 || IMG == 3 || IMG == 5 || MODE == BWD_FWI);
    constexpr bool RMW_ILL = BWD && (IMG == 1 || IMG == 2 || IMG == 4 || IMG == 5);
    constexpr bool ABC = MODE == FWD_ABC;
    constexpr int LW = 64 + 8;
    constexpr int LH = BY + 8;
    constexpr int HY = (8 + BY - 1) / BY;  // y-halo rows loaded per wave
    // centre planes of both shots, double buffered: [buffer][shot][row][col],
    // each entry a cell's NF fields side by side, so that one 8-byte LDS
    // access moves both fields and every neighbour of a thread is an
    // immediate offset from one address
    using CellT = typename std::conditional<NF == 2, float2, float>::type;
    __shared__ CellT lds[2][2][LH][LW];

    const Layout& L = a.L;
    const int tx = threadIdx.x;
    const int ty = __builtin_amdgcn_readfirstlane((int)threadIdx.y);

    // forward receiver gather of the current field (as step_kernel)
    if constexpr (!BWD) {
        if (a.seis_row) {
            const int nrec = a.rec_start[a.nshots];
            const int per = (nrec + gridDim.x - 1) / gridDim.x;
            const int beg = blockIdx.x * per;
            const int end = min(beg + per, nrec);
            for (int q = beg + threadIdx.y * 64 + tx; q < end; q += 64 * BY) {
                float v = a.cur[a.rec_off[q]];
                if (a.rec_damp) v = v * a.rec_damp[q];
                a.seis_row[q] = v;
            }
        }
    }

    const long vol_b = L.vol * 4;
    const unsigned plane_b = (unsigned)(L.pp * 4);

    // Item order (step_grid launches 8 * range blocks per shot group):
    // workgroups are dispatched round-robin over the 8 XCDs, so blocks b,
    // b+8, b+16, ... share an L2, and in block order.  The shot groups (the
    // slowest index of an item) are taken one after another by the whole
    // GPU, and within a group each XCD takes a contiguous range of tiles, so
    // the blocks running at any time hold neighbouring tiles of one group
    // (whose halo rows they share) rather than tiles of several groups.
    const int per_grp = a.ntx * a.nty * a.nzc;
    const int range = (per_grp + 7) >> 3;
    const int xcd = blockIdx.x & 7;
    const int nslots = range * (a.nitems / per_grp);

    for (int slot = blockIdx.x >> 3; slot < nslots; slot += gridDim.x >> 3) {
        const int gq = slot / range;
        const int t = xcd * range + (slot - gq * range);
        if (t >= per_grp) continue;  // uniform per block
        const int item = t + gq * per_grp;
        const int tile_x = item % a.ntx;
        const int r1 = item / a.ntx;
        const int tile_y = r1 % a.nty;
        const int r2 = r1 / a.nty;
        const int cz = r2 % a.nzc;
        const int grp = r2 / a.nzc;  // shot pair

        const int x = tile_x * 64 + tx;
        const int y0 = 4 + tile_y * BY;
        const int y = y0 + ty;
        const int zb = 4 + cz * a.zlen;
        const int ze = min(zb + a.zlen, L.nz - 4);
        if (zb >= ze) continue;  // uniform per block
        const bool xin = x >= 4 && x < L.nx - 4;
        const bool xld = x < L.nx;
        const int hxx = tx < 4 ? x - 4 : x + 4;
        const int hxc = tx < 4 ? tx : tx + 8;
        const bool hx_lane = tx < 4 || tx >= 60;
        const bool hxok = hx_lane && hxx >= 0 && hxx < L.nx;

        const int cell = x + y * L.px;
        const unsigned qo = (xld && y < L.ny) ? (unsigned)cell * 4u : OOB;
        const unsigned co = (xin && y < L.ny - 4) ? (unsigned)cell * 4u : OOB;
        const unsigned hxo = (hxok && y < L.ny) ? (unsigned)(hxx + y * L.px) * 4u : OOB;
        int hy_row[HY];
        unsigned hyo[HY];
#pragma unroll
        for (int m = 0; m < HY; ++m) {
            const int h = ty + m * BY;
            const int yy = h < 4 ? y0 - 4 + h : y0 + BY + (h - 4);
            hy_row[m] = h < 4 ? h : BY + h;
            hyo[m] = (h < 8 && xld && yy < L.ny) ? (unsigned)(x + yy * L.px) * 4u : OOB;
        }
        // interior in x and y (the second shot's velocity is the first's in
        // the interior planes: its load takes the offset co_sh, out of range
        // there); a.share_nb < 0: no sharing
        const int snb = a.share_nb;
        const bool lane_xy_shared =
            snb >= 0 && x >= snb && x < L.nx - snb && y >= snb && y < L.ny - snb;
        const unsigned co_sh = lane_xy_shared ? OOB : co;
        auto z_shared = [&](int zz) { return snb >= 0 && zz >= snb && zz < L.nz - snb; };  // block-uniform

        DampAxis dmp_x{0.0f, false};
        DampXY dmp_xy{};
        bool tile_damped = false;
        if constexpr (ABC) {
            dmp_x = damp_axis(a.b1d, a.nb, x, L.nx);
            dmp_xy = damp_xy(dmp_x, damp_axis(a.b1d, a.nb, y, L.ny));
            const int xa = tile_x * 64;
            tile_damped = xa < a.nb || xa + 63 >= L.nx - a.nb || y0 < a.nb || y0 + BY - 1 >= L.ny - a.nb;
        }

        // the pair's arrays: one wave-uniform buffer resource per array over
        // both shots' volumes (the host launches pairs only while two
        // volumes stay below 2 GiB); shot s of the pair at scalar offset
        // s * pair_b
        const long pbase = (long)(2 * grp) * a.sstride;
        const unsigned pair_b = (unsigned)(a.sstride * 4);
        Rsrc fr[NF], orr, vr, rr;
        fr[0] = make_rsrc(a.cur + pbase, 2 * vol_b);
        orr = make_rsrc(a.old + pbase, 2 * vol_b);
        vr = make_rsrc(a.vp + pbase, 2 * vol_b);
        if constexpr (BWD) {
            fr[1] = make_rsrc(a.rcur + pbase, 2 * vol_b);
            rr = make_rsrc(a.rold + pbase, 2 * vol_b);
        }
        int src_z[2] = {-1, -1}, src_c[2] = {-1, -1};
        if constexpr (!BWD) {
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const long so = a.src_off[2 * grp + s];
                if (so >= 0) {
                    src_z[s] = (int)(so / L.pp);
                    src_c[s] = (int)(so - (long)src_z[s] * L.pp);
                }
            }
        }
        const Rsrc ir = make_rsrc(a.image + (long)grp * a.istride, vol_b);
        const Rsrc lr_ = make_rsrc(a.illum + (long)grp * a.istride, vol_b);
        // scalar offset of plane zz (clamped to the volume) of shot s
        auto soff = [&](int zz, int s) { return (unsigned)min(max(zz, 0), L.nz - 1) * plane_b + s * pair_b; };

        // pipeline registers (two stream sets, [set]: plane z and plane
        // z+1; the two shots' velocities in separate arrays, so that the
        // per-lane choice between them is a select of values)
        float q[2][NF][9];
        float qn[2][NF];
        float hx[2][NF];
        float hy[2][NF][HY];
        float so[2][2], sro[2][2], sv0[2], sv1[2], sim[2], sil[2];

        auto load_plane = [&](int zz, auto sset) {
            constexpr int S = decltype(sset)::value;
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int f = 0; f < NF; ++f) {
                    hx[s][f] = bload(fr[f], hxo, soff(zz, s));
#pragma unroll
                    for (int m = 0; m < HY; ++m)  // wave-uniform
                        if (ty + m * BY < 8) hy[s][f][m] = bload(fr[f], hyo[m], soff(zz, s));
                }
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                so[S][s] = bload<kNT>(orr, co, soff(zz, s));
                if constexpr (BWD) sro[S][s] = bload<kNT>(rr, co, soff(zz, s));
            }
            sv0[S] = bload<kNT>(vr, co, soff(zz, 0));
            sv1[S] = bload<kNT>(vr, z_shared(zz) ? co_sh : co, soff(zz, 1));
            if constexpr (RMW_IMG) sim[S] = bload<kNT>(ir, co, soff(zz, 0));
            if constexpr (RMW_ILL) sil[S] = bload<kNT>(lr_, co, soff(zz, 0));
        };
        auto cell_of = [&](const float(&v)[NF]) {
            if constexpr (NF == 2)
                return make_float2(v[0], v[1]);
            else
                return v[0];
        };
        auto stage_plane = [&](int buf, int qi) {
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                CellT(*T)[LW] = lds[buf][s];
                float cv[NF];
#pragma unroll
                for (int f = 0; f < NF; ++f) cv[f] = q[s][f][qi];
                // x halos without a per-lane branch: a lane with no halo
                // cell rewrites its own centre value
                T[ty + 4][tx + 4] = cell_of(cv);
                T[ty + 4][hx_lane ? hxc : tx + 4] = hx_lane ? cell_of(hx[s]) : cell_of(cv);
#pragma unroll
                for (int m = 0; m < HY; ++m)
                    if (ty + m * BY < 8) {
                        float hv[NF];
#pragma unroll
                        for (int f = 0; f < NF; ++f) hv[f] = hy[s][f][m];
                        T[hy_row[m]][tx + 4] = cell_of(hv);
                    }
            }
        };

        // prologue: planes zb-4 .. zb+4 into the queues, plane zb into LDS
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int f = 0; f < NF; ++f)
#pragma unroll
                for (int d = 0; d < 9; ++d) q[s][f][d] = bload(fr[f], qo, soff(zb - 4 + d, s));
        load_plane(zb, Parity<0>{});
        stage_plane(0, 4);
        __syncthreads();

        // one plane; P = (z - zb) & 1 selects the LDS buffer and the stream set
        auto plane = [&](auto par, const int z) {
            constexpr int P = decltype(par)::value;
            constexpr int SC = P, SN = P ^ 1;
            // 1. prefetch plane z+1 (queue head z+5 clamped to the volume;
            // planes past the chunk are loaded and not used)
            {
                const int ph = min(z + 5, L.nz - 1);
#pragma unroll
                for (int s = 0; s < 2; ++s)
#pragma unroll
                    for (int f = 0; f < NF; ++f) qn[s][f] = bload(fr[f], qo, soff(ph, s));
                load_plane(z + 1, Parity<SN>{});
            }

            // 2. compute plane z
            DampAxis dmp_z{0.0f, false};
            bool damp = false;  // block-uniform
            if constexpr (ABC) {
                dmp_z = damp_axis(a.b1d, a.nb, z, L.nz);
                damp = tile_damped || dmp_z.on;
            }
            // the second shot's velocity: the first's where shared
            const float v_pair1 = (z_shared(z) && lane_xy_shared) ? sv0[SC] : sv1[SC];
            float res[2][NF];
            float t_ill[2], t_img[2], p1[2], p2[2], pimg[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int lr = ty + 4, lc = tx + 4;
                const CellT(*T)[LW] = lds[P][s];
                // the 16 x and y neighbours, both fields each
                const CellT xm1 = T[lr][lc - 1], xp1 = T[lr][lc + 1], xm2 = T[lr][lc - 2], xp2 = T[lr][lc + 2];
                const CellT xm3 = T[lr][lc - 3], xp3 = T[lr][lc + 3], xm4 = T[lr][lc - 4], xp4 = T[lr][lc + 4];
                const CellT ym1 = T[lr - 1][lc], yp1 = T[lr + 1][lc], ym2 = T[lr - 2][lc], yp2 = T[lr + 2][lc];
                const CellT ym3 = T[lr - 3][lc], yp3 = T[lr + 3][lc], ym4 = T[lr - 4][lc], yp4 = T[lr + 4][lc];
                auto fx = [](const CellT& v, auto fc) {
                    if constexpr (NF == 2)
                        return decltype(fc)::value == 0 ? v.x : v.y;
                    else
                        return v;
                };
                const float c = q[s][0][4];
                const float dxx = d2<ARITH>(c, fx(xm1, Parity<0>{}), fx(xp1, Parity<0>{}), fx(xm2, Parity<0>{}),
                                            fx(xp2, Parity<0>{}), fx(xm3, Parity<0>{}), fx(xp3, Parity<0>{}),
                                            fx(xm4, Parity<0>{}), fx(xp4, Parity<0>{}), a.idh2);
                const float dyy = d2<ARITH>(c, fx(ym1, Parity<0>{}), fx(yp1, Parity<0>{}), fx(ym2, Parity<0>{}),
                                            fx(yp2, Parity<0>{}), fx(ym3, Parity<0>{}), fx(yp3, Parity<0>{}),
                                            fx(ym4, Parity<0>{}), fx(yp4, Parity<0>{}), a.idh2);
                const float* qz = q[s][0];
                const float dzz = d2<ARITH>(c, qz[3], qz[5], qz[2], qz[6], qz[1], qz[7], qz[0], qz[8], a.idh2);
                const float vel = s == 0 ? sv0[SC] : v_pair1;
                const float vp2 = vel * vel;
                const float K = a.dt2 * vp2;
                if constexpr (!BWD) {
                    float oo = so[SC][s];
                    float dmp = 1.0f;
                    if constexpr (ABC) {
                        if (damp) {
                            dmp = dmp_z.on ? damper_zon(dmp_z.b, dmp_x.b, dmp_xy) : dmp_xy.dxy;
                            oo = oo * dmp;
                        }
                    }
                    float nv = K * ((dxx + dyy) + dzz);
                    nv = nv + 2.0f * c;
                    nv = nv - oo;
                    if constexpr (ABC) {
                        if (damp) nv = nv * dmp;
                    }
                    if (z == src_z[s] && cell == src_c[s]) nv = nv + a.inj;
                    res[s][0] = nv;
                } else {
                    const float cr = q[s][1][4];
                    const float rxx = d2<ARITH>(cr, fx(xm1, Parity<1>{}), fx(xp1, Parity<1>{}), fx(xm2, Parity<1>{}),
                                                fx(xp2, Parity<1>{}), fx(xm3, Parity<1>{}), fx(xp3, Parity<1>{}),
                                                fx(xm4, Parity<1>{}), fx(xp4, Parity<1>{}), a.idh2);
                    const float ryy = d2<ARITH>(cr, fx(ym1, Parity<1>{}), fx(yp1, Parity<1>{}), fx(ym2, Parity<1>{}),
                                                fx(yp2, Parity<1>{}), fx(ym3, Parity<1>{}), fx(yp3, Parity<1>{}),
                                                fx(ym4, Parity<1>{}), fx(yp4, Parity<1>{}), a.idh2);
                    const float* rz = q[s][1];
                    const float rzz = d2<ARITH>(cr, rz[3], rz[5], rz[2], rz[6], rz[1], rz[7], rz[0], rz[8], a.idh2);
                    const float o = so[SC][s];
                    float sn = K * ((dxx + dyy) + dzz);
                    sn = sn + 2.0f * c;
                    sn = sn - o;
                    float rn = K * ((rxx + ryy) + rzz);
                    rn = rn + 2.0f * cr;
                    rn = rn - sro[SC][s];
                    res[s][0] = sn;
                    res[s][1] = rn;
                    t_ill[s] = sn * sn;
                    if constexpr (MODE == BWD_RTM) {
                        t_img[s] = sn * cr;
                    } else {
                        float gterm = a.dt * cr;
                        gterm = gterm * (dxx + dzz);
                        t_img[s] = gterm * vp2;
                    }
                    // the earlier steps' terms this launch can still form:
                    // illumination of steps i-2 (new source field o) and
                    // i-1 (c), image of step i-1 (c with the receiver field
                    // it read, this launch's old receiver value)
                    p2[s] = o * o;
                    p1[s] = c * c;
                    pimg[s] = c * sro[SC][s];
                }
                // one shot at a time: its 16 neighbour reads are live only
                // during its own derivatives
                __builtin_amdgcn_sched_barrier(0);
            }
            // image and illumination: each step's terms in shot order, the
            // steps in order
            float ill = 0.0f, img = 0.0f;
            if constexpr (BWD) {
                if constexpr (RMW_ILL) {
                    ill = sil[SC];
                    if constexpr (IMG == 4 || IMG == 5) ill = (ill + p2[0]) + p2[1];
                    if constexpr (IMG == 2 || IMG == 4 || IMG == 5) ill = (ill + p1[0]) + p1[1];
                    ill = (ill + t_ill[0]) + t_ill[1];
                }
                if constexpr (RMW_IMG) {
                    img = sim[SC];
                    if constexpr (MODE == BWD_RTM && (IMG == 2 || IMG == 3 || IMG == 5)) img = (img + pimg[0]) + pimg[1];
                    img = (img + t_img[0]) + t_img[1];
                }
            }

            // keep the compute above this point: staging waits for the
            // prefetched loads, which stay in flight during the compute
            __builtin_amdgcn_sched_barrier(0);

            // 3. store plane z; stage plane z+1 into the other LDS buffer;
            // advance the queues
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                bstore<kNT>(orr, co, soff(z, s), res[s][0]);
                if constexpr (BWD) bstore<kNT>(rr, co, soff(z, s), res[s][1]);
            }
            if constexpr (RMW_ILL) bstore<kNT>(lr_, co, soff(z, 0), ill);
            if constexpr (RMW_IMG) bstore<kNT>(ir, co, soff(z, 0), img);
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int f = 0; f < NF; ++f) {
#pragma unroll
                    for (int d = 0; d < 8; ++d) q[s][f][d] = q[s][f][d + 1];
                    q[s][f][8] = qn[s][f];
                }
            stage_plane(P ^ 1, 4);
            __syncthreads();
        };
        int z = zb;
        for (; z + 1 < ze; z += 2) {
            plane(Parity<0>{}, z);
            plane(Parity<1>{}, z + 1);
        }
        if (z < ze) plane(Parity<0>{}, z);
    }
}

template <int MODE, int ARITH, int BY, int IMG>
hipError_t launch_pair(const StepArgs& a, dim3 grid, hipStream_t st) {
    pair_kernel<MODE, ARITH, BY, IMG><<<grid, dim3(64, BY), 0, st>>>(a);
    return hipGetLastError();
}

}  // namespace

}  // namespace fwi

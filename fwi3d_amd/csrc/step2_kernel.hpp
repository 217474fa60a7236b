// step2_kernel.hpp -- two time steps per launch (temporal blocking) for the
// random-boundary forward pass and the RTM backward pass.
//
// A launch reads the two time levels (cur, old) of each field once and
// writes the next two (the step-1 result becomes the new old, the step-2
// result the new cur) into two other buffers, so that no block overwrites
// what a neighbour still reads.  Per cell and two steps the compulsory HBM
// traffic is: forward 2 reads + 2 writes + vp = 20 B (16 B per step with one
// step per launch: 32 B); RTM 4 reads + 4 writes + vp + the image and
// illumination read-modify-write once for both steps = 52 B (one step per
// launch: 2 x 34.7 B with deferred imaging).
//
// Work decomposition (gfx950, wave64).  A block of BY waves owns a 64 x TY
// region of the x-y plane (RPT rows per thread, TY = BY * RPT) and marches
// along z.  Step 1 is computed on the whole region, step 2 on its inner
// 56 x (TY - 8) cells, which are the block's output; neighbouring blocks'
// regions overlap by 8 cells, and each computes the step-1 values of the
// overlap itself.  Iteration z computes step 1 at plane z + 4 and step 2 at
// plane z:
//   q1   the z-samples z .. z+8 of the step-1 stencil field (forward: the
//        current field; RTM: the current source field So and the receiver
//        field R), a register queue per row;
//   q2   the step-1 results of planes z-4 .. z+4 (z+4 computed this
//        iteration), the z-queue of step 2;
//   lds1 the step-1 stencil field of plane z+4 with a 4-cell halo
//        (double-buffered), for the x and y derivatives of step 1;
//   lds2 the step-1 results of plane z over the region (double-buffered),
//        for the x and y derivatives of step 2.
// The old field of step 2 is q1's plane z (the step-1 stencil field).  Eight
// iterations before the chunk's first plane compute step 1 only (priming);
// step-1 values outside the updated interior are zero, as the fields there
// (the host uses these launches only while the fields' 4-cell rims are zero:
// the output buffers rotate through four, whose rims nothing writes).
//
// Between the two steps, as the one-step launches do between launches: the
// forward pass adds the source term of the second step to the step-1 field
// and records the seismogram of the second step from it; the RTM pass
// injects the data of the second step into the step-1 receiver field (per
// cell, in receiver order).  Receiver cells are found through a per-shot map
// of each receiver plane (cell index or -1).  The arithmetic is the one-step
// kernel's, operation for operation, and the image / illumination terms are
// added in step order, so the results are bit-identical to two one-step
// launches.
#pragma once

#include "step_kernel.hpp"

namespace fwi {
namespace {

template <int MODE, int ARITH, int BY, int RPT>
__global__ void __launch_bounds__(64 * BY) step2_kernel(const Step2Args a) {
    constexpr int NF = MODE == BWD_RTM ? 2 : 1;
    constexpr bool BWD = NF == 2;
    static_assert(MODE == FWD_RBC || MODE == BWD_RTM, "two-step launches: random-boundary forward and RTM");
    constexpr int TY = BY * RPT;
    constexpr int LW = 64 + 8;
    constexpr int LH = TY + 8;
    constexpr int HY = (8 + BY - 1) / BY;
    __shared__ float lds1[2][NF][LH][LW];
    // same shape as lds1: the 4-cell margin is never written; lanes of the
    // region's rim read it and discard their step-2 results
    __shared__ float lds2[2][NF][LH][LW];

    const Layout& L = a.L;
    const int tx = threadIdx.x;
    const int ty = __builtin_amdgcn_readfirstlane((int)threadIdx.y);

    // receiver gather of the first step from the current field (read-only
    // in this launch: the outputs go to other buffers)
    if constexpr (!BWD) {
        if (a.seis_row0) {
            const int nrec = a.rec_start[a.nshots];
            const int per = (nrec + gridDim.x - 1) / gridDim.x;
            const int beg = blockIdx.x * per;
            const int end = min(beg + per, nrec);
            for (int q = beg + threadIdx.y * 64 + tx; q < end; q += 64 * BY) a.seis_row0[q] = a.cur[a.rec_off[q]];
        }
    }

    const long vol_b = L.vol * 4;
    const unsigned plane_b = (unsigned)(L.pp * 4);
    const long plane_n = (long)L.ny * L.px;  // cells of a receiver-plane map

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
        const int shot = r2 / a.nzc;

        const int zb = 4 + cz * a.zlen;
        const int ze = min(zb + a.zlen, L.nz - 4);
        if (zb >= ze) continue;
        const int ox = 4 + tile_x * 56;         // first output column
        const int oy = 4 + tile_y * (TY - 8);   // first output row
        const int x = ox - 4 + tx;
        const bool xld = x < L.nx;
        const bool xin = x >= 4 && x < L.nx - 4;
        const bool xout = tx >= 4 && tx < 60 && x < L.nx - 4;
        const int hxx = tx < 4 ? x - 4 : x + 4;
        const int hxc = tx < 4 ? tx : tx + 8;
        const bool hx_lane = tx < 4 || tx >= 60;
        const bool hxok = hx_lane && hxx >= 0 && hxx < L.nx;

        // offsets: qo = step-1 stencil field (any cell of the volume), co =
        // updated cell (step-1 streams), wo = output cell (stores, step-2
        // streams), hxo / hyo = halo cells; OOB where there is none
        // (the step-1 and output offsets are derived from co on use: a cell
        // of the region is loaded if it lies in the volume, updated if it
        // lies in the updated interior, an output if it is also inner)
        unsigned co[RPT], hxo[RPT];
        unsigned rowbits = 0;  // bit r: row r loaded; bit RPT + r: updated; bit 2 RPT + r: output
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int row = ty + r * BY;
            const int yr = oy - 4 + row;
            const bool yld = yr >= 0 && yr < L.ny;
            const bool yin = yr >= 4 && yr < L.ny - 4;
            const bool yout = row >= 4 && row < TY - 4 && yr < L.ny - 4;
            co[r] = (unsigned)(x + yr * L.px) * 4u;
            rowbits |= ((xld && yld) ? 1u : 0u) << r;
            rowbits |= ((xin && yin) ? 1u : 0u) << (RPT + r);
            rowbits |= ((xout && yout) ? 1u : 0u) << (2 * RPT + r);
            hxo[r] = (hxok && yld) ? (unsigned)(hxx + yr * L.px) * 4u : OOB;
        }
        auto qo = [&](int r) { return ((rowbits >> r) & 1u) ? co[r] : OOB; };
        auto uo = [&](int r) { return ((rowbits >> (RPT + r)) & 1u) ? co[r] : OOB; };
        auto wo = [&](int r) { return ((rowbits >> (2 * RPT + r)) & 1u) ? co[r] : OOB; };
        auto upd = [&](int r) { return ((rowbits >> (RPT + r)) & 1u) != 0; };
        int hy_row[HY];
        unsigned hyo[HY];
#pragma unroll
        for (int m = 0; m < HY; ++m) {
            const int h = ty + m * BY;
            const int yy = h < 4 ? oy - 8 + h : oy - 4 + TY + (h - 4);
            hy_row[m] = h < 4 ? h : TY + h;
            hyo[m] = (h < 8 && xld && yy >= 0 && yy < L.ny) ? (unsigned)(x + yy * L.px) * 4u : OOB;
        }

        const long sbase = (long)shot * a.sstride;
        Vol<false> F1[NF], O1[NF], W1[NF], W2[NF];
        F1[0] = Vol<false>(a.cur + sbase, vol_b, L, plane_b);
        O1[0] = Vol<false>(a.old + sbase, vol_b, L, plane_b);
        W1[0] = Vol<false>(a.out_old + sbase, vol_b, L, plane_b);
        W2[0] = Vol<false>(a.out_cur + sbase, vol_b, L, plane_b);
        if constexpr (BWD) {
            F1[1] = Vol<false>(a.rcur + sbase, vol_b, L, plane_b);
            O1[1] = Vol<false>(a.rold + sbase, vol_b, L, plane_b);
            W1[1] = Vol<false>(a.rout_old + sbase, vol_b, L, plane_b);
            W2[1] = Vol<false>(a.rout_cur + sbase, vol_b, L, plane_b);
        }
        const Vol<false> VP(a.vp + sbase, vol_b, L, plane_b);
        const Vol<false> IM(a.image + (long)shot * a.istride, vol_b, L, plane_b);
        const Vol<false> IL(a.illum + (long)shot * a.istride, vol_b, L, plane_b);
        int src_z = -1, src_c = -1;
        if constexpr (!BWD) {
            const long so = a.src_off[shot];
            if (so >= 0) {
                src_z = (int)(so / L.pp);
                src_c = (int)(so - (long)src_z * L.pp);
            }
        }
        // receiver planes of this shot (the map is used on these planes only)
        int rp_z[kMaxRecPlanes];
#pragma unroll
        for (int k = 0; k < kMaxRecPlanes; ++k) rp_z[k] = a.rp_z ? a.rp_z[shot * kMaxRecPlanes + k] : -1;
        // the maps as a buffer: the cell's byte offset in a map is co
        const Rsrc rpm = make_rsrc(reinterpret_cast<const float*>(a.rp_map),
                                   a.rp_map ? (long)a.nshots * kMaxRecPlanes * plane_n * 4 : 0);

        float q1[NF][RPT][9], q2[NF][RPT][9];
        float qn[NF][RPT];
        float hx[NF][RPT], hy[NF][HY];
        float o1s[2][NF][RPT], v1s[2][RPT];  // step-1 streams (old, vp), two sets by parity
        float v2[RPT], im[RPT], il[RPT];     // step-2 streams of the current iteration

        // halos of plane zz (the staged plane) and the step-1 streams of
        // plane zz into set S
        auto load_step1 = [&](int zz, auto sset) {
            constexpr int S = decltype(sset)::value;
#pragma unroll
            for (int f = 0; f < NF; ++f) {
#pragma unroll
                for (int r = 0; r < RPT; ++r) hx[f][r] = F1[f].ld(hxo[r], zz);
#pragma unroll
                for (int m = 0; m < HY; ++m)
                    if (ty + m * BY < 8) hy[f][m] = F1[f].ld(hyo[m], zz);
#pragma unroll
                for (int r = 0; r < RPT; ++r) o1s[S][f][r] = O1[f].template ld<kNT>(uo(r), zz);
            }
#pragma unroll
            for (int r = 0; r < RPT; ++r) v1s[S][r] = VP.template ld<kNT>(uo(r), zz);
        };
        auto stage1 = [&](int buf) {
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                float(*T)[LW] = lds1[buf][f];
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
                    const float cv = q1[f][r][4];
                    T[ty + r * BY + 4][tx + 4] = cv;
                    T[ty + r * BY + 4][hx_lane ? hxc : tx + 4] = hx_lane ? hx[f][r] : cv;
                }
#pragma unroll
                for (int m = 0; m < HY; ++m)
                    if (ty + m * BY < 8) T[hy_row[m]][tx + 4] = hy[f][m];
            }
        };
        auto stage2 = [&](int buf, int qi) {
#pragma unroll
            for (int f = 0; f < NF; ++f)
#pragma unroll
                for (int r = 0; r < RPT; ++r) lds2[buf][f][ty + r * BY + 4][tx + 4] = q2[f][r][qi];
        };

        // prologue: q1 = planes z0 .. z0+8 (z0 = zb - 8; planes outside the
        // volume read 0), step-1 streams and halos of plane z0 + 4, staged
        const int z0 = zb - 8;
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
#pragma unroll
                for (int d = 0; d < 9; ++d) q1[f][r][d] = F1[f].ld(qo(r), z0 + d);
#pragma unroll
                for (int d = 0; d < 9; ++d) q2[f][r][d] = 0.0f;
            }
        load_step1(z0 + 4, Parity<0>{});
        stage1(0);
        __syncthreads();

        auto iteration = [&](auto par, const int z) {
            constexpr int P = decltype(par)::value;
            const int p = z + 4;          // plane of step 1
            const bool main = z >= zb;    // block-uniform: step 2 at plane z
            // 1. prefetch for the next iteration: queue head z+9, halos and
            // step-1 streams of plane z+5
            {
                const int ph = min(z + 9, L.nz - 1);
#pragma unroll
                for (int f = 0; f < NF; ++f)
#pragma unroll
                    for (int r = 0; r < RPT; ++r) qn[f][r] = F1[f].ld(qo(r), ph);
                load_step1(z + 5, Parity<P ^ 1>{});
            }
            // 2. step-2 streams of plane z (used after step 1)
            if (main) {
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
                    v2[r] = VP.template ld<kNT>(wo(r), z);
                    if constexpr (BWD) {
                        im[r] = IM.template ld<kNT>(wo(r), z);
                        il[r] = IL.template ld<kNT>(wo(r), z);
                    }
                }
            }
            // 3. step 1 at plane p
            const bool pin = p >= 4 && p < L.nz - 4;     // block-uniform
            const bool pown = p >= zb && p < ze;          // this chunk's plane
            int rk = -1;                                  // receiver plane slot of p
#pragma unroll
            for (int k = 0; k < kMaxRecPlanes; ++k)
                if (rp_z[k] == p) rk = k;
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const int lr = ty + r * BY + 4;
                const int lc = tx + 4;
                const float vel = v1s[P][r];
                const float vp2 = vel * vel;
                const float K = a.dt2 * vp2;
                float res[NF];
#pragma unroll
                for (int f = 0; f < NF; ++f) {
                    const float(*T)[LW] = lds1[P][f];
                    const float c = q1[f][r][4];
                    const float dxx = d2<ARITH>(c, T[lr][lc - 1], T[lr][lc + 1], T[lr][lc - 2], T[lr][lc + 2],
                                                T[lr][lc - 3], T[lr][lc + 3], T[lr][lc - 4], T[lr][lc + 4], a.idh2);
                    const float dyy = d2<ARITH>(c, T[lr - 1][lc], T[lr + 1][lc], T[lr - 2][lc], T[lr + 2][lc],
                                                T[lr - 3][lc], T[lr + 3][lc], T[lr - 4][lc], T[lr + 4][lc], a.idh2);
                    const float* qz = q1[f][r];
                    const float dzz = d2<ARITH>(c, qz[3], qz[5], qz[2], qz[6], qz[1], qz[7], qz[0], qz[8], a.idh2);
                    float nv = K * ((dxx + dyy) + dzz);
                    nv = nv + 2.0f * c;
                    nv = nv - o1s[P][f][r];
                    res[f] = nv;
                }
                if constexpr (!BWD) {
                    if (p == src_z && (int)(co[r] >> 2) == src_c) res[0] = res[0] + a.inj1;
                }
                if (!(upd(r) && pin)) {
#pragma unroll
                    for (int f = 0; f < NF; ++f) res[f] = 0.0f;
                }
                if (rk >= 0 && upd(r)) {  // receiver plane: inject (RTM) or record (forward)
                    const unsigned moff = (unsigned)(((long)shot * kMaxRecPlanes + rk) * plane_n * 4);
                    const int cidx = __builtin_bit_cast(int, bload<0>(rpm, co[r], moff));
                    if (cidx >= 0) {
                        if constexpr (BWD) {
                            if (a.data_row1)
                                for (int q = a.cell_start[cidx]; q < a.cell_start[cidx + 1]; ++q)
                                    res[1] = res[1] + a.idh3 * a.data_row1[a.cell_rec[q]];
                        } else {
                            if (a.seis_row1 && pown && wo(r) != OOB)
                                for (int q = a.cell_start[cidx]; q < a.cell_start[cidx + 1]; ++q)
                                    a.seis_row1[a.cell_rec[q]] = res[0];
                        }
                    }
                }
#pragma unroll
                for (int f = 0; f < NF; ++f) {
                    q2[f][r][8] = res[f];
                    if (pown) W1[f].template st<kNT>(wo(r), p, res[f]);
                }
                // one row at a time: keeps the stencil operands of only one
                // row live
                __builtin_amdgcn_sched_barrier(0);
            }
            // 4. step 2 at plane z
            if (main) {
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
                    const int cr = ty + r * BY + 4;
                    const int cc = tx + 4;
                    const float vel = v2[r];
                    const float vp2 = vel * vel;
                    const float K = a.dt2 * vp2;
                    float res[NF];
#pragma unroll
                    for (int f = 0; f < NF; ++f) {
                        const float(*T)[LW] = lds2[P][f];
                        const float c = q2[f][r][4];
                        const float dxx = d2<ARITH>(c, T[cr][cc - 1], T[cr][cc + 1], T[cr][cc - 2], T[cr][cc + 2],
                                                    T[cr][cc - 3], T[cr][cc + 3], T[cr][cc - 4], T[cr][cc + 4],
                                                    a.idh2);
                        const float dyy = d2<ARITH>(c, T[cr - 1][cc], T[cr + 1][cc], T[cr - 2][cc], T[cr + 2][cc],
                                                    T[cr - 3][cc], T[cr + 3][cc], T[cr - 4][cc], T[cr + 4][cc],
                                                    a.idh2);
                        const float* qz = q2[f][r];
                        const float dzz =
                            d2<ARITH>(c, qz[3], qz[5], qz[2], qz[6], qz[1], qz[7], qz[0], qz[8], a.idh2);
                        float nv = K * ((dxx + dyy) + dzz);
                        nv = nv + 2.0f * c;
                        nv = nv - q1[f][r][0];
                        res[f] = nv;
                    }
                    if constexpr (!BWD) {
                        if (z == src_z && (int)(co[r] >> 2) == src_c) res[0] = res[0] + a.inj2;
                    }
#pragma unroll
                    for (int f = 0; f < NF; ++f) W2[f].template st<kNT>(wo(r), z, res[f]);
                    if constexpr (BWD) {
                        // the two steps' terms in step order: the reconstructed
                        // source field with the receiver field it is paired with
                        const float s1 = q2[0][r][4], s2 = res[0];
                        float ill = il[r] + s1 * s1;
                        ill = ill + s2 * s2;
                        float img = im[r] + s1 * q1[1][r][0];
                        img = img + s2 * q2[1][r][4];
                        IL.template st<kNT>(wo(r), z, ill);
                        IM.template st<kNT>(wo(r), z, img);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
            // 5. stage plane z+5 of the step-1 field and plane z+1 of the
            // step-1 results, advance the queues
            stage2(P ^ 1, 5);
#pragma unroll
            for (int f = 0; f < NF; ++f)
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
#pragma unroll
                    for (int d = 0; d < 8; ++d) {
                        q1[f][r][d] = q1[f][r][d + 1];
                        q2[f][r][d] = q2[f][r][d + 1];
                    }
                    q1[f][r][8] = qn[f][r];
                }
            stage1(P ^ 1);
            __syncthreads();
        };
        // z0 = zb - 8: eight priming iterations (step 1 only), then the
        // chunk; an even count keeps the parity of z - z0 in step with the
        // compile-time set index
        int z = z0;
        for (; z + 1 < ze; z += 2) {
            iteration(Parity<0>{}, z);
            iteration(Parity<1>{}, z + 1);
        }
        if (z < ze) iteration(Parity<0>{}, z);
    }
}

template <int MODE, int ARITH, int BY, int RPT>
hipError_t launch_two(const Step2Args& a, dim3 grid, hipStream_t st) {
    step2_kernel<MODE, ARITH, BY, RPT><<<grid, dim3(64, BY), 0, st>>>(a);
    return hipGetLastError();
}

}  // namespace

using Launch2Fn = hipError_t (*)(const Step2Args&, dim3, hipStream_t);

struct Entry2 {
    int mode, arith, by, rpt;
    Launch2Fn fn;
};

}  // namespace fwi

// step_kernel.hpp -- the time-step kernel template of the wave-propagation
// path and the table entry type of its instances.  The instances are
// compiled in several translation units (step_inst_*.hip) so that they build
// in parallel; kernels_step.hip picks one per launch.
//
// One launch advances every cell 4 <= idx < n-4 of one or more shots by one
// leapfrog step of the 8th-order acoustic scheme:
//   forward  (compute_pressure, src/modeling/modeling.cu:324-382)
//       new = dt^2 vp^2 L(cur) + 2 cur - old            [+ Cerjan damping]
//   backward (cross_correlation, src/migration/migration.cu:181-202, and
//             build_gradient, src/inversion/inversion.cu:400-468)
//       S   = dt^2 vp^2 L(So) + 2 So - S   (time-reversed source field)
//       Ro  = dt^2 vp^2 L(R)  + 2 R  - Ro  (receiver field)
//       illum += S^2;  image += S R   |  grad += dt R (Dxx + Dzz)(So) vp^2
//   with the optional backward corrections of StepArgs::phys: the source
//   term put back into S during reconstruction, So instead of S in the
//   imaging condition, Dyy in the gradient.  They are compiled only into
//   the CORR instances, so the default backward kernels carry none of
//   their branches or registers.  With aligned imaging the previous step's
//   deferred terms (below) use its current source field, this launch's
//   old value, instead of its new one.
//
// Deferred imaging (IMG).  A backward launch holds, besides this step's
// fields, the two values the previous step's imaging terms were made of:
// its new source field is this launch's current one (c) and the receiver
// field it read is this launch's old receiver value (ro).  So the image and
// illumination read-modify-write can be done every other launch: IMG = 0
// leaves them out (the gradient of the FWI mode, whose term needs the
// previous step's derivatives, is still updated), IMG = 2 adds the previous
// step's terms (c*ro, c*c) and then this step's, in that order, and IMG = 1
// is the plain per-step update.  In batched launches the other groups pass
// both steps' terms to group 0, which adds all groups' terms of the
// previous step before any of this step's.  The sums are bit for bit those
// of per-step updates; the image/illumination traffic halves (single shot,
// RTM: 44 -> 36 B per cell and step on average).
// Except with aligned imaging, the illumination term of the step before the
// previous one is known as well (its new source field is this launch's old
// value o), so the illumination is updated every third launch and the
// image every other one: IMG = 3 adds the image terms of the previous step
// and this one, IMG = 4 the illumination terms of the two previous steps
// (o*o, c*c) and this one, IMG = 5 both (single shot, RTM: 36 -> 34.7 B per
// cell and step on average); batched launches again add the steps in order,
// each step's terms in shot order.  Terms still pending when a pass or a
// segment ends are added by a separate elementwise launch.
//
// Work decomposition (gfx950, wave64):
//  * a block owns a 64 x TY tile of the x-y plane (TY = BY*RPT rows; each of
//    the BY waves covers one 64-float row, 256 B, per register row) and
//    marches along z over one z-chunk;
//  * each thread keeps the 9 z-samples z-4..z+4 of its column in a register
//    queue (Dzz needs no LDS and no re-read);
//  * the centre plane z is staged with a 4-cell halo in LDS, double
//    buffered, so one barrier per plane suffices; Dxx and Dyy read LDS;
//  * HBM traffic per cell and step: read cur, old, vp; write new (16 B),
//    plus the halo rows a tile re-reads from its neighbours (8 of every TY
//    rows and 8 of every 64 columns of the stencil field; the XCD-aware
//    item order keeps neighbouring tiles on one XCD, but most of these
//    re-reads still miss its L2 at the 401 x 401 x 201 size).
//  * old and new share one buffer: each cell reads its own old value and
//    overwrites it, so no second field write is needed.  The ABC damping of
//    the current field (the reference's P *= damper, which races with the
//    neighbours' reads) is applied instead when the cell is next read as
//    "old", giving the same values in a single pass.
//  * batched launches put SB shots of the same tile in one block, one shot
//    per group of BY waves (so each thread carries one shot's registers, as
//    in a single-shot launch); in the backward modes only group 0 reads and
//    writes image and illumination, after adding the other groups' terms
//    of the plane, which they pass through LDS: that read-modify-write is
//    shared by the SB shots.
//
// Arithmetic.  ARITH_REF reproduces the reference kernel literally: float
// pair sums, double coefficients and accumulation per axis, rounding to
// float per axis, then the float update in source order; the library is
// built with -ffp-contract=off so each operation rounds once, which makes
// the result bit-identical to the CPU oracle.  ARITH_FP32 evaluates the
// derivative with FP32 FMAs.
#pragma once

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

// A thread's x-y part of damper(z, x, y), fixed while it marches along z:
// the whole damper when the plane is outside the z layer (dxy), and what
// the three cases with the z axis on need (the same float operations as
// damper() with i = z, j = x, k = y).
struct DampXY {
    float dxy;  // damper(off, x, y)
    float e;    // one of x, y on: that axis' b
    float t2;   // both on: b_x + b_y
    float by;
    int cnt;    // axes on among x, y
};

__device__ __forceinline__ DampXY damp_xy(DampAxis x, DampAxis y) {
    return {damper(DampAxis{0.0f, false}, x, y), y.on ? y.b : x.b, x.b + y.b, y.b, (int)x.on + (int)y.on};
}

// damper(z, x, y) for a plane inside the z layer (z.on)
__device__ __forceinline__ float damper_zon(float bz, float bx, const DampXY& d) {
    if (d.cnt == 0) return bz;
    if (d.cnt == 1) return (bz + d.e) - 1.0f;
    const float t1 = bz + bx, t3 = bz + d.by;
    return ((t1 + d.t2) + t3) - 5.0f;
}

// Buffer loads and stores.  Every volume of a launch is addressed through a
// wave-uniform buffer resource (base, size in bytes) held in scalar
// registers; a lane supplies one 32-bit in-plane byte offset and the plane
// offset is a scalar.  Lanes with nothing to load or store use the offset
// OOB, which the hardware range check turns into a load of 0 and a dropped
// store: the loop body has no per-lane branches, so the compiler can keep
// several planes of loads and stores in flight with exact wait counts.
// For volumes below 2 GiB an in-range offset plus the plane offset never
// reaches OOB; larger volumes are addressed per plane (Vol<true> below).
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr unsigned OOB = 0x80000000u;
constexpr int kBufferFlags = 0x00020000;  // gfx9 buffer resource word 3: 32-bit data, raw addressing
constexpr int kNT = 2;                    // cache policy: non-temporal stream

__device__ __forceinline__ Rsrc make_rsrc(const float* p, long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)bytes, kBufferFlags);
}
template <int AUX = 0>
__device__ __forceinline__ float bload(Rsrc r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, AUX));
}
template <int AUX = 0>
__device__ __forceinline__ void bstore(Rsrc r, unsigned voff, unsigned soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff, AUX);
}

template <int P>
struct Parity {
    static constexpr int value = P;
};

// One volume of a launch, addressed by plane.  Below 2 GiB (BIG = false): a
// wave-uniform buffer resource over the whole volume, the plane's byte
// offset as the scalar offset.  From 2 GiB on (BIG = true), where 32-bit
// byte offsets no longer reach every plane: a resource per access, based at
// the plane (64-bit base) and covering one plane; planes outside the volume
// get an empty range.  A plane index outside [0, nz) reads the nearest plane
// of the volume below 2 GiB and 0 from 2 GiB on (callers discard such
// values); the lane offset OOB reads 0 and drops stores.
template <bool BIG>
struct Vol {
    Rsrc r;
    const float* base;
    long pp;
    int nz;
    unsigned plane_b;
    __device__ __forceinline__ Vol(const float* p, long vol_b, const Layout& L, unsigned pb)
        : r(make_rsrc(p, BIG ? 0 : vol_b)), base(p), pp(L.pp), nz(L.nz), plane_b(pb) {}
    __device__ __forceinline__ Vol() : r(), base(nullptr), pp(0), nz(0), plane_b(0) {}
    __device__ __forceinline__ Rsrc plane(int zz, unsigned* soff) const {
        if constexpr (!BIG) {
            *soff = (unsigned)min(max(zz, 0), nz - 1) * plane_b;
            return r;
        } else {
            const bool in = zz >= 0 && zz < nz;
            *soff = 0;
            return make_rsrc(base + (in ? (long)zz * pp : 0), in ? (long)plane_b : 0);
        }
    }
    template <int AUX = 0>
    __device__ __forceinline__ float ld(unsigned voff, int zz) const {
        unsigned so;
        const Rsrc rz = plane(zz, &so);
        return bload<AUX>(rz, voff, so);
    }
    template <int AUX = 0>
    __device__ __forceinline__ void st(unsigned voff, int zz, float v) const {
        unsigned so;
        const Rsrc rz = plane(zz, &so);
        bstore<AUX>(rz, voff, so, v);
    }
};

template <int MODE, int ARITH, int BY, int RPT, int SB, bool CORR, int IMG, int PD, bool BIG>
__global__ void __launch_bounds__(64 * BY * SB) step_kernel(const StepArgs a) {
    constexpr int NF = (MODE == BWD_RTM || MODE == BWD_FWI) ? 2 : 1;
    constexpr bool BWD = NF == 2;
    static_assert(IMG == 1 || (BWD && (!CORR || SB == 1)), "deferred imaging: backward launches only");
    static_assert(IMG <= 2 || ((!CORR || SB == 1) && (IMG == 4 || MODE == BWD_RTM)),
                  "image every other launch, illumination every third: backward launches only");
    static_assert(PD == 1 || (PD == 2 && SB == 1), "two-plane prefetch: one shot group per block");
    // image (or gradient) and illumination read-modify-writes of this launch
    constexpr bool RMW_IMG = BWD && (IMG == 1 || IMG == 2 || IMG == 3 || IMG == 5 || MODE == BWD_FWI);
    constexpr bool RMW_ILL = BWD && (IMG == 1 || IMG == 2 || IMG == 4 || IMG == 5);
    constexpr bool ABC = MODE == FWD_ABC;
    constexpr int TY = BY * RPT;
    constexpr int LW = 64 + 8;
    constexpr int LH = TY + 8;
    constexpr int HY = (8 + BY - 1) / BY;  // y-halo rows loaded per wave
    __shared__ float lds[2][SB * NF][LH][LW];
    // batched backward: the image/illumination terms of groups 1..SB-1,
    // double-buffered by plane parity, summed by group 0
    // (two terms per group and step; IMG = 2 passes two steps' terms, and
    // group 0 parks its own terms of this step in the last two slots;
    // IMG >= 3: five terms per group, slots 5k .. 5k+4 = the illumination
    // terms of steps i-2, i-1, i, the image terms of steps i-1, i, and
    // group 0 parks its illumination terms of i-1, i and image term of i)
    constexpr int NX = (BWD && SB > 1)
                           ? (IMG >= 3 ? 5 * (SB - 1) + 3 : IMG == 2 ? 4 * (SB - 1) + 2 : 2 * (SB - 1))
                           : 1;
    __shared__ float xch[2][NX][TY][64];

    const Layout& L = a.L;
    const int tx = threadIdx.x;
    // wave group = shot within the block (wave-uniform: kept in a scalar register)
    const int g = SB == 1 ? 0 : __builtin_amdgcn_readfirstlane((int)threadIdx.y / BY);
    // wave row within the group's tile (wave-uniform)
    const int ty = __builtin_amdgcn_readfirstlane(SB == 1 ? (int)threadIdx.y : (int)threadIdx.y % BY);

    // Forward receiver gather of the current field (compute_seismogram,
    // modeling.cu:384-393).  The current field is read-only in this launch
    // (new values go to the old buffer), so any block may read it.
    if constexpr (!BWD) {
        if (a.seis_row) {
            const int nrec = a.rec_start[a.nshots];
            const int per = (nrec + gridDim.x - 1) / gridDim.x;
            const int beg = blockIdx.x * per;
            const int end = min(beg + per, nrec);
            for (int q = beg + threadIdx.y * 64 + tx; q < end; q += 64 * BY * SB) {
                float v = a.cur[a.rec_off[q]];
                if (a.rec_damp) v = v * a.rec_damp[q];
                a.seis_row[q] = v;
            }
        }
    }

    const long vol_b = L.vol * 4;                  // bytes of one volume (< 2 GiB)
    const unsigned plane_b = (unsigned)(L.pp * 4);  // bytes of one z-plane

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

        // in-plane byte offsets: qo = stencil-field load (any cell of the
        // volume), co = updated cell (loads of the streams, stores), hxo /
        // hyo = halo cells; OOB where the lane has no such cell
        unsigned qo[RPT], co[RPT], hxo[RPT];
        int cell[RPT];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int yr = y0 + ty + r * BY;
            cell[r] = x + yr * L.px;
            qo[r] = (xld && yr < L.ny) ? (unsigned)cell[r] * 4u : OOB;
            co[r] = (xin && yr < L.ny - 4) ? (unsigned)cell[r] * 4u : OOB;
            hxo[r] = (hxok && yr < L.ny) ? (unsigned)(hxx + yr * L.px) * 4u : OOB;
        }
        int hy_row[HY];
        unsigned hyo[HY];
#pragma unroll
        for (int m = 0; m < HY; ++m) {
            const int h = ty + m * BY;
            const int yy = h < 4 ? y0 - 4 + h : y0 + TY + (h - 4);
            hy_row[m] = h < 4 ? h : TY + h;
            hyo[m] = (h < 8 && xld && yy < L.ny) ? (unsigned)(x + yy * L.px) * 4u : OOB;
        }

        DampAxis dmp_x{0.0f, false};
        DampXY dmp_xy[RPT];
        // block-uniform: does the tile reach into the absorbing layer in x
        // or y?  If not, its cells are damped only on planes within nb of a
        // z face; elsewhere the damper is exactly 1 and is skipped.
        bool tile_damped = false;
        if constexpr (ABC) {
            dmp_x = damp_axis(a.b1d, a.nb, x, L.nx);
#pragma unroll
            for (int r = 0; r < RPT; ++r) dmp_xy[r] = damp_xy(dmp_x, damp_axis(a.b1d, a.nb, y0 + ty + r * BY, L.ny));
            const int xa = tile_x * 64, ya = y0;
            tile_damped = xa < a.nb || xa + 63 >= L.nx - a.nb || ya < a.nb || ya + TY - 1 >= L.ny - a.nb;
        }

        // this group's shot: buffer resources (wave-uniform)
        const int shot = grp * SB + g;
        const long sbase = (long)shot * a.sstride;
        Vol<BIG> fb[NF], ob, vb, rb;
        fb[0] = Vol<BIG>(a.cur + sbase, vol_b, L, plane_b);
        ob = Vol<BIG>(a.old + sbase, vol_b, L, plane_b);
        vb = Vol<BIG>(a.vp + sbase, vol_b, L, plane_b);
        if constexpr (BWD) {
            fb[1] = Vol<BIG>(a.rcur + sbase, vol_b, L, plane_b);
            rb = Vol<BIG>(a.rold + sbase, vol_b, L, plane_b);
        }
        int src_z = -1, src_c = -1;
        if (!BWD || (CORR && (a.phys & PHYS_SUBTRACT_SOURCE))) {  // uniform
            const long so = a.src_off[shot];
            if (so >= 0) {
                src_z = (int)(so / L.pp);
                src_c = (int)(so - (long)src_z * L.pp);
            }
        }
        // image/illumination of the shot group: read and written by group 0
        const Vol<BIG> ib(a.image + (long)grp * a.istride, vol_b, L, plane_b);
        const Vol<BIG> lb(a.illum + (long)grp * a.istride, vol_b, L, plane_b);
        const bool owner = g == 0;

        // Pipeline registers.  The loads of plane z + PD (PD = 1 or 2, the
        // prefetch depth) are issued while plane z is computed.
        //   q       z-samples z-4..z+4 of the stencil fields;
        //   qn[k]   queue heads in flight (plane z+5, and z+6 with PD = 2);
        //   hx, hy  halo values of the planes in flight (one set per plane);
        //   s*      per-cell streams (old, vp, receiver old, image,
        //           illumination): plane z and the PD planes in flight.
        // The loop body is instantiated once per phase of the set rotation
        // (two phases with PD = 1, six with PD = 2), so the rotation costs
        // no register moves.
        constexpr int NS = PD + 1;
        float q[NF][RPT][9];
        float qn[PD][NF][RPT];
        float hx[PD][NF][RPT];
        float hy[PD][NF][HY];
        float so[NS][RPT], sv[NS][RPT], sro[NS][RPT], sim[NS][RPT], sil[NS][RPT];
        // results of the plane just computed; img_r/ill_r: image and
        // illumination with this group's term added (group 0: the running
        // sums, completed with the other groups' terms one plane later)
        float res[NF][RPT], img_r[RPT], ill_r[RPT];
        // IMG = 2, one group: this step's terms, added after the previous
        // step's (batched launches keep them in LDS instead)
        float img_c[RPT], ill_c[RPT];

        // halos of plane zz into halo set H and its streamed per-cell values
        // into stream set S (planes at or past nz read 0: out of range)
        auto load_plane = [&](int zz, auto sset, auto hset) {
            constexpr int S = decltype(sset)::value, H = decltype(hset)::value;
            float(&oo)[RPT] = so[S];
            float(&vv)[RPT] = sv[S];
            float(&rr)[RPT] = sro[S];
            float(&ii)[RPT] = sim[S];
            float(&ll)[RPT] = sil[S];
#pragma unroll
            for (int f = 0; f < NF; ++f) {
#pragma unroll
                for (int r = 0; r < RPT; ++r) hx[H][f][r] = fb[f].ld(hxo[r], zz);
#pragma unroll
                for (int m = 0; m < HY; ++m)  // wave-uniform: waves holding y-halo rows
                    if (ty + m * BY < 8) hy[H][f][m] = fb[f].ld(hyo[m], zz);
            }
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                oo[r] = ob.template ld<kNT>(co[r], zz);
                vv[r] = vb.template ld<kNT>(co[r], zz);
                if constexpr (BWD) rr[r] = rb.template ld<kNT>(co[r], zz);
            }
            if constexpr (BWD) {
                if (owner) {
#pragma unroll
                    for (int r = 0; r < RPT; ++r) {
                        if constexpr (RMW_IMG) ii[r] = ib.template ld<kNT>(co[r], zz);
                        if constexpr (RMW_ILL) ll[r] = lb.template ld<kNT>(co[r], zz);
                    }
                }
            }
        };
        auto stage_plane = [&](int buf, int qi, auto hset) {
            // Centre values (queue slot qi) and the halos of set H into LDS
            // buffer buf.  x halos without a per-lane branch (a lane with no
            // halo cell rewrites its own centre value); y-halo rows belong
            // to whole waves, so that test is wave-uniform.
            constexpr int H = decltype(hset)::value;
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                float(*T)[LW] = lds[buf][g * NF + f];
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
                    const float cv = q[f][r][qi];
                    T[ty + r * BY + 4][tx + 4] = cv;
                    T[ty + r * BY + 4][hx_lane ? hxc : tx + 4] = hx_lane ? hx[H][f][r] : cv;
                }
#pragma unroll
                for (int m = 0; m < HY; ++m)
                    if (ty + m * BY < 8) T[hy_row[m]][tx + 4] = hy[H][f][m];
            }
        };
        // group 0: add the other groups' terms of plane zz (parity buffer
        // xb), in shot order, and store image and illumination
        auto finish_plane = [&](int zz, int xb) {
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                float ill = ill_r[r], img = img_r[r];
                const int row = ty + r * BY;
                if constexpr (IMG == 2) {
                    // previous step: the other groups' terms; this step:
                    // group 0's, then the other groups' (shot order)
#pragma unroll
                    for (int k = 0; k < SB - 1; ++k) ill = ill + xch[xb][4 * k][row][tx];
                    ill = ill + (SB > 1 ? xch[xb][NX - 2][row][tx] : ill_c[r]);
#pragma unroll
                    for (int k = 0; k < SB - 1; ++k) ill = ill + xch[xb][4 * k + 2][row][tx];
                    if constexpr (MODE == BWD_RTM) {
#pragma unroll
                        for (int k = 0; k < SB - 1; ++k) img = img + xch[xb][4 * k + 1][row][tx];
                        img = img + (SB > 1 ? xch[xb][NX - 1][row][tx] : img_c[r]);
                    }
#pragma unroll
                    for (int k = 0; k < SB - 1; ++k) img = img + xch[xb][4 * k + 3][row][tx];
                } else if constexpr (IMG >= 3 && SB == 1) {
                    // one group: this step's deferred terms, after the earlier ones
                    if constexpr (RMW_ILL) ill = ill + ill_c[r];
                    if constexpr (MODE == BWD_RTM && RMW_IMG) img = img + img_c[r];
                } else if constexpr (IMG >= 3) {
                    // step by step (i-2, i-1, i), each in shot order; group
                    // 0's term of step i-2 (i-1 for the image) is in ill_r
                    if constexpr (RMW_ILL) {
#pragma unroll
                        for (int k = 0; k < SB - 1; ++k) ill = ill + xch[xb][5 * k][row][tx];
                        ill = ill + xch[xb][NX - 3][row][tx];
#pragma unroll
                        for (int k = 0; k < SB - 1; ++k) ill = ill + xch[xb][5 * k + 1][row][tx];
                        ill = ill + xch[xb][NX - 2][row][tx];
#pragma unroll
                        for (int k = 0; k < SB - 1; ++k) ill = ill + xch[xb][5 * k + 2][row][tx];
                    }
                    if constexpr (MODE == BWD_RTM) {
                        if constexpr (RMW_IMG) {
#pragma unroll
                            for (int k = 0; k < SB - 1; ++k) img = img + xch[xb][5 * k + 3][row][tx];
                            img = img + xch[xb][NX - 1][row][tx];
#pragma unroll
                            for (int k = 0; k < SB - 1; ++k) img = img + xch[xb][5 * k + 4][row][tx];
                        }
                    } else {
#pragma unroll
                        for (int k = 0; k < SB - 1; ++k) img = img + xch[xb][5 * k + 4][row][tx];
                    }
                } else if constexpr (SB > 1) {
#pragma unroll
                    for (int k = 0; k < SB - 1; ++k) {
                        if constexpr (RMW_ILL) ill = ill + xch[xb][2 * k][row][tx];
                        if constexpr (RMW_IMG) img = img + xch[xb][2 * k + 1][row][tx];
                    }
                }
                if constexpr (RMW_ILL) lb.template st<kNT>(co[r], zz, ill);
                if constexpr (RMW_IMG) ib.template st<kNT>(co[r], zz, img);
            }
        };

        // prologue: planes zb-4 .. zb+4 into the queue, plane zb into LDS
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int d = 0; d < 9; ++d)
#pragma unroll
                for (int r = 0; r < RPT; ++r) q[f][r][d] = fb[f].ld(qo[r], zb - 4 + d);
        load_plane(zb, Parity<0>{}, Parity<0>{});
        if constexpr (PD == 2) {
            // plane zb+1 in flight: its queue head (zb+5, consumed at the
            // end of plane zb, from head set 1) and its halos and streams
            const int ph = min(zb + 5, L.nz - 1);
#pragma unroll
            for (int f = 0; f < NF; ++f)
#pragma unroll
                for (int r = 0; r < RPT; ++r) qn[1][f][r] = fb[f].ld(qo[r], ph);
            load_plane(zb + 1, Parity<1>{}, Parity<1>{});
        }
        stage_plane(0, 4, Parity<0>{});
        __syncthreads();

        // One plane.  The phase is a compile-time argument (Parity<P>).
        // PD = 1: P = (z - zb) & 1 selects the LDS buffer and which stream
        // set is current, so no moves or address updates are needed between
        // planes; blocks of 1024 threads with several shot groups, and
        // groups of four rows per thread, keep a runtime buffer index and
        // move the next stream set into place (P = 2): the doubled loop body
        // would not fit their register budget.  PD = 2: P = (z - zb) mod 6;
        // plane z's streams are set P mod 3, plane z+2's loads go to stream
        // set (P+2) mod 3, halo set P mod 2 and head set P mod 2, and the
        // LDS buffer, the halos staged at the end (plane z+1) and the queue
        // head consumed there (plane z+5) follow from the phase as well.
        auto plane = [&](auto par, const int z, const int dyn_buf) {
            constexpr int P = decltype(par)::value;
            constexpr int SC = PD == 1 ? (P == 1 ? 1 : 0) : P % 3;      // streams of plane z
            constexpr int SN = PD == 1 ? (P == 1 ? 0 : 1) : (P + 2) % 3;  // streams loaded now
            constexpr int HL = PD == 1 ? 0 : P % 2;                     // halo / head set loaded now
            constexpr int HS = PD == 1 ? 0 : (P + 1) % 2;               // halo / head set used at the end
            const int buf = (PD == 1 && P == 2) ? dyn_buf : P % 2;
            float(&o)[RPT] = so[SC];
            float(&v)[RPT] = sv[SC];
            float(&ro)[RPT] = sro[SC];
            float(&im)[RPT] = sim[SC];
            float(&il)[RPT] = sil[SC];
            // 1. prefetch plane z+PD: queue head z+4+PD, halos, streams.  The
            // loop body is one straight block (no branch on the last planes,
            // whose prefetched values are simply unused; the queue head is
            // clamped to the last plane of the volume, and planes past it
            // read 0).
            {
                const int ph = min(z + 4 + PD, L.nz - 1);
#pragma unroll
                for (int f = 0; f < NF; ++f)
#pragma unroll
                    for (int r = 0; r < RPT; ++r) qn[HL][f][r] = fb[f].ld(qo[r], ph);
                load_plane(z + PD, Parity<SN>{}, Parity<HL>{});
            }

            // batched backward: complete plane z-1 (its terms were
            // exchanged before the barrier that ended the previous plane)
            if constexpr (BWD && SB > 1) {
                if (owner && z > zb) finish_plane(z - 1, buf ^ 1);
            }

            // 2. compute plane z (results kept in registers)
            DampAxis dmp_z{0.0f, false};
            bool damp = false;  // block-uniform
            if constexpr (ABC) {
                dmp_z = damp_axis(a.b1d, a.nb, z, L.nz);
                damp = tile_damped || dmp_z.on;
            }
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const int lr = ty + r * BY + 4;
                const int lc = tx + 4;
                const float(*P)[LW] = lds[buf][g * NF];
                const float c = q[0][r][4];
                const float dxx = d2<ARITH>(c, P[lr][lc - 1], P[lr][lc + 1], P[lr][lc - 2], P[lr][lc + 2],
                                            P[lr][lc - 3], P[lr][lc + 3], P[lr][lc - 4], P[lr][lc + 4], a.idh2);
                const float dyy = d2<ARITH>(c, P[lr - 1][lc], P[lr + 1][lc], P[lr - 2][lc], P[lr + 2][lc],
                                            P[lr - 3][lc], P[lr + 3][lc], P[lr - 4][lc], P[lr + 4][lc], a.idh2);
                const float* qz = q[0][r];
                const float dzz = d2<ARITH>(c, qz[3], qz[5], qz[2], qz[6], qz[1], qz[7], qz[0], qz[8], a.idh2);
                const float vel = v[r];
                const float vp2 = vel * vel;
                const float K = a.dt2 * vp2;
                if constexpr (!BWD) {
                    float oo = o[r];
                    float dmp = 1.0f;
                    if constexpr (ABC) {
                        if (damp) {
                            dmp = dmp_z.on ? damper_zon(dmp_z.b, dmp_x.b, dmp_xy[r]) : dmp_xy[r].dxy;
                            oo = oo * dmp;
                        }
                    }
                    float nv = K * ((dxx + dyy) + dzz);
                    nv = nv + 2.0f * c;
                    nv = nv - oo;
                    if constexpr (ABC) {
                        if (damp) nv = nv * dmp;
                    }
                    if (z == src_z && cell[r] == src_c) nv = nv + a.inj;
                    res[0][r] = nv;
                } else {
                    const float(*R)[LW] = lds[buf][g * NF + 1];
                    const float cr = q[1][r][4];
                    const float rxx = d2<ARITH>(cr, R[lr][lc - 1], R[lr][lc + 1], R[lr][lc - 2], R[lr][lc + 2],
                                                R[lr][lc - 3], R[lr][lc + 3], R[lr][lc - 4], R[lr][lc + 4], a.idh2);
                    const float ryy = d2<ARITH>(cr, R[lr - 1][lc], R[lr + 1][lc], R[lr - 2][lc], R[lr + 2][lc],
                                                R[lr - 3][lc], R[lr + 3][lc], R[lr - 4][lc], R[lr + 4][lc], a.idh2);
                    const float* rz = q[1][r];
                    const float rzz = d2<ARITH>(cr, rz[3], rz[5], rz[2], rz[6], rz[1], rz[7], rz[0], rz[8], a.idh2);
                    float sn = K * ((dxx + dyy) + dzz);
                    sn = sn + 2.0f * c;
                    sn = sn - o[r];
                    if (CORR && (a.phys & PHYS_SUBTRACT_SOURCE) && z == src_z && cell[r] == src_c) sn = sn + a.inj;
                    float rn = K * ((rxx + ryy) + rzz);
                    rn = rn + 2.0f * cr;
                    rn = rn - ro[r];
                    res[0][r] = sn;
                    res[1][r] = rn;
                    // source field paired with R: the reconstructed one (one
                    // step behind R, the reference) or the current one
                    const float si = (CORR && (a.phys & PHYS_ALIGNED_IMAGING)) ? c : sn;
                    const float t_ill = si * si;
                    float t_img;
                    if constexpr (MODE == BWD_RTM) {
                        t_img = si * cr;
                    } else {
                        float gterm = a.dt * cr;
                        gterm = gterm * ((CORR && (a.phys & PHYS_GRAD_DYY)) ? (dxx + dyy) + dzz : dxx + dzz);
                        t_img = gterm * vp2;
                    }
                    const int row = ty + r * BY;
                    if constexpr (IMG == 1) {
                        if (owner) {
                            ill_r[r] = il[r] + t_ill;
                            img_r[r] = im[r] + t_img;
                        } else if constexpr (SB > 1) {
                            xch[buf][2 * (g - 1)][row][tx] = t_ill;
                            xch[buf][2 * (g - 1) + 1][row][tx] = t_img;
                        }
                    } else if constexpr (IMG == 0) {
                        if constexpr (MODE == BWD_FWI) {  // the gradient term is not deferred
                            if (owner)
                                img_r[r] = im[r] + t_img;
                            else if constexpr (SB > 1)
                                xch[buf][2 * (g - 1) + 1][row][tx] = t_img;
                        }
                    } else if constexpr (IMG >= 3) {
                        // the illumination terms of the two previous steps
                        // (new source fields o and c), the image term of the
                        // previous one (c with the receiver field it read,
                        // ro); the later terms are added after them in
                        // finish_plane.  Not with aligned imaging, whose
                        // term of two steps back is no longer held (the host
                        // launches these instances only without it).
                        const float p2 = o[r] * o[r], p1 = c * c;
                        if (owner) {
                            if constexpr (RMW_ILL) {
                                if constexpr (SB == 1) {
                                    ill_r[r] = (il[r] + p2) + p1;
                                    ill_c[r] = t_ill;
                                } else {
                                    ill_r[r] = il[r] + p2;
                                    xch[buf][NX - 3][row][tx] = p1;
                                    xch[buf][NX - 2][row][tx] = t_ill;
                                }
                            }
                            if constexpr (MODE == BWD_RTM) {
                                if constexpr (RMW_IMG) {
                                    img_r[r] = im[r] + c * ro[r];
                                    if constexpr (SB == 1)
                                        img_c[r] = t_img;
                                    else
                                        xch[buf][NX - 1][row][tx] = t_img;
                                }
                            } else {
                                img_r[r] = im[r] + t_img;
                            }
                        } else if constexpr (SB > 1) {
                            const int k = 5 * (g - 1);
                            if constexpr (RMW_ILL) {
                                xch[buf][k][row][tx] = p2;
                                xch[buf][k + 1][row][tx] = p1;
                                xch[buf][k + 2][row][tx] = t_ill;
                            }
                            if constexpr (MODE == BWD_RTM) {
                                if constexpr (RMW_IMG) {
                                    xch[buf][k + 3][row][tx] = c * ro[r];
                                    xch[buf][k + 4][row][tx] = t_img;
                                }
                            } else {
                                xch[buf][k + 4][row][tx] = t_img;
                            }
                        }
                    } else {
                        // the previous step's terms: its new source field c
                        // (aligned imaging: its current one, this launch's old
                        // value o) with the receiver field it read (ro)
                        const float ps = (CORR && (a.phys & PHYS_ALIGNED_IMAGING)) ? o[r] : c;
                        const float p_ill = ps * ps;
                        if (owner) {
                            ill_r[r] = il[r] + p_ill;
                            if constexpr (SB > 1)
                                xch[buf][NX - 2][row][tx] = t_ill;
                            else
                                ill_c[r] = t_ill;
                            if constexpr (MODE == BWD_RTM) {
                                img_r[r] = im[r] + ps * ro[r];
                                if constexpr (SB > 1)
                                    xch[buf][NX - 1][row][tx] = t_img;
                                else
                                    img_c[r] = t_img;
                            } else {
                                img_r[r] = im[r] + t_img;
                            }
                        } else if constexpr (SB > 1) {
                            xch[buf][4 * (g - 1)][row][tx] = p_ill;
                            xch[buf][4 * (g - 1) + 2][row][tx] = t_ill;
                            if constexpr (MODE == BWD_RTM) xch[buf][4 * (g - 1) + 1][row][tx] = ps * ro[r];
                            xch[buf][4 * (g - 1) + 3][row][tx] = t_img;
                        }
                    }
                }
            }

            // Keep the compute above this point: staging waits for the
            // prefetched loads, which must stay in flight during the compute.
            __builtin_amdgcn_sched_barrier(0);

            // 3. store plane z; stage plane z+1 into the other LDS buffer
            // (the staging waits only for the prefetched loads), advance the queue
            {
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
                    ob.template st<kNT>(co[r], z, res[0][r]);
                    if constexpr (BWD) rb.template st<kNT>(co[r], z, res[1][r]);
                }
                if constexpr (BWD && SB == 1) finish_plane(z, 0);
            }
            stage_plane(buf ^ 1, 5, Parity<HS>{});
#pragma unroll
            for (int f = 0; f < NF; ++f)
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
#pragma unroll
                    for (int d = 0; d < 8; ++d) q[f][r][d] = q[f][r][d + 1];
                    q[f][r][8] = qn[HS][f][r];
                }
            __syncthreads();
        };
        if constexpr (PD == 2) {
            // six phases per round; the loop leaves after the chunk's last
            // plane (a block-uniform test)
            for (int z = zb;; z += 6) {
                plane(Parity<0>{}, z, 0);
                if (z + 1 >= ze) break;
                plane(Parity<1>{}, z + 1, 0);
                if (z + 2 >= ze) break;
                plane(Parity<2>{}, z + 2, 0);
                if (z + 3 >= ze) break;
                plane(Parity<3>{}, z + 3, 0);
                if (z + 4 >= ze) break;
                plane(Parity<4>{}, z + 4, 0);
                if (z + 5 >= ze) break;
                plane(Parity<5>{}, z + 5, 0);
                if (z + 6 >= ze) break;
            }
        } else if constexpr (SB == 1 || (BY * SB <= 8 && RPT <= 2)) {
            int z = zb;
            for (; z + 1 < ze; z += 2) {
                plane(Parity<0>{}, z, 0);
                plane(Parity<1>{}, z + 1, 1);
            }
            if (z < ze) plane(Parity<0>{}, z, 0);
        } else {
            for (int z = zb; z < ze; ++z) {
                plane(Parity<2>{}, z, (z - zb) & 1);
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
                    so[0][r] = so[1][r];
                    sv[0][r] = sv[1][r];
                    if constexpr (BWD) {
                        sro[0][r] = sro[1][r];
                        sim[0][r] = sim[1][r];
                        sil[0][r] = sil[1][r];
                    }
                }
            }
        }
        // batched backward: complete the last plane (its terms were
        // exchanged before the loop's last barrier)
        if constexpr (BWD && SB > 1) {
            if (owner) finish_plane(ze - 1, (ze - 1 - zb) & 1);
        }
    }
}

template <int MODE, int ARITH, int BY, int RPT, int SB, bool CORR, int IMG, int PD = 1, bool BIG = false>
hipError_t launch_one(const StepArgs& a, dim3 grid, hipStream_t st) {
    step_kernel<MODE, ARITH, BY, RPT, SB, CORR, IMG, PD, BIG><<<grid, dim3(64, BY * SB), 0, st>>>(a);
    return hipGetLastError();
}

}  // namespace

using LaunchFn = hipError_t (*)(const StepArgs&, dim3, hipStream_t);

struct Entry {
    int mode, arith, by, rpt, sb;
    int pd;            // prefetch depth (planes in flight)
    int big;           // per-plane addressing (volumes of 2 GiB and more)
    LaunchFn fn;       // per-step imaging, no corrections
    LaunchFn fn_corr;  // with the backward corrections (backward modes)
    LaunchFn fn_img0;  // deferred imaging: none this launch
    LaunchFn fn_img2;  // deferred imaging: previous step's and this step's terms
    LaunchFn fn_corr_img0, fn_corr_img2;  // the same with the corrections (one shot group per block)
    // image every other launch, illumination every third (IMG 3: image,
    // 4: illumination, 5: both; 4 also for FWI)
    LaunchFn fn_img3, fn_img4, fn_img5;
    LaunchFn fn_corr_img3, fn_corr_img4, fn_corr_img5;  // with the corrections (not aligned imaging)
    int spt = 1;       // shots per thread (2: pair_kernel, sb = 2 shots per block)
};

constexpr bool is_bwd(int m) { return m == BWD_RTM || m == BWD_FWI; }

// backward instances: with the corrections, the two deferred-imaging ones
// (RTM, and FWI in blocks of up to 512 threads: in larger batched blocks
// the FWI variants would spill registers), and those two with the
// corrections for one shot group per block; other entries name the plain
// instance there (never chosen)
constexpr bool defers(int m, int by, int sb) {
    return m == BWD_RTM || (m == BWD_FWI && (sb == 1 || by * sb <= 8));
}
constexpr bool defers_corr(int m, int sb) { return is_bwd(m) && sb == 1; }
// image every other launch, illumination every third: wherever deferred
// imaging is, except in batched blocks of 1024 threads with two or more
// rows per thread, whose instances would spill registers (those keep the
// two-launch pattern)
constexpr bool cad3_fits(int by, int rpt, int sb) { return sb == 1 || !(by * sb == 16 && rpt >= 2); }
constexpr bool cad3_ill(int m, int by, int rpt, int sb) {
    return is_bwd(m) && defers(m, by, sb) && cad3_fits(by, rpt, sb);
}
constexpr bool cad3_img(int m, int by, int rpt, int sb) { return m == BWD_RTM && cad3_fits(by, rpt, sb); }
constexpr bool cad3_img_corr(int m, int sb) { return m == BWD_RTM && defers_corr(m, sb); }
#define FWI_E(M, A, BY, RPT, SB)                                                                      \
    {M, A, BY, RPT, SB, 1, 0, &launch_one<M, A, BY, RPT, SB, false, 1>,                                \
     is_bwd(M) ? &launch_one<M, A, BY, RPT, SB, is_bwd(M), 1> : nullptr,                               \
     defers(M, BY, SB) ? &launch_one<M, A, BY, RPT, SB, false, defers(M, BY, SB) ? 0 : 1> : nullptr,   \
     defers(M, BY, SB) ? &launch_one<M, A, BY, RPT, SB, false, defers(M, BY, SB) ? 2 : 1> : nullptr,   \
     defers_corr(M, SB) ? &launch_one<M, A, BY, RPT, SB, defers_corr(M, SB), defers_corr(M, SB) ? 0 : 1> \
                        : nullptr,                                                                     \
     defers_corr(M, SB) ? &launch_one<M, A, BY, RPT, SB, defers_corr(M, SB), defers_corr(M, SB) ? 2 : 1> \
                        : nullptr,                                                                     \
     cad3_img(M, BY, RPT, SB) ? &launch_one<M, A, BY, RPT, SB, false, cad3_img(M, BY, RPT, SB) ? 3 : 1>     \
                              : nullptr,                                                              \
     cad3_ill(M, BY, RPT, SB) ? &launch_one<M, A, BY, RPT, SB, false, cad3_ill(M, BY, RPT, SB) ? 4 : 1>     \
                              : nullptr,                                                              \
     cad3_img(M, BY, RPT, SB) ? &launch_one<M, A, BY, RPT, SB, false, cad3_img(M, BY, RPT, SB) ? 5 : 1>     \
                              : nullptr,                                                              \
     cad3_img_corr(M, SB) ? &launch_one<M, A, BY, RPT, SB, cad3_img_corr(M, SB), cad3_img_corr(M, SB) ? 3 : 1> \
                          : nullptr,                                                                  \
     defers_corr(M, SB) ? &launch_one<M, A, BY, RPT, SB, defers_corr(M, SB), defers_corr(M, SB) ? 4 : 1>     \
                        : nullptr,                                                                    \
     cad3_img_corr(M, SB) ? &launch_one<M, A, BY, RPT, SB, cad3_img_corr(M, SB), cad3_img_corr(M, SB) ? 5 : 1> \
                          : nullptr}
#define FWI_TILES(M, A, SB)                                                                              \
    FWI_E(M, A, 4, 1, SB), FWI_E(M, A, 4, 2, SB), FWI_E(M, A, 8, 1, SB), FWI_E(M, A, 8, 2, SB),         \
        FWI_E(M, A, 4, 4, SB), FWI_E(M, A, 16, 1, SB), FWI_E(M, A, 16, 2, SB)

// two planes in flight (one shot per block): the per-step and deferred
// imaging instances without the corrections (a launch with corrections
// takes the one-plane instance of the same shape)
#define FWI_E2(M, A, BY, RPT)                                                                            \
    {M, A, BY, RPT, 1, 2, 0, &launch_one<M, A, BY, RPT, 1, false, 1, 2>, nullptr,                         \
     is_bwd(M) ? &launch_one<M, A, BY, RPT, 1, false, is_bwd(M) ? 0 : 1, 2> : nullptr,                   \
     is_bwd(M) ? &launch_one<M, A, BY, RPT, 1, false, is_bwd(M) ? 2 : 1, 2> : nullptr, nullptr, nullptr,  \
     M == BWD_RTM ? &launch_one<M, A, BY, RPT, 1, false, M == BWD_RTM ? 3 : 1, 2> : nullptr,             \
     is_bwd(M) ? &launch_one<M, A, BY, RPT, 1, false, is_bwd(M) ? 4 : 1, 2> : nullptr,                   \
     M == BWD_RTM ? &launch_one<M, A, BY, RPT, 1, false, M == BWD_RTM ? 5 : 1, 2> : nullptr,             \
     nullptr, nullptr, nullptr}
// (shapes whose instances fit 128 registers without spilling)
#define FWI_TILES2(M, A) FWI_E2(M, A, 16, 1), FWI_E2(M, A, 8, 1)

// volumes of 2 GiB and more: per-plane addressing, one shot per block,
// image and illumination updated every step (with and without corrections)
#define FWI_EB(M, A, BY, RPT)                                                                            \
    {M, A, BY, RPT, 1, 1, 1, &launch_one<M, A, BY, RPT, 1, false, 1, 1, true>,                            \
     is_bwd(M) ? &launch_one<M, A, BY, RPT, 1, is_bwd(M), 1, 1, true> : nullptr, nullptr, nullptr, nullptr,  \
     nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr}
#define FWI_TILESB(M, A) FWI_EB(M, A, 16, 1), FWI_EB(M, A, 8, 1), FWI_EB(M, A, 4, 2)

// batched: SB wave groups per block, one shot each (BY * SB <= 16 waves)
#define FWI_BATCH(M)                                                                                     \
    FWI_E(M, ARITH_REF, 4, 1, 2), FWI_E(M, ARITH_REF, 4, 2, 2), FWI_E(M, ARITH_REF, 8, 1, 2),            \
        FWI_E(M, ARITH_REF, 8, 2, 2), FWI_E(M, ARITH_REF, 2, 2, 4), FWI_E(M, ARITH_REF, 4, 1, 4),         \
        FWI_E(M, ARITH_REF, 4, 2, 4), FWI_E(M, ARITH_REF, 2, 4, 4)

}  // namespace fwi

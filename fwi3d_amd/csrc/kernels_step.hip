// kernels_step.hip -- launch selection of the time-step kernels: picks the
// instance of step_kernel (step_kernel.hpp) for a mode, arithmetic, launch
// shape and imaging cadence from the tables of the step_inst_*.hip units.
#include "step_kernel.hpp"

namespace fwi {

const Entry* step_table_fwd(int* count);
const Entry* step_table_rtm(int* count);
const Entry* step_table_fwi(int* count);
const Entry* step_table_batch(int* count);
const Entry* step_table_prefetch(int* count);
const Entry* step_table_big(int* count);
const Entry* step_table_pair(int* count);

namespace {

const Entry* find(int mode, int arith, const Tuning& t) {
    int counts[7];
    const Entry* tables[] = {step_table_fwd(&counts[0]),      step_table_rtm(&counts[1]),
                             step_table_fwi(&counts[2]),      step_table_batch(&counts[3]),
                             step_table_prefetch(&counts[4]), step_table_big(&counts[5]),
                             step_table_pair(&counts[6])};
    for (int k = 0; k < 7; ++k)
        for (int i = 0; i < counts[k]; ++i) {
            const Entry& e = tables[k][i];
            if (e.mode == mode && e.arith == arith && e.by == t.by && e.rpt == t.rpt && e.sb == t.sb &&
                e.pd == t.pd && e.big == t.big && e.spt == t.spt)
                return &e;
        }
    return nullptr;
}

// The entry a launch uses: the two-plane-prefetch instances carry no
// corrections, so a launch with corrections takes the one-plane instance of
// the same shape.
const Entry* resolve(int mode, int arith, const Tuning& t, bool corrections) {
    const Entry* e = find(mode, arith, t);
    if (e && e->pd != 1 && corrections) {
        Tuning t1 = t;
        t1.pd = 1;
        e = find(mode, arith, t1);
    }
    return e;
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
    const int per_grp = a.ntx * a.nty * a.nzc;
    a.nitems = per_grp * (nshots / sb);
    // the kernels' item order: 8 blocks per tile range of an XCD, per group
    const int blocks = (per_grp + 7) / 8 * 8 * (nshots / sb);
    grid = dim3(blocks);
    block = dim3(64, t.by * sb / t.spt);
}

bool step_defer_supported(int mode, int arith, const Tuning& t, bool corrections) {
    const Entry* e = resolve(mode, arith, t, corrections);
    if (!e) return false;
    return corrections ? (e->fn_corr_img0 && e->fn_corr_img2) : (e->fn_img0 && e->fn_img2);
}

bool step_defer3_supported(int mode, int arith, const Tuning& t, int phys) {
    const Entry* e = resolve(mode, arith, t, phys != 0);
    if (!e || (phys & PHYS_ALIGNED_IMAGING)) return false;
    if (phys)
        return e->fn_corr_img0 && e->fn_corr_img4 && (mode == BWD_FWI || (e->fn_corr_img3 && e->fn_corr_img5));
    return e->fn_img0 && e->fn_img4 && (mode == BWD_FWI || (e->fn_img3 && e->fn_img5));
}

hipError_t launch_step(int mode, int arith, const Tuning& t, const StepArgs& a, hipStream_t st) {
    const bool corr = is_bwd(mode) && a.phys != 0;
    const Entry* e = resolve(mode, arith, t, corr);
    if (!e || a.nshots % t.sb != 0) return hipErrorInvalidConfiguration;
    if (a.img != 1 && !is_bwd(mode)) return hipErrorInvalidConfiguration;
    if (a.img < 0 || a.img > 5) return hipErrorInvalidConfiguration;
    if (a.img >= 3 && (a.phys & PHYS_ALIGNED_IMAGING)) return hipErrorInvalidConfiguration;
    LaunchFn fn = nullptr;
    switch (a.img) {
        case 0: fn = corr ? e->fn_corr_img0 : e->fn_img0; break;
        case 1: fn = corr ? e->fn_corr : e->fn; break;
        case 2: fn = corr ? e->fn_corr_img2 : e->fn_img2; break;
        case 3: fn = corr ? e->fn_corr_img3 : e->fn_img3; break;
        case 4: fn = corr ? e->fn_corr_img4 : e->fn_img4; break;
        default: fn = corr ? e->fn_corr_img5 : e->fn_img5; break;
    }
    if (!fn) return hipErrorInvalidConfiguration;
    StepArgs args = a;
    dim3 grid, block;
    step_grid(a.L, t, a.nshots, t.sb, args, grid, block);
    return fn(args, grid, st);
}

}  // namespace fwi

// abi.hip -- the C ABI of libfwi3d_hip.so (declared in include/fwi3d.h).
//
// A context owns one HIP stream and the device buffers of `nshots` shot
// slots: per-slot padded model (each shot has its own random boundary),
// source-side and receiver-side fields, the flattened receiver lists of all
// slots, the time-major seismogram and backward data, and the image /
// illumination / gradient accumulators (one partial per group of shots that
// share a block; folded into group 0 after a pass).  The time loops of the
// reference's forward_solver (modeling.cu:309-322) and backward_propagation
// (migration.cu:61-81, inversion.cu:235-253) are issued here: one step
// kernel per time step advances every slot, optionally captured in a
// hipGraph.  A single-shot context is the nshots = 1 case; the entry points
// without a slot argument address slot 0.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/fwi3d.h"
#include "common.hpp"

using fwi::Layout;
using fwi::StepArgs;
using fwi::Step2Args;
using fwi::Tuning;

struct fwi_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    fwi_grid g{};
    Layout L{};
    int nshots = 1;
    int arith = FWI_ARITH_REFERENCE;
    int phys = 0;                        // FWI_PHYS_* bits of the backward passes
    bool defer_imaging = true;           // image/illumination updated every second or third backward launch
    float idh2 = 0, idh3 = 0, dt2 = 0;

    float* vp = nullptr;                 // nshots volumes
    // source-side fields: f[fcur] = d_P (current), f[fold] = d_Pold; the
    // two spare buffers receive the output of two-step launches
    float* f[4] = {nullptr, nullptr, nullptr, nullptr};
    int fcur = 0, fold = 1;
    // receiver-side fields: r[rcur] = d_Pr, r[rold] = d_Prold, likewise
    float* r[4] = {nullptr, nullptr, nullptr, nullptr};
    int rcur = 0, rold = 1;
    float* image = nullptr;              // nshots partial volumes each; group 0 = result
    float* illum = nullptr;
    float* grad = nullptr;
    float* staging = nullptr;            // one padded volume, reference layout

    std::vector<float> wavelet;
    float* b1d = nullptr;
    std::vector<float> b1d_h;
    bool have_model = false;

    std::vector<int> sx, sy, sz;         // per slot, -1 = unset
    long* src_off = nullptr;             // device [nshots], offsets within a slot
    std::vector<std::vector<int>> rx, ry, rz;
    bool geom_dirty = true;
    std::vector<int> rec_start_h;        // nshots + 1
    int nrec_total = 0;
    long* rec_off = nullptr;             // absolute offsets (slot*vol + cell)
    float* rec_damp = nullptr;
    int* rec_start = nullptr;
    int ncells = 0;
    long* cell_off = nullptr;
    int* cell_start = nullptr;
    int* cell_rec = nullptr;
    // receiver planes of the two-step launches: rp_z [nshots][kMaxRecPlanes]
    // (-1: none) and per plane a map [ny*px] of receiver cell index or -1;
    // rp_ok is false when a slot's receivers lie on more planes than that
    int* rp_z = nullptr;
    int* rp_map = nullptr;
    bool rp_ok = false;
    std::vector<long> cell_off_h;        // host copy of cell_off
    // the 4-cell rims of the source-side / receiver-side field buffers are
    // zero (never updated; two-step launches rely on it, since their output
    // buffers rotate through four)
    bool f_rims_zero = true, r_rims_zero = true;
    float* seis_tm = nullptr;            // [t][q] over all slots' receivers
    float* data_tm = nullptr;            // [t][q] backward data
    float* seis_tmp = nullptr;           // 2 * nt * nrec_total scratch for transposes
    double* red_partial = nullptr;
    double* red_result = nullptr;

    Tuning tune[4];
    bool use_graph = false;
    // Instantiated pass graphs, replayed while nothing their launches
    // captured has changed: every setter that changes a launch argument
    // (wavelet, tuning, physics, geometry, buffers) advances `gen`.
    struct PassGraph {
        std::string key;
        uint64_t gen;
        hipGraphExec_t exec;
        int fcur_after, fold_after, rcur_after, rold_after;
    };
    std::vector<PassGraph> graphs;
    uint64_t gen = 0;
    int graph_instantiations = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    int last_launches = 0;
    int fwd_done = 0;                    // 1 after a full forward pass
    bool ktime = false;                  // per-step-kernel events
    std::vector<hipEvent_t> kev;
    int kev_used = 0;
    double kernel_ms = 0.0;
    std::string err;
};

namespace {

thread_local std::string g_create_error;

int fail(fwi_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                                 \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return fail(ctx, FWI_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));             \
    } while (0)

#define CHKRC(expr)             \
    do {                        \
        int rc_ = (expr);       \
        if (rc_) return rc_;    \
    } while (0)

long nvol_host(const fwi_grid& g) { return (long)g.nxx * g.nyy * g.nzz; }
long ninterior(const fwi_grid& g) {
    return (long)(g.nxx - 2 * g.nb) * (g.nyy - 2 * g.nb) * (g.nzz - 2 * g.nb);
}

// (a buffer allocated here keeps its address until the context is destroyed,
// and a pass allocates what it uses before it is captured, so allocations do
// not invalidate captured pass graphs)
int alloc_zero(fwi_ctx* c, float** p, long nfloats) {
    if (*p) return FWI_OK;
    HIPCHK(c, hipMalloc((void**)p, nfloats * sizeof(float)));
    HIPCHK(c, hipMemsetAsync(*p, 0, nfloats * sizeof(float), c->st));
    return FWI_OK;
}

int alloc_slots(fwi_ctx* c, float** p) { return alloc_zero(c, p, c->L.vol * c->nshots); }

// Host twin of the device damper (same float operations).
float host_damper(const std::vector<float>& b1, int nb, int x, int y, int z, const Layout& L) {
    auto axis = [&](int q, int n, float* b) {
        const int d = q < nb ? q : (q >= n - nb ? n - 1 - q : -1);
        *b = d >= 0 ? b1[d] : 0.0f;
        return d >= 0;
    };
    float bi, bj, bk;
    const bool oi = axis(z, L.nz, &bi), oj = axis(x, L.nx, &bj), ok = axis(y, L.ny, &bk);
    const int cnt = (int)oi + (int)oj + (int)ok;
    if (cnt == 0) return 1.0f;
    if (cnt == 1) return oi ? bi : (oj ? bj : bk);
    if (cnt == 2) {
        const float a = oi ? bi : bj;
        const float b = ok ? bk : bj;
        return (a + b) - 1.0f;
    }
    const float t1 = bi + bj, t2 = bj + bk, t3 = bi + bk;
    return ((t1 + t2) + t3) - 5.0f;
}

bool in_updated(const fwi_ctx* c, int x, int y, int z) {
    return x >= 4 && x < c->g.nxx - 4 && y >= 4 && y < c->g.nyy - 4 && z >= 4 && z < c->g.nzz - 4;
}

int check_slot(fwi_ctx* c, int slot) {
    if (slot < 0 || slot >= c->nshots) return fail(c, FWI_ERR_ARG, "shot slot out of range");
    return FWI_OK;
}

bool two_step_wanted(const fwi_ctx* c) {
    for (const Tuning& t : c->tune)
        if (t.tb == 2) return true;
    return false;
}

// Receiver planes of every slot for the two-step launches, which find the
// receiver cells of a plane through a per-slot map of that plane; coff are
// the receiver cells (absolute offsets) in cell-index order.
int build_rec_planes(fwi_ctx* c, const std::vector<long>& coff) {
    c->rp_ok = false;
    if (!two_step_wanted(c)) return FWI_OK;
    const int S = c->nshots, K = fwi::kMaxRecPlanes;
    const long plane_n = (long)c->L.ny * c->L.px;
    if ((long)S * K * plane_n * (long)sizeof(int) >= (1L << 31)) return FWI_OK;  // offsets of the map buffer
    if (!c->rp_z) HIPCHK(c, hipMalloc((void**)&c->rp_z, (size_t)S * K * sizeof(int)));
    if (!c->rp_map) HIPCHK(c, hipMalloc((void**)&c->rp_map, (size_t)S * K * plane_n * sizeof(int)));
    std::vector<int> z((size_t)S * K, -1);
    std::vector<long> pos(coff.size());
    for (size_t j = 0; j < coff.size(); ++j) {
        const int s = (int)(coff[j] / c->L.vol);
        const long w = coff[j] - (long)s * c->L.vol;
        const int zz = (int)(w / c->L.pp);
        int k = 0;
        while (k < K && z[s * K + k] != -1 && z[s * K + k] != zz) ++k;
        if (k == K) return FWI_OK;  // more receiver planes than a launch looks up
        z[s * K + k] = zz;
        pos[j] = ((long)s * K + k) * plane_n + (w - (long)zz * c->L.pp);
    }
    HIPCHK(c, hipMemcpy(c->rp_z, z.data(), z.size() * sizeof(int), hipMemcpyHostToDevice));
    long* d_pos = nullptr;
    if (!pos.empty()) {
        HIPCHK(c, hipMalloc((void**)&d_pos, pos.size() * sizeof(long)));
        hipError_t e = hipMemcpy(d_pos, pos.data(), pos.size() * sizeof(long), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = fwi::launch_cell_map(c->rp_map, (long)S * K * plane_n, d_pos, (int)pos.size(), c->st);
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);
        (void)hipFree(d_pos);
        if (e != hipSuccess) return fail(c, FWI_ERR_HIP, std::string("receiver-plane maps: ") + hipGetErrorString(e));
    } else {
        HIPCHK(c, fwi::launch_cell_map(c->rp_map, (long)S * K * plane_n, nullptr, 0, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
    }
    c->rp_ok = true;
    return FWI_OK;
}

// (Re)build the flattened receiver arrays of all slots after a geometry
// change; recorded seismograms and uploaded data are cleared.
int sync_geometry(fwi_ctx* c) {
    if (!c->geom_dirty) return FWI_OK;
    c->gen++;
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (void* p : {(void*)c->rec_off, (void*)c->rec_damp, (void*)c->seis_tm, (void*)c->data_tm, (void*)c->seis_tmp,
                    (void*)c->cell_off, (void*)c->cell_start, (void*)c->cell_rec})
        if (p) HIPCHK(c, hipFree(p));
    c->rec_off = nullptr;
    c->rec_damp = nullptr;
    c->seis_tm = c->data_tm = c->seis_tmp = nullptr;
    c->cell_off = nullptr;
    c->cell_start = c->cell_rec = nullptr;
    c->ncells = 0;
    const int S = c->nshots;
    c->rec_start_h.assign(S + 1, 0);
    for (int s = 0; s < S; ++s) c->rec_start_h[s + 1] = c->rec_start_h[s] + (int)c->rx[s].size();
    const int n = c->nrec_total = c->rec_start_h[S];
    HIPCHK(c, hipMemcpy(c->rec_start, c->rec_start_h.data(), (S + 1) * sizeof(int), hipMemcpyHostToDevice));
    c->geom_dirty = false;
    c->cell_off_h.clear();
    if (n == 0) return build_rec_planes(c, c->cell_off_h);
    std::vector<long> off(n);
    std::vector<float> damp(n, 1.0f);
    for (int s = 0; s < S; ++s)
        for (size_t q = 0; q < c->rx[s].size(); ++q) {
            const int k = c->rec_start_h[s] + (int)q;
            off[k] = (long)s * c->L.vol + c->L.at(c->rx[s][q], c->ry[s][q], c->rz[s][q]);
            if (!c->b1d_h.empty()) damp[k] = host_damper(c->b1d_h, c->g.nb, c->rx[s][q], c->ry[s][q], c->rz[s][q], c->L);
        }
    // receivers grouped by cell, ascending receiver order inside a cell
    std::vector<int> order(n);
    for (int q = 0; q < n; ++q) order[q] = q;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return off[a] < off[b]; });
    std::vector<long> coff;
    std::vector<int> cstart;
    for (int q = 0; q < n; ++q)
        if (q == 0 || off[order[q]] != off[order[q - 1]]) {
            coff.push_back(off[order[q]]);
            cstart.push_back(q);
        }
    cstart.push_back(n);
    c->ncells = (int)coff.size();
    const long ns = (long)c->g.nt * n;
    HIPCHK(c, hipMalloc((void**)&c->rec_off, n * sizeof(long)));
    HIPCHK(c, hipMalloc((void**)&c->rec_damp, n * sizeof(float)));
    HIPCHK(c, hipMalloc((void**)&c->cell_off, coff.size() * sizeof(long)));
    HIPCHK(c, hipMalloc((void**)&c->cell_start, cstart.size() * sizeof(int)));
    HIPCHK(c, hipMalloc((void**)&c->cell_rec, n * sizeof(int)));
    HIPCHK(c, hipMalloc((void**)&c->seis_tm, ns * sizeof(float)));
    HIPCHK(c, hipMalloc((void**)&c->data_tm, ns * sizeof(float)));
    HIPCHK(c, hipMalloc((void**)&c->seis_tmp, 2 * ns * sizeof(float)));
    HIPCHK(c, hipMemcpy(c->rec_off, off.data(), n * sizeof(long), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->rec_damp, damp.data(), n * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->cell_off, coff.data(), coff.size() * sizeof(long), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->cell_start, cstart.data(), cstart.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->cell_rec, order.data(), n * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(c->seis_tm, 0, ns * sizeof(float)));
    HIPCHK(c, hipMemset(c->data_tm, 0, ns * sizeof(float)));
    c->cell_off_h = coff;
    return build_rec_planes(c, coff);
}

float* vol_ptr(fwi_ctx* c, int slot, int which, bool alloc, int* rc) {
    *rc = FWI_OK;
    float** p = nullptr;
    bool per_slot = true;
    switch (which) {
        case FWI_VOL_MODEL: p = &c->vp; break;
        case FWI_VOL_FIELD: p = &c->f[c->fcur]; break;
        case FWI_VOL_FIELD_OLD: p = &c->f[c->fold]; break;
        case FWI_VOL_RECV: p = &c->r[c->rcur]; break;
        case FWI_VOL_RECV_OLD: p = &c->r[c->rold]; break;
        case FWI_VOL_IMAGE: p = &c->image; per_slot = false; break;
        case FWI_VOL_ILLUM: p = &c->illum; per_slot = false; break;
        case FWI_VOL_GRAD: p = &c->grad; per_slot = false; break;
        default: *rc = fail(c, FWI_ERR_ARG, "unknown volume id"); return nullptr;
    }
    if (!*p && alloc) {
        *rc = alloc_slots(c, p);
        if (*rc) return nullptr;
    }
    if (!*p) return nullptr;
    return *p + (per_slot ? (long)slot * c->L.vol : 0);
}

int ensure_backward_buffers(fwi_ctx* c, int kind) {
    CHKRC(alloc_slots(c, &c->r[0]));
    CHKRC(alloc_slots(c, &c->r[1]));
    CHKRC(alloc_slots(c, &c->illum));
    CHKRC(alloc_slots(c, kind == 0 ? &c->image : &c->grad));
    if (kind == 0 && c->tune[fwi::BWD_RTM].tb == 2)  // the outputs of two-step launches
        for (float** p : {&c->f[2], &c->f[3], &c->r[2], &c->r[3]}) CHKRC(alloc_slots(c, p));
    return FWI_OK;
}

int kernel_event(fwi_ctx* c, int slot, hipEvent_t* ev) {
    while ((int)c->kev.size() <= slot) {
        hipEvent_t e;
        HIPCHK(c, hipEventCreate(&e));
        c->kev.push_back(e);
    }
    *ev = c->kev[slot];
    return FWI_OK;
}

int launch_step_timed(fwi_ctx* c, int mode, const Tuning& t, const StepArgs& a) {
    if (!c->ktime || c->use_graph) {
        HIPCHK(c, fwi::launch_step(mode, c->arith, t, a, c->st));
        return FWI_OK;
    }
    hipEvent_t e0, e1;
    CHKRC(kernel_event(c, 2 * c->kev_used, &e0));
    CHKRC(kernel_event(c, 2 * c->kev_used + 1, &e1));
    HIPCHK(c, hipEventRecord(e0, c->st));
    HIPCHK(c, fwi::launch_step(mode, c->arith, t, a, c->st));
    HIPCHK(c, hipEventRecord(e1, c->st));
    c->kev_used++;
    return FWI_OK;
}

StepArgs base_args(fwi_ctx* c) {
    StepArgs a{};
    a.img = 1;
    a.L = c->L;
    a.sstride = c->L.vol;
    a.nshots = c->nshots;
    a.vp = c->vp;
    // every slot's model is slot 0's (fwi_set_model / _interior replicate
    // it) except within nb of a face, where the random boundaries write
    a.share_nb = c->g.nb;
    a.istride = c->L.vol;
    a.idh2 = c->idh2;
    a.dt2 = c->dt2;
    a.dt = c->g.dt;
    a.src_off = c->src_off;
    a.b1d = c->b1d;
    a.nb = c->g.nb;
    a.rec_off = c->rec_off;
    a.rec_start = c->rec_start;
    return a;
}

// The one-step tuning's two-step counterpart (tile and z chunks of the
// two-step launches).
Tuning two_step_tuning(const Tuning& t) {
    Tuning u = t;
    u.by = t.tby;
    u.rpt = t.trpt;
    u.zchunks = t.tzc;
    return u;
}

// Two-step launches for this pass: requested, available for the mode,
// arithmetic and tile, one shot per block, no per-plane addressing, the
// receiver planes mapped, and (backward) no corrections.
bool two_step_eligible(const fwi_ctx* c, int mode) {
    const Tuning& t = c->tune[mode];
    if (t.tb != 2 || !c->rp_ok || t.sb != 1 || t.big) return false;
    if (mode == fwi::BWD_RTM && c->phys != 0) return false;
    if (mode != fwi::FWD_RBC && mode != fwi::BWD_RTM) return false;
    if (!c->f_rims_zero || (mode == fwi::BWD_RTM && !c->r_rims_zero)) return false;
    return fwi::step2_supported(mode, c->arith, two_step_tuning(t));
}

// ... and the spare buffers it writes are allocated
bool two_step_active(const fwi_ctx* c, int mode) {
    if (!c->f[2] || !c->f[3] || (mode == fwi::BWD_RTM && (!c->r[2] || !c->r[3]))) return false;
    return two_step_eligible(c, mode);
}

// The two buffer indices (of four) not in {a, b}.
void spare_pair(int a, int b, int* x, int* y) {
    int k = 0, out[2] = {0, 0};
    for (int i = 0; i < 4 && k < 2; ++i)
        if (i != a && i != b) out[k++] = i;
    *x = out[0];
    *y = out[1];
}

Step2Args base_args2(fwi_ctx* c) {
    Step2Args a{};
    a.L = c->L;
    a.sstride = c->L.vol;
    a.nshots = c->nshots;
    a.vp = c->vp;
    a.istride = c->L.vol;
    a.idh2 = c->idh2;
    a.dt2 = c->dt2;
    a.idh3 = c->idh3;
    a.src_off = c->src_off;
    a.rec_off = c->rec_off;
    a.rec_start = c->rec_start;
    a.rp_z = c->rp_z;
    a.rp_map = c->rp_map;
    a.cell_start = c->cell_start;
    a.cell_rec = c->cell_rec;
    return a;
}

int launch_step2_timed(fwi_ctx* c, int mode, const Tuning& t, const Step2Args& a) {
    if (!c->ktime || c->use_graph) {
        HIPCHK(c, fwi::launch_step2(mode, c->arith, t, a, c->st));
        return FWI_OK;
    }
    hipEvent_t e0, e1;
    CHKRC(kernel_event(c, 2 * c->kev_used, &e0));
    CHKRC(kernel_event(c, 2 * c->kev_used + 1, &e1));
    HIPCHK(c, hipEventRecord(e0, c->st));
    HIPCHK(c, fwi::launch_step2(mode, c->arith, t, a, c->st));
    HIPCHK(c, hipEventRecord(e1, c->st));
    c->kev_used++;
    return FWI_OK;
}

// Forward steps [n0, n1) of every slot from the current state: the current
// field excludes the source term of step n0, which is added first; each step
// kernel adds the source term of the next step to the field it writes
// (except after the last step).  record = false skips the receiver gather.
int run_forward(fwi_ctx* c, int boundary, int n0, int n1, bool record = true) {
    const int mode = boundary == FWI_BOUNDARY_ABC ? fwi::FWD_ABC : fwi::FWD_RBC;
    const Tuning& t = c->tune[mode];
    StepArgs a = base_args(c);
    if (boundary == FWI_BOUNDARY_ABC) a.rec_damp = c->rec_damp;
    if (n0 < n1 && n0 < c->g.nt)
        HIPCHK(c, fwi::launch_add_points(c->f[c->fcur], c->src_off, c->nshots, c->L.vol, c->idh3 * c->wavelet[n0],
                                         c->st));
    auto seis_row = [&](int n) -> float* {
        return (record && c->nrec_total > 0 && n > c->g.tlag && n - c->g.tlag < c->g.nt)
                   ? c->seis_tm + (long)(n - c->g.tlag) * c->nrec_total
                   : nullptr;
    };
    auto inj = [&](int n) { return (n < n1 && n < c->g.nt) ? c->idh3 * c->wavelet[n] : 0.0f; };
    int n = n0;
    if (two_step_active(c, mode)) {
        // steps n, n+1 per launch: the step-1 field goes to one spare buffer
        // (the new old field), the step-2 field to the other (the new current)
        const Tuning t2 = two_step_tuning(t);
        Step2Args a2 = base_args2(c);
        for (; n + 1 < n1; n += 2) {
            int w1, w2;
            spare_pair(c->fcur, c->fold, &w1, &w2);
            a2.cur = c->f[c->fcur];
            a2.old = c->f[c->fold];
            a2.out_old = c->f[w1];
            a2.out_cur = c->f[w2];
            a2.inj1 = inj(n + 1);
            a2.inj2 = inj(n + 2);
            a2.seis_row0 = seis_row(n);
            a2.seis_row1 = seis_row(n + 1);
            CHKRC(launch_step2_timed(c, mode, t2, a2));
            c->fold = w1;
            c->fcur = w2;
        }
    }
    for (; n < n1; ++n) {
        a.cur = c->f[c->fcur];
        a.old = c->f[c->fold];
        a.inj = inj(n + 1);
        a.seis_row = seis_row(n);
        CHKRC(launch_step_timed(c, mode, t, a));
        std::swap(c->fcur, c->fold);
    }
    return FWI_OK;
}

// Shot pairs address both shots' volumes through one buffer resource with
// 32-bit offsets: two volumes must stay below 2 GiB.
bool pairs_fit(const fwi_ctx* c) { return 2 * c->L.vol * (long)sizeof(float) < (1L << 31); }

// The tuning of a backward pass: the two-shots-per-thread instances carry no
// corrections, so with corrections a pass of shot pairs runs the
// two-groups-per-block instances (the same summation order) on 64 x 4 tiles.
Tuning backward_tuning(const fwi_ctx* c, int mode) {
    Tuning t = c->tune[mode];
    if (t.spt == 2 && c->phys != 0) {
        t.spt = 1;
        t.by = 4;
        t.rpt = 1;
    }
    return t;
}

// Backward steps [b0, b1) of every slot from the current state (receiver
// data injected before each step b < nt, in receiver order per cell).
int run_backward(fwi_ctx* c, int kind, int b0, int b1) {
    const int mode = kind == 0 ? fwi::BWD_RTM : fwi::BWD_FWI;
    const Tuning t = backward_tuning(c, mode);
    const int nt = c->g.nt;
    StepArgs a = base_args(c);
    a.image = kind == 0 ? c->image : c->grad;
    a.illum = c->illum;
    if (b0 < b1 && b0 < nt)
        HIPCHK(c, fwi::launch_inject(c->r[c->rcur], c->cell_off, c->cell_start, c->cell_rec, c->ncells,
                                     c->data_tm + (long)(nt - 1 - b0) * c->nrec_total, c->idh3, c->st));
    a.phys = c->phys;
    const int nsteps = nt + c->g.tlag;  // steps of the forward pass being reversed
    // deferred imaging: pairs of launches share one image/illumination
    // read-modify-write, and a pass (or segment) of odd length ends with a
    // per-step launch; without aligned imaging the image is updated every
    // other launch and the illumination every third, and the terms still
    // pending at the end are added by one elementwise launch.  Either way no
    // term is left over when the pass ends.
    // two-step launches add both steps' image and illumination terms
    // themselves; a segment of odd length ends with a one-step launch that
    // does the same
    const bool two = two_step_active(c, mode);
    const bool defer = !two && c->defer_imaging && fwi::step_defer_supported(mode, c->arith, t, c->phys != 0);
    const bool defer3 = defer && fwi::step_defer3_supported(mode, c->arith, t, c->phys);
    int b = b0;
    if (two) {
        // steps b, b+1 per launch; the data of step b+1 is injected into the
        // step-1 receiver field inside the launch
        const Tuning t2 = two_step_tuning(t);
        Step2Args a2 = base_args2(c);
        a2.image = c->image;
        a2.illum = c->illum;
        for (; b + 1 < b1; b += 2) {
            int w1, w2, v1, v2;
            spare_pair(c->fcur, c->fold, &w1, &w2);
            spare_pair(c->rcur, c->rold, &v1, &v2);
            a2.cur = c->f[c->fold];
            a2.old = c->f[c->fcur];
            a2.out_old = c->f[w1];
            a2.out_cur = c->f[w2];
            a2.rcur = c->r[c->rcur];
            a2.rold = c->r[c->rold];
            a2.rout_old = c->r[v1];
            a2.rout_cur = c->r[v2];
            a2.data_row1 =
                (b + 1 < nt && c->nrec_total > 0) ? c->data_tm + (long)(nt - 2 - b) * c->nrec_total : nullptr;
            CHKRC(launch_step2_timed(c, mode, t2, a2));
            // as after two one-step launches: fold holds the newest
            // source-side field, rcur the newest receiver field
            c->fcur = w1;
            c->fold = w2;
            c->rold = v1;
            c->rcur = v2;
            if (b + 2 < b1 && b + 2 < nt)
                HIPCHK(c, fwi::launch_inject(c->r[c->rcur], c->cell_off, c->cell_start, c->cell_rec, c->ncells,
                                             c->data_tm + (long)(nt - 3 - b) * c->nrec_total, c->idh3, c->st));
        }
    }
    for (; b < b1; ++b) {
        const int i = b - b0;
        if (defer3) {
            const bool img_now = mode == fwi::BWD_RTM && i % 2 == 1;
            const bool ill_now = i % 3 == 2;
            a.img = img_now ? (ill_now ? 5 : 3) : (ill_now ? 4 : 0);
        } else {
            a.img = !defer ? 1 : (i % 2 == 1 ? 2 : (b + 1 < b1 ? 0 : 1));
        }
        a.cur = c->f[c->fold];
        a.old = c->f[c->fcur];
        a.rcur = c->r[c->rcur];
        a.rold = c->r[c->rold];
        // the newer source-side field of step b is the forward field of step
        // nsteps - b, which carries that step's source term
        const int ns = nsteps - b;
        a.inj = (c->phys & FWI_PHYS_SUBTRACT_SOURCE) && ns >= 0 && ns < nt ? c->idh3 * c->wavelet[ns] : 0.0f;
        CHKRC(launch_step_timed(c, mode, t, a));
        std::swap(c->fcur, c->fold);
        std::swap(c->rcur, c->rold);
        if (b + 1 < b1 && b + 1 < nt)
            HIPCHK(c, fwi::launch_inject(c->r[c->rcur], c->cell_off, c->cell_start, c->cell_rec, c->ncells,
                                         c->data_tm + (long)(nt - 2 - b) * c->nrec_total, c->idh3, c->st));
    }
    if (defer3 && b0 < b1) {
        // the terms of the steps after the last update: the illumination's
        // of the last one or two steps, the image's of the last step when
        // the segment has odd length
        const int last = b1 - b0 - 1;
        const int ill_terms = (last % 3 + 1) % 3;
        const int img_term = mode == fwi::BWD_RTM && last % 2 == 0;
        HIPCHK(c, fwi::launch_flush_imaging(c->f[c->fold], c->f[c->fcur], c->r[c->rold], a.image, a.illum,
                                            c->L, c->nshots, t.sb, c->L.vol, ill_terms, img_term, c->st));
    }
    const int groups = c->nshots / t.sb;
    if (b0 < b1 && groups > 1) {
        HIPCHK(c, fwi::launch_fold_groups(kind == 0 ? c->image : c->grad, c->L.vol, groups, c->st));
        HIPCHK(c, fwi::launch_fold_groups(c->illum, c->L.vol, groups, c->st));
    }
    return FWI_OK;
}

void drop_graphs(fwi_ctx* c) {
    for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    c->graphs.clear();
}

// Run `body` directly on the stream or as a hipGraph, bracketed by timing
// events.  A pass's graph is captured and instantiated once and replayed
// while its key (the pass, the field parity at its start, the setter
// generation) is unchanged.
template <typename F>
int timed_pass(fwi_ctx* c, int launches, F body, const char* pass_key = nullptr) {
    c->kev_used = 0;
    c->kernel_ms = 0.0;
    if (c->use_graph) {
        // the key names the buffers the pass starts from: forward passes use
        // the source-side fields only, backward passes both sides
        char key[160];
        const bool bwd_pass = pass_key && pass_key[0] == 'b';
        std::snprintf(key, sizeof key, "%s|f%d%d r%d%d", pass_key ? pass_key : "-", c->fcur, c->fold,
                      bwd_pass ? c->rcur : 0, bwd_pass ? c->rold : 1);
        // graphs of an older generation can never match again
        for (size_t i = 0; i < c->graphs.size();)
            if (c->graphs[i].gen != c->gen) {
                (void)hipGraphExecDestroy(c->graphs[i].exec);
                c->graphs.erase(c->graphs.begin() + i);
            } else {
                ++i;
            }
        const fwi_ctx::PassGraph* hit = nullptr;
        if (pass_key)
            for (const auto& g : c->graphs)
                if (g.key == key) hit = &g;
        hipGraphExec_t exec = hit ? hit->exec : nullptr;
        if (!hit) {
            hipGraph_t graph = nullptr;
            HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
            const int saved_f = c->fcur, saved_fo = c->fold, saved_r = c->rcur, saved_ro = c->rold;
            int rc = body();
            hipError_t e = hipStreamEndCapture(c->st, &graph);
            if (rc || e != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                c->fcur = saved_f;
                c->fold = saved_fo;
                c->rcur = saved_r;
                c->rold = saved_ro;
                if (rc) return rc;
                return fail(c, FWI_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            }
            hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ei != hipSuccess) {
                c->fcur = saved_f;
                c->fold = saved_fo;
                c->rcur = saved_r;
                c->rold = saved_ro;
                return fail(c, FWI_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
            }
            c->graph_instantiations++;
            if (pass_key) {
                if (c->graphs.size() >= 8) {
                    (void)hipGraphExecDestroy(c->graphs.front().exec);
                    c->graphs.erase(c->graphs.begin());
                }
                c->graphs.push_back({key, c->gen, exec, c->fcur, c->fold, c->rcur, c->rold});
            }
        } else {
            c->fcur = hit->fcur_after;
            c->fold = hit->fold_after;
            if (bwd_pass) {  // a forward pass leaves the receiver side as it is
                c->rcur = hit->rcur_after;
                c->rold = hit->rold_after;
            }
        }
        HIPCHK(c, hipEventRecord(c->ev0, c->st));
        HIPCHK(c, hipGraphLaunch(exec, c->st));
        HIPCHK(c, hipEventRecord(c->ev1, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        if (!pass_key) (void)hipGraphExecDestroy(exec);
    } else {
        HIPCHK(c, hipEventRecord(c->ev0, c->st));
        CHKRC(body());
        HIPCHK(c, hipEventRecord(c->ev1, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
    }
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_launches = launches;
    for (int i = 0; i < c->kev_used; ++i) {
        float k = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&k, c->kev[2 * i], c->kev[2 * i + 1]));
        c->kernel_ms += k;
    }
    return FWI_OK;
}

int check_ready_forward(fwi_ctx* c, int boundary) {
    if (boundary != FWI_BOUNDARY_ABC && boundary != FWI_BOUNDARY_RBC) return fail(c, FWI_ERR_ARG, "boundary");
    if (!c->have_model) return fail(c, FWI_ERR_STATE, "model not set");
    if ((int)c->wavelet.size() != c->g.nt) return fail(c, FWI_ERR_STATE, "wavelet not set");
    for (int s = 0; s < c->nshots; ++s)  // -2: a disabled slot (fwi_shot_disable)
        if (c->sx[s] == -1) return fail(c, FWI_ERR_STATE, "source not set for every shot slot");
    if (boundary == FWI_BOUNDARY_ABC && !c->b1d) return fail(c, FWI_ERR_STATE, "dampers not set");
    if (boundary == FWI_BOUNDARY_RBC && c->tune[fwi::FWD_RBC].tb == 2)  // the outputs of two-step launches
        for (float** p : {&c->f[2], &c->f[3]}) CHKRC(alloc_slots(c, p));
    return sync_geometry(c);
}

// max/min/non-finite check of every slot's model.
int check_model_stats(fwi_ctx* c) {
    for (int s = 0; s < c->nshots; ++s) {
        HIPCHK(c, fwi::launch_stats(c->vp + (long)s * c->L.vol, c->L, c->red_partial, c->red_result, c->st));
        double h[3];
        HIPCHK(c, hipMemcpyAsync(h, c->red_result, sizeof(h), hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        if (h[2] > 0) return fail(c, FWI_ERR_NONFINITE, "model holds NaN/Inf");
        const double cfl = std::max(std::fabs(h[0]), std::fabs(h[1])) * c->g.dt / c->g.dh;
        if (cfl > fwi::CFL_MAX) {
            char buf[160];
            std::snprintf(buf, sizeof buf, "CFL violated: vmax*dt/dh = %.6f > %.6f", cfl, fwi::CFL_MAX);
            return fail(c, FWI_ERR_CFL, buf);
        }
    }
    return FWI_OK;
}

// copy slot 0's model into the other slots
int replicate_model(fwi_ctx* c) {
    for (int s = 1; s < c->nshots; ++s)
        HIPCHK(c, hipMemcpyAsync(c->vp + (long)s * c->L.vol, c->vp, c->L.vol * sizeof(float), hipMemcpyDeviceToDevice,
                                 c->st));
    return FWI_OK;
}

const int kTileChoices[][2] = {{4, 1}, {4, 2}, {8, 1}, {8, 2}, {4, 4}, {16, 1}, {16, 2}, {2, 2}, {2, 4}};

// Copy the columns of slot s out of a time-major [nt][nrec_total] matrix
// into a compact [nt][nrec_s] one (to_slot = false) or back (true).
int slot_columns(fwi_ctx* c, int s, float* tm, float* compact, bool to_slot) {
    const int q0 = c->rec_start_h[s], ns = c->rec_start_h[s + 1] - q0;
    const size_t wide = (size_t)c->nrec_total * sizeof(float), narrow = (size_t)ns * sizeof(float);
    if (to_slot)
        HIPCHK(c, hipMemcpy2DAsync(tm + q0, wide, compact, narrow, narrow, c->g.nt, hipMemcpyDeviceToDevice, c->st));
    else
        HIPCHK(c, hipMemcpy2DAsync(compact, narrow, tm + q0, wide, narrow, c->g.nt, hipMemcpyDeviceToDevice, c->st));
    return FWI_OK;
}

// Seeded generator of the host-side random draws (splitmix64 stream).
struct HostRng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
};

// Blobs of the reference-style random boundary (modeling.cu:558-592):
// Poisson-disk points with minimum distance `ratio` in the padded box
// [0, (n-1) dh)^3 (Bridson's algorithm: 30 candidates in the shell
// [ratio, 2 ratio) around a random active point, grid cells of ratio/sqrt(3)
// so that the 5x5x5 neighbourhood covers the exclusion ball), keeping those
// outside the inner box 0.5*nb*dh away from the faces; each gets a radius
// r ~ U[0.2 ratio, ratio), an amplitude A ~ U[0.5 var_vp, var_vp) and a sign
// (-1 with probability 1/3).  Output rows: xc, yc, zc, r, +-A.
// The sample is grown only outside a core box (the inner box shrunk by
// 2 ratio): points there would be dropped, and a kept point's neighbourhood
// (its exclusion ball and the candidates around it) never reaches the core,
// so the kept points are a Poisson-disk sample of the shell as before, at
// about half the cost for a 50-cell boundary.
std::vector<float> make_blobs(uint64_t seed, int shot, int nxx, int nyy, int nzz, int nb, float dh, float ratio,
                              float var_vp) {
    HostRng rng{(seed * 0xD6E8FEB86659FD93ull) ^ ((uint64_t)(uint32_t)shot * 0xA0761D6478BD642Full) ^
                0x2545F4914F6CDD1Dull};
    const double xm = (double)((float)(nxx - 1) * dh), ym = (double)((float)(nyy - 1) * dh),
                 zm = (double)((float)(nzz - 1) * dh);
    const double rad = ratio, cell = rad / std::sqrt(3.0);
    const int gx = std::max(1, (int)std::ceil(xm / cell)), gy = std::max(1, (int)std::ceil(ym / cell)),
              gz = std::max(1, (int)std::ceil(zm / cell));
    std::vector<int> grid((size_t)gx * gy * gz, -1);
    std::vector<double> pts;  // x, y, z triples
    std::vector<int> active;
    auto gidx = [&](double x, double y, double z, int& i, int& j, int& k) {
        i = std::min(gx - 1, (int)(x / cell));
        j = std::min(gy - 1, (int)(y / cell));
        k = std::min(gz - 1, (int)(z / cell));
    };
    const double fnb = (double)((float)nb * dh), cm = 0.5 * fnb + 2.0 * rad;
    auto in_core = [&](double x, double y, double z) {
        return x > cm && x < xm - cm && y > cm && y < ym - cm && z > cm && z < zm - cm;
    };
    auto valid = [&](double x, double y, double z) {
        if (x < 0 || y < 0 || z < 0 || x >= xm || y >= ym || z >= zm || in_core(x, y, z)) return false;
        int i, j, k;
        gidx(x, y, z, i, j, k);
        for (int c = std::max(0, k - 2); c <= std::min(gz - 1, k + 2); ++c)
            for (int b = std::max(0, j - 2); b <= std::min(gy - 1, j + 2); ++b)
                for (int a = std::max(0, i - 2); a <= std::min(gx - 1, i + 2); ++a) {
                    const int q = grid[(size_t)a + (size_t)gx * (b + (size_t)gy * c)];
                    if (q < 0) continue;
                    const double dx = pts[3 * q] - x, dy = pts[3 * q + 1] - y, dz = pts[3 * q + 2] - z;
                    if (dx * dx + dy * dy + dz * dz < rad * rad) return false;
                }
        return true;
    };
    auto insert = [&](double x, double y, double z) {
        int i, j, k;
        gidx(x, y, z, i, j, k);
        const int q = (int)(pts.size() / 3);
        pts.push_back(x);
        pts.push_back(y);
        pts.push_back(z);
        grid[(size_t)i + (size_t)gx * (j + (size_t)gy * k)] = q;
        active.push_back(q);
    };
    if (xm > 0 && ym > 0 && zm > 0 && rad > 0) {
        // first point: uniform in the box, redrawn while in the core (the
        // shell holds at least the faces' outer half-boundary, so this ends)
        double x, y, z;
        do {
            x = rng.uniform() * xm;
            y = rng.uniform() * ym;
            z = rng.uniform() * zm;
        } while (in_core(x, y, z));
        insert(x, y, z);
    }
    while (!active.empty()) {
        const size_t a = (size_t)(rng.uniform() * (double)active.size());
        const int p = active[a];
        bool found = false;
        for (int t = 0; t < 30 && !found; ++t) {
            const double u = 2.0 * rng.uniform() - 1.0, th = 2.0 * M_PI * rng.uniform();
            const double rr = rad * (1.0 + rng.uniform()), sq = std::sqrt(std::max(0.0, 1.0 - u * u));
            const double x = pts[3 * p] + rr * sq * std::cos(th), y = pts[3 * p + 1] + rr * sq * std::sin(th),
                         z = pts[3 * p + 2] + rr * u;
            if (valid(x, y, z)) {
                insert(x, y, z);
                found = true;
            }
        }
        if (!found) {
            active[a] = active.back();
            active.pop_back();
        }
    }
    std::vector<float> out;
    for (size_t q = 0; q < pts.size() / 3; ++q) {
        const double x = pts[3 * q], y = pts[3 * q + 1], z = pts[3 * q + 2];
        const bool inner = x > 0.5 * fnb && x < xm - 0.5 * fnb && y > 0.5 * fnb && y < ym - 0.5 * fnb &&
                           z > 0.5 * fnb && z < zm - 0.5 * fnb;
        if (inner) continue;
        const float r = (float)(rad * (0.2 + 0.8 * rng.uniform()));
        const float A = (float)((double)var_vp * (0.5 + 0.5 * rng.uniform()));
        const float sign = rng.next() % 3 == 0 ? -1.0f : 1.0f;
        out.insert(out.end(), {(float)x, (float)y, (float)z, r, sign * A});
    }
    return out;
}

}  // namespace

namespace fwi {
// the message fwi_last_error(NULL) returns (entry points without a context)
void set_global_error(const std::string& msg) { g_create_error = msg; }
}  // namespace fwi

extern "C" {

int fwi_create_batch(const fwi_grid* grid, int nshots, int device, fwi_ctx** out) {
    if (!grid || !out) return fail(nullptr, FWI_ERR_ARG, "NULL argument");
    *out = nullptr;
    const fwi_grid& g = *grid;
    if (g.nxx < 9 || g.nyy < 9 || g.nzz < 9 || g.nb < 0 || g.nt < 1 || g.tlag < 0 || !(g.dh > 0) || !(g.dt > 0))
        return fail(nullptr, FWI_ERR_ARG, "invalid grid");
    if (nshots < 1) return fail(nullptr, FWI_ERR_ARG, "nshots must be >= 1");
    // the step kernels address a plane with 32-bit byte offsets below 2^31
    // (volumes of 2 GiB and more are addressed plane by plane)
    if (fwi::make_layout(g.nxx, g.nyy, g.nzz).pp * 4 >= (1L << 31))
        return fail(nullptr, FWI_ERR_RANGE, "padded x-y plane exceeds 2 GiB");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(nullptr, FWI_ERR_NODEV, "no HIP device");
    if (device < 0 || device >= ndev) return fail(nullptr, FWI_ERR_ARG, "device index out of range");
    if (hipSetDevice(device) != hipSuccess) return fail(nullptr, FWI_ERR_HIP, "hipSetDevice failed");

    fwi_ctx* c = new fwi_ctx();
    c->device = device;
    c->g = g;
    c->nshots = nshots;
    c->L = fwi::make_layout(g.nxx, g.nyy, g.nzz);
    c->idh2 = 1.0f / (g.dh * g.dh);
    c->idh3 = 1.0f / (g.dh * g.dh * g.dh);
    c->dt2 = g.dt * g.dt;
    c->sx.assign(nshots, -1);
    c->sy.assign(nshots, -1);
    c->sz.assign(nshots, -1);
    c->rx.assign(nshots, {});
    c->ry.assign(nshots, {});
    c->rz.assign(nshots, {});
    auto bail = [&](int rc) {
        g_create_error = c->err;
        fwi_destroy(c);
        return rc;
    };
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(c, FWI_ERR_HIP, "hipStreamCreate failed"));
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return bail(fail(c, FWI_ERR_HIP, "hipEventCreate failed"));
    int rc;
    if ((rc = alloc_slots(c, &c->vp)) || (rc = alloc_slots(c, &c->f[0])) || (rc = alloc_slots(c, &c->f[1])))
        return bail(rc);
    if (hipMalloc((void**)&c->staging, nvol_host(g) * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&c->red_partial, fwi::kReduceBlocks * 3 * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&c->red_result, 4 * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&c->src_off, nshots * sizeof(long)) != hipSuccess ||
        hipMalloc((void**)&c->rec_start, (nshots + 1) * sizeof(int)) != hipSuccess)
        return bail(fail(c, FWI_ERR_HIP, "hipMalloc failed"));
    if (hipMemset(c->src_off, 0, nshots * sizeof(long)) != hipSuccess)
        return bail(fail(c, FWI_ERR_HIP, "hipMemset failed"));
    if ((rc = sync_geometry(c))) return bail(rc);
    const bool big = c->L.vol * 4 >= (1L << 31);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1)
        cus = 256;
    for (int k = 0; k < 4; ++k) {
        Tuning& t = c->tune[k];
        t = Tuning{};
        if (big) {  // the per-plane-addressed instances
            t.big = 1;
            t.by = 16;
            t.rpt = 1;
        } else if (nshots == 1) {
            // one-shot contexts: 64 x 32 tiles for the forward passes and
            // 64 x 16 for the backward ones (16 waves per block), with the
            // fewest z-chunks that give about one block per CU -- the shapes
            // the autotuner picks on large grids (batched contexts keep the
            // smaller tiles that their shots-per-block instances take)
            t.by = 16;
            t.rpt = k < 2 ? 2 : 1;
            const long tiles = (long)((c->L.nx + 63) / 64) * ((c->L.ny - 8 + t.by * t.rpt - 1) / (t.by * t.rpt));
            while (tiles * t.zchunks < cus * 15L / 16 && t.zchunks < 16 && t.zchunks < g.nzz - 8) ++t.zchunks;
        }
    }
    *out = c;
    return FWI_OK;
}

int fwi_create(const fwi_grid* grid, int device, fwi_ctx** out) { return fwi_create_batch(grid, 1, device, out); }

int fwi_destroy(fwi_ctx* c) {
    if (!c) return FWI_OK;
    (void)hipSetDevice(c->device);
    if (c->st) (void)hipStreamSynchronize(c->st);
    drop_graphs(c);
    float* vols[] = {c->vp, c->f[0], c->f[1], c->f[2], c->f[3], c->r[0], c->r[1], c->r[2], c->r[3], c->image,
                     c->illum, c->grad, c->staging,
                     c->b1d, c->rec_damp, c->seis_tm, c->data_tm, c->seis_tmp};
    for (float* p : vols)
        if (p) (void)hipFree(p);
    void* others[] = {c->src_off, c->rec_off, c->rec_start, c->cell_off, c->cell_start, c->cell_rec,
                      c->red_partial, c->red_result, c->rp_z, c->rp_map};
    for (void* p : others)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : c->kev) (void)hipEventDestroy(e);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return FWI_OK;
}

const char* fwi_last_error(const fwi_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int fwi_set_arith(fwi_ctx* c, int arith) {
    if (!c) return FWI_ERR_ARG;
    c->gen++;  // launch arguments may change: recapture pass graphs
    if (arith != FWI_ARITH_REFERENCE && arith != FWI_ARITH_FP32) return fail(c, FWI_ERR_ARG, "arith");
    for (int k = 0; k < 4; ++k)
        if (!fwi::step_config_supported(k, arith, c->tune[k])) {
            c->tune[k].by = 4;
            c->tune[k].rpt = 2;
            c->tune[k].sb = 1;
            c->tune[k].spt = 1;
            c->tune[k].pd = 1;
        }
    c->arith = arith;
    return FWI_OK;
}

int fwi_set_model(fwi_ctx* c, const float* vp) {
    if (!c || !vp) return c ? fail(c, FWI_ERR_ARG, "NULL model") : FWI_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(c->staging, vp, nvol_host(c->g) * sizeof(float), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, fwi::launch_to_device(c->staging, c->vp, c->L, c->st));
    CHKRC(replicate_model(c));
    int rc = check_model_stats(c);
    c->have_model = rc == FWI_OK;
    return rc;
}

int fwi_set_model_interior(fwi_ctx* c, const float* vp) {
    if (!c || !vp) return c ? fail(c, FWI_ERR_ARG, "NULL model") : FWI_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    const long n = ninterior(c->g);
    if (n <= 0) return fail(c, FWI_ERR_ARG, "empty interior");
    HIPCHK(c, hipMemcpyAsync(c->staging, vp, n * sizeof(float), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, fwi::launch_expand(c->staging, c->vp, c->L, c->g.nb, c->st));
    CHKRC(replicate_model(c));
    int rc = check_model_stats(c);
    c->have_model = rc == FWI_OK;
    return rc;
}

int fwi_shot_random_boundary(fwi_ctx* c, int slot, uint64_t seed, int shot, float amp, float vclamp) {
    if (!c) return FWI_ERR_ARG;
    CHKRC(check_slot(c, slot));
    if (!c->have_model) return fail(c, FWI_ERR_STATE, "model not set");
    if (!(amp >= 0.0f && amp < 1.0f)) return fail(c, FWI_ERR_ARG, "amp must be in [0, 1)");
    if (c->g.nb < 1) return fail(c, FWI_ERR_ARG, "random boundary needs nb >= 1");
    if (ninterior(c->g) <= 0 || c->g.nxx <= 2 * c->g.nb || c->g.nyy <= 2 * c->g.nb)
        return fail(c, FWI_ERR_ARG, "random boundary needs a non-empty interior");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, fwi::launch_random_boundary(c->vp + (long)slot * c->L.vol, c->L, c->g.nb, seed, shot, amp, vclamp,
                                          c->st));
    int rc = check_model_stats(c);
    c->have_model = rc == FWI_OK;
    return rc;
}

int fwi_random_boundary(fwi_ctx* c, uint64_t seed, int shot, float amp, float vclamp) {
    return fwi_shot_random_boundary(c, 0, seed, shot, amp, vclamp);
}

int fwi_rbc_blobs(uint64_t seed, int shot, int nxx, int nyy, int nzz, int nb, float dh, float ratio, float var_vp,
                  float* out, int cap, int* count) {
    if (!count || nxx < 1 || nyy < 1 || nzz < 1 || nb < 0 || !(dh > 0) || !(ratio > 0) || !(var_vp >= 0))
        return FWI_ERR_ARG;
    const std::vector<float> b = make_blobs(seed, shot, nxx, nyy, nzz, nb, dh, ratio, var_vp);
    *count = (int)(b.size() / 5);
    if (out && cap >= *count) std::memcpy(out, b.data(), b.size() * sizeof(float));
    return FWI_OK;
}

int fwi_shot_random_boundary_blobs(fwi_ctx* c, int slot, uint64_t seed, int shot, float ratio, float var_vp,
                                   float vclamp) {
    if (!c) return FWI_ERR_ARG;
    CHKRC(check_slot(c, slot));
    if (!c->have_model) return fail(c, FWI_ERR_STATE, "model not set");
    if (!(ratio > 0) || !(var_vp >= 0)) return fail(c, FWI_ERR_ARG, "ratio must be > 0 and var_vp >= 0");
    if (c->g.nb < 1) return fail(c, FWI_ERR_ARG, "random boundary needs nb >= 1");
    if (ninterior(c->g) <= 0 || c->g.nxx <= 2 * c->g.nb || c->g.nyy <= 2 * c->g.nb)
        return fail(c, FWI_ERR_ARG, "random boundary needs a non-empty interior");
    HIPCHK(c, hipSetDevice(c->device));
    float* vol = c->vp + (long)slot * c->L.vol;
    // the clamp bounds of random_boundary_gp: the extremes of the interior
    // model (modeling.cu:164-171) -/+ var_vp; the interior is a sub-layout
    // with the same pitches, so earlier boundaries do not widen the bounds
    const int nb = c->g.nb;
    fwi::Layout inner = c->L;
    inner.nx -= 2 * nb;
    inner.ny -= 2 * nb;
    inner.nz -= 2 * nb;
    HIPCHK(c, fwi::launch_stats(vol + c->L.at(nb, nb, nb), inner, c->red_partial, c->red_result, c->st));
    double h[3];
    HIPCHK(c, hipMemcpyAsync(h, c->red_result, sizeof(h), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    const std::vector<float> b = make_blobs(seed, shot, c->g.nxx, c->g.nyy, c->g.nzz, nb, c->g.dh, ratio, var_vp);
    const int nblobs = (int)(b.size() / 5);
    float* d_blobs = nullptr;
    if (nblobs > 0) {
        HIPCHK(c, hipMalloc((void**)&d_blobs, b.size() * sizeof(float)));
        hipError_t e = hipMemcpy(d_blobs, b.data(), b.size() * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(d_blobs);
            return fail(c, FWI_ERR_HIP, std::string("blob upload: ") + hipGetErrorString(e));
        }
    }
    hipError_t e = fwi::launch_blob_boundary(vol, c->L, nb, c->g.dh, seed, shot, var_vp, d_blobs, nblobs,
                                             (float)h[1] - var_vp, (float)h[0] + var_vp, vclamp, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (d_blobs) (void)hipFree(d_blobs);
    if (e != hipSuccess) return fail(c, FWI_ERR_HIP, std::string("blob boundary: ") + hipGetErrorString(e));
    int rc = check_model_stats(c);
    c->have_model = rc == FWI_OK;
    return rc;
}

int fwi_random_boundary_blobs(fwi_ctx* c, uint64_t seed, int shot, float ratio, float var_vp, float vclamp) {
    return fwi_shot_random_boundary_blobs(c, 0, seed, shot, ratio, var_vp, vclamp);
}

int fwi_set_wavelet(fwi_ctx* c, const float* w) {
    if (!c || !w) return c ? fail(c, FWI_ERR_ARG, "NULL wavelet") : FWI_ERR_ARG;
    std::vector<float> v(w, w + c->g.nt);
    for (float x : v)
        if (!std::isfinite(x)) return fail(c, FWI_ERR_NONFINITE, "wavelet holds NaN/Inf");
    c->wavelet.swap(v);
    c->gen++;  // the source terms are kernel arguments
    return FWI_OK;
}

int fwi_set_dampers(fwi_ctx* c, const float* b1d) {
    if (!c || !b1d) return c ? fail(c, FWI_ERR_ARG, "NULL dampers") : FWI_ERR_ARG;
    const int nb = c->g.nb;
    if (nb < 1) return fail(c, FWI_ERR_ARG, "ABC needs nb >= 1");
    HIPCHK(c, hipSetDevice(c->device));
    c->b1d_h.assign(b1d, b1d + nb);
    if (!c->b1d) HIPCHK(c, hipMalloc((void**)&c->b1d, nb * sizeof(float)));
    HIPCHK(c, hipMemcpy(c->b1d, b1d, nb * sizeof(float), hipMemcpyHostToDevice));
    c->geom_dirty = true;  // receiver damping factors depend on the table
    return FWI_OK;
}

int fwi_shot_set_source(fwi_ctx* c, int slot, int sx, int sy, int sz) {
    if (!c) return FWI_ERR_ARG;
    CHKRC(check_slot(c, slot));
    if (!in_updated(c, sx, sy, sz)) return fail(c, FWI_ERR_RANGE, "source outside the updated interior");
    HIPCHK(c, hipSetDevice(c->device));
    c->sx[slot] = sx;
    c->sy[slot] = sy;
    c->sz[slot] = sz;
    const long off = c->L.at(sx, sy, sz);
    HIPCHK(c, hipMemcpy(c->src_off + slot, &off, sizeof off, hipMemcpyHostToDevice));
    return FWI_OK;
}

int fwi_set_source(fwi_ctx* c, int sx, int sy, int sz) { return fwi_shot_set_source(c, 0, sx, sy, sz); }

int fwi_shot_disable(fwi_ctx* c, int slot) {
    if (!c) return FWI_ERR_ARG;
    CHKRC(check_slot(c, slot));
    HIPCHK(c, hipSetDevice(c->device));
    c->sx[slot] = c->sy[slot] = c->sz[slot] = -2;
    const long off = -1;  // no source cell: the injection kernels skip the slot
    HIPCHK(c, hipMemcpy(c->src_off + slot, &off, sizeof off, hipMemcpyHostToDevice));
    c->rx[slot].clear();
    c->ry[slot].clear();
    c->rz[slot].clear();
    c->geom_dirty = true;
    return sync_geometry(c);
}

int fwi_shot_set_receivers(fwi_ctx* c, int slot, int nrec, const int* rx, const int* ry, const int* rz) {
    if (!c) return FWI_ERR_ARG;
    CHKRC(check_slot(c, slot));
    if (nrec < 0 || (nrec > 0 && (!rx || !ry || !rz))) return fail(c, FWI_ERR_ARG, "bad receiver arrays");
    for (int q = 0; q < nrec; ++q)
        if (!in_updated(c, rx[q], ry[q], rz[q])) return fail(c, FWI_ERR_RANGE, "receiver outside the updated interior");
    c->rx[slot].assign(rx, rx + nrec);
    c->ry[slot].assign(ry, ry + nrec);
    c->rz[slot].assign(rz, rz + nrec);
    c->geom_dirty = true;
    HIPCHK(c, hipSetDevice(c->device));
    return sync_geometry(c);
}

int fwi_set_receivers(fwi_ctx* c, int nrec, const int* rx, const int* ry, const int* rz) {
    return fwi_shot_set_receivers(c, 0, nrec, rx, ry, rz);
}

int fwi_forward_steps(fwi_ctx* c, int boundary, int n0, int n1) {
    if (!c) return FWI_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    CHKRC(check_ready_forward(c, boundary));
    if (n0 < 0 || n1 < n0 || n1 > c->g.nt + c->g.tlag) return fail(c, FWI_ERR_ARG, "step range");
    char key[64];
    std::snprintf(key, sizeof key, "fwd b%d %d-%d", boundary, n0, n1);
    int rc = timed_pass(c, n1 - n0, [&] { return run_forward(c, boundary, n0, n1); }, key);
    c->fwd_done = 0;
    return rc;
}

int fwi_forward(fwi_ctx* c, int boundary) {
    if (!c) return FWI_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    CHKRC(check_ready_forward(c, boundary));
    const long vols = c->L.vol * c->nshots;
    for (float* p : c->f)
        if (p) HIPCHK(c, hipMemsetAsync(p, 0, vols * sizeof(float), c->st));
    if (!c->f_rims_zero) c->gen++;
    c->f_rims_zero = true;
    if (c->nrec_total > 0)
        HIPCHK(c, hipMemsetAsync(c->seis_tm, 0, (long)c->g.nt * c->nrec_total * sizeof(float), c->st));
    c->fcur = 0;
    c->fold = 1;
    const int N = c->g.nt + c->g.tlag;
    char key[64];
    std::snprintf(key, sizeof key, "fwd b%d 0-%d", boundary, N);
    int rc = timed_pass(c, N, [&] { return run_forward(c, boundary, 0, N); }, key);
    c->fwd_done = rc == FWI_OK ? 1 : 0;
    return rc;
}

int fwi_shot_get_seismogram(fwi_ctx* c, int slot, float* out) {
    if (!c || !out) return c ? fail(c, FWI_ERR_ARG, "NULL output") : FWI_ERR_ARG;
    CHKRC(check_slot(c, slot));
    HIPCHK(c, hipSetDevice(c->device));
    CHKRC(sync_geometry(c));
    const int ns = c->rec_start_h[slot + 1] - c->rec_start_h[slot];
    if (ns == 0) return FWI_OK;
    float* compact = c->seis_tmp;
    float* tr = c->seis_tmp + (long)c->g.nt * c->nrec_total;
    CHKRC(slot_columns(c, slot, c->seis_tm, compact, false));
    HIPCHK(c, fwi::launch_transpose2d(compact, tr, c->g.nt, ns, c->st));
    HIPCHK(c, hipMemcpyAsync(out, tr, (long)c->g.nt * ns * sizeof(float), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return FWI_OK;
}

int fwi_get_seismogram(fwi_ctx* c, float* out) { return fwi_shot_get_seismogram(c, 0, out); }

int fwi_shot_set_backward_data(fwi_ctx* c, int slot, const float* data) {
    if (!c || !data) return c ? fail(c, FWI_ERR_ARG, "NULL data") : FWI_ERR_ARG;
    CHKRC(check_slot(c, slot));
    HIPCHK(c, hipSetDevice(c->device));
    CHKRC(sync_geometry(c));
    const int ns = c->rec_start_h[slot + 1] - c->rec_start_h[slot];
    if (ns == 0) return FWI_OK;
    float* up = c->seis_tmp;
    float* compact = c->seis_tmp + (long)c->g.nt * c->nrec_total;
    HIPCHK(c, hipMemcpyAsync(up, data, (long)c->g.nt * ns * sizeof(float), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, fwi::launch_transpose2d(up, compact, ns, c->g.nt, c->st));
    CHKRC(slot_columns(c, slot, c->data_tm, compact, true));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return FWI_OK;
}

int fwi_set_backward_data(fwi_ctx* c, const float* data) { return fwi_shot_set_backward_data(c, 0, data); }

int fwi_shot_get_backward_data(fwi_ctx* c, int slot, float* out) {
    if (!c || !out) return c ? fail(c, FWI_ERR_ARG, "NULL output") : FWI_ERR_ARG;
    CHKRC(check_slot(c, slot));
    HIPCHK(c, hipSetDevice(c->device));
    CHKRC(sync_geometry(c));
    const int ns = c->rec_start_h[slot + 1] - c->rec_start_h[slot];
    if (ns == 0) return FWI_OK;
    float* compact = c->seis_tmp;
    float* tr = c->seis_tmp + (long)c->g.nt * c->nrec_total;
    CHKRC(slot_columns(c, slot, c->data_tm, compact, false));
    HIPCHK(c, fwi::launch_transpose2d(compact, tr, c->g.nt, ns, c->st));
    HIPCHK(c, hipMemcpyAsync(out, tr, (long)c->g.nt * ns * sizeof(float), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return FWI_OK;
}

int fwi_backward_steps(fwi_ctx* c, int kind, int b0, int b1) {
    if (!c) return FWI_ERR_ARG;
    if (kind != 0 && kind != 1) return fail(c, FWI_ERR_ARG, "kind");
    if (!c->have_model) return fail(c, FWI_ERR_STATE, "model not set");
    if (b0 < 0 || b1 < b0 || b1 > c->g.nt + c->g.tlag) return fail(c, FWI_ERR_ARG, "step range");
    HIPCHK(c, hipSetDevice(c->device));
    CHKRC(sync_geometry(c));
    CHKRC(ensure_backward_buffers(c, kind));
    char key[64];
    std::snprintf(key, sizeof key, "bwd k%d %d-%d", kind, b0, b1);
    return timed_pass(c, b1 - b0, [&] { return run_backward(c, kind, b0, b1); }, key);
}

static int full_backward(fwi_ctx* c, int kind) {
    if (!c) return FWI_ERR_ARG;
    if (!c->fwd_done) return fail(c, FWI_ERR_STATE, "backward pass needs a completed forward pass");
    HIPCHK(c, hipSetDevice(c->device));
    CHKRC(sync_geometry(c));
    CHKRC(ensure_backward_buffers(c, kind));
    const long vols = c->L.vol * c->nshots;
    for (float* p : c->r)
        if (p) HIPCHK(c, hipMemsetAsync(p, 0, vols * sizeof(float), c->st));
    if (!c->r_rims_zero) c->gen++;
    c->r_rims_zero = true;
    c->rcur = 0;
    c->rold = 1;
    const int N = c->g.nt + c->g.tlag;
    char key[64];
    std::snprintf(key, sizeof key, "bwd k%d 0-%d", kind, N);
    int rc = timed_pass(c, N, [&] { return run_backward(c, kind, 0, N); }, key);
    c->fwd_done = 0;
    return rc;
}

int fwi_rtm_backward(fwi_ctx* c) { return full_backward(c, 0); }
int fwi_grad_backward(fwi_ctx* c) { return full_backward(c, 1); }

int fwi_zero_volume(fwi_ctx* c, int which) {
    if (!c) return FWI_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    float* p = vol_ptr(c, 0, which, true, &rc);
    if (!p) return rc;
    HIPCHK(c, hipMemsetAsync(p, 0, c->L.vol * c->nshots * sizeof(float), c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return FWI_OK;
}

int fwi_shot_get_volume(fwi_ctx* c, int slot, int which, float* out) {
    if (!c || !out) return c ? fail(c, FWI_ERR_ARG, "NULL output") : FWI_ERR_ARG;
    CHKRC(check_slot(c, slot));
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    float* p = vol_ptr(c, slot, which, true, &rc);
    if (!p) return rc;
    HIPCHK(c, fwi::launch_to_host(p, c->staging, c->L, 0, c->st));
    HIPCHK(c, hipMemcpyAsync(out, c->staging, nvol_host(c->g) * sizeof(float), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return FWI_OK;
}

int fwi_get_volume(fwi_ctx* c, int which, float* out) { return fwi_shot_get_volume(c, 0, which, out); }

// Any non-zero value in the 4-cell rim of a host-layout volume (z fastest,
// then x, then y)?
bool rim_nonzero(const fwi_grid& g, const float* v) {
    for (int y = 0; y < g.nyy; ++y)
        for (int x = 0; x < g.nxx; ++x) {
            const float* row = v + ((long)y * g.nxx + x) * g.nzz;
            const bool edge = y < 4 || y >= g.nyy - 4 || x < 4 || x >= g.nxx - 4;
            for (int z = 0; z < g.nzz; z = (edge || z < 3 || z >= g.nzz - 5) ? z + 1 : g.nzz - 4)
                if (row[z] != 0.0f) return true;
        }
    return false;
}

int fwi_set_volume(fwi_ctx* c, int which, const float* in) {
    if (!c || !in) return c ? fail(c, FWI_ERR_ARG, "NULL input") : FWI_ERR_ARG;
    if (which == FWI_VOL_MODEL) return fwi_set_model(c, in);
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    float* p = vol_ptr(c, 0, which, true, &rc);
    if (!p) return rc;
    HIPCHK(c, hipMemcpyAsync(c->staging, in, nvol_host(c->g) * sizeof(float), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, fwi::launch_to_device(c->staging, p, c->L, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (which == FWI_VOL_FIELD || which == FWI_VOL_FIELD_OLD) c->fwd_done = 1;
    if (rim_nonzero(c->g, in)) {
        c->gen++;  // the next passes may run other launches: recapture pass graphs
        if (which == FWI_VOL_FIELD || which == FWI_VOL_FIELD_OLD) c->f_rims_zero = false;
        if (which == FWI_VOL_RECV || which == FWI_VOL_RECV_OLD) c->r_rims_zero = false;
    }
    return FWI_OK;
}

static int export_interior(fwi_ctx* c, const float* num, int zfirst, float* out, float* gmax) {
    const int nb = c->g.nb;
    const long n = ninterior(c->g);
    HIPCHK(c, fwi::launch_export_image(num, c->illum, c->staging, c->L, nb, c->idh2, zfirst, c->st));
    if (gmax) {
        HIPCHK(c, fwi::launch_absmax(c->staging, n, c->red_partial, c->red_result, c->st));
        HIPCHK(c, fwi::launch_scale_by_inverse(c->staging, n, c->red_result, c->st));
        double g = 0;
        HIPCHK(c, hipMemcpyAsync(&g, c->red_result, sizeof g, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipMemcpyAsync(out, c->staging, n * sizeof(float), hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        *gmax = (float)g;
        return FWI_OK;
    }
    HIPCHK(c, hipMemcpyAsync(out, c->staging, n * sizeof(float), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return FWI_OK;
}

int fwi_export_image(fwi_ctx* c, float* out) {
    if (!c || !out) return c ? fail(c, FWI_ERR_ARG, "NULL output") : FWI_ERR_ARG;
    if (!c->image || !c->illum) return fail(c, FWI_ERR_STATE, "no image accumulated");
    HIPCHK(c, hipSetDevice(c->device));
    return export_interior(c, c->image, 0, out, nullptr);
}

int fwi_export_gradient(fwi_ctx* c, float zmask, float* out, float* gmax) {
    if (!c || !out || !gmax) return c ? fail(c, FWI_ERR_ARG, "NULL output") : FWI_ERR_ARG;
    if (!c->grad || !c->illum) return fail(c, FWI_ERR_STATE, "no gradient accumulated");
    HIPCHK(c, hipSetDevice(c->device));
    // inversion.cu:152 masks i <= (int)(zmask/dh); a negative mask would make
    // the reference read below the first row, so it is clamped to 0 here
    const int zfirst = std::max(0, (int)(zmask / c->g.dh));
    return export_interior(c, c->grad, zfirst, out, gmax);
}

int fwi_shot_residual(fwi_ctx* c, int slot, const float* obs, double* sum_sq) {
    if (!c || !obs || !sum_sq) return c ? fail(c, FWI_ERR_ARG, "NULL argument") : FWI_ERR_ARG;
    CHKRC(check_slot(c, slot));
    *sum_sq = 0.0;
    HIPCHK(c, hipSetDevice(c->device));
    CHKRC(sync_geometry(c));
    const int ns = c->rec_start_h[slot + 1] - c->rec_start_h[slot];
    if (ns == 0) return FWI_OK;
    const long n = (long)c->g.nt * ns;
    float* obs_tm = c->seis_tmp;
    float* cal = c->seis_tmp + (long)c->g.nt * c->nrec_total;
    // obs (t + r*nt) -> [t][r] in obs_tm via cal as upload buffer
    HIPCHK(c, hipMemcpyAsync(cal, obs, n * sizeof(float), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, fwi::launch_transpose2d(cal, obs_tm, ns, c->g.nt, c->st));
    CHKRC(slot_columns(c, slot, c->seis_tm, cal, false));
    HIPCHK(c, fwi::launch_residual(obs_tm, cal, n, c->red_partial, c->red_result, c->st));
    CHKRC(slot_columns(c, slot, c->data_tm, cal, true));
    HIPCHK(c, hipMemcpyAsync(sum_sq, c->red_result, sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return FWI_OK;
}

int fwi_residual(fwi_ctx* c, const float* obs, double* sum_sq) { return fwi_shot_residual(c, 0, obs, sum_sq); }

int fwi_get_tuning(fwi_ctx* c, int kind, int* by, int* rpt, int* zchunks) {
    if (!c || kind < 0 || kind > 3 || !by || !rpt || !zchunks) return FWI_ERR_ARG;
    *by = c->tune[kind].by;
    *rpt = c->tune[kind].rpt;
    *zchunks = c->tune[kind].zchunks;
    return FWI_OK;
}

int fwi_set_tuning(fwi_ctx* c, int kind, int by, int rpt, int zchunks) {
    if (!c) return FWI_ERR_ARG;
    c->gen++;  // launch arguments may change: recapture pass graphs
    if (kind < 0 || kind > 3) return fail(c, FWI_ERR_ARG, "kind");
    Tuning t = c->tune[kind];
    t.by = by;
    t.rpt = rpt;
    t.zchunks = zchunks;
    if (!fwi::step_config_supported(kind, c->arith, t)) t.pd = 1;  // shapes without two-plane prefetch
    if (zchunks < 1 || zchunks > (c->g.nzz - 8) || !fwi::step_config_supported(kind, c->arith, t))
        return fail(c, FWI_ERR_ARG, "unsupported tile configuration");
    c->tune[kind] = t;
    return FWI_OK;
}

int fwi_set_prefetch_depth(fwi_ctx* c, int kind, int pd) {
    if (!c) return FWI_ERR_ARG;
    c->gen++;  // launch arguments may change: recapture pass graphs
    if (kind < 0 || kind > 3 || (pd != 1 && pd != 2)) return fail(c, FWI_ERR_ARG, "kind / prefetch depth");
    Tuning t = c->tune[kind];
    t.pd = pd;
    if (!fwi::step_config_supported(kind, c->arith, t))
        return fail(c, FWI_ERR_ARG, "no two-plane prefetch instance for this launch shape");
    c->tune[kind] = t;
    return FWI_OK;
}

int fwi_get_prefetch_depth(fwi_ctx* c, int kind, int* pd) {
    if (!c || kind < 0 || kind > 3 || !pd) return FWI_ERR_ARG;
    *pd = c->tune[kind].pd;
    return FWI_OK;
}

int fwi_set_two_step(fwi_ctx* c, int kind, int enable, int by, int rpt, int zchunks) {
    if (!c) return FWI_ERR_ARG;
    c->gen++;  // launch arguments may change: recapture pass graphs
    if (kind < 0 || kind > 3) return fail(c, FWI_ERR_ARG, "kind");
    Tuning t = c->tune[kind];
    t.tb = enable ? 2 : 1;
    if (enable) {
        t.tby = by;
        t.trpt = rpt;
        t.tzc = zchunks;
        Tuning u = t;
        u.by = by;
        u.rpt = rpt;
        u.zchunks = zchunks;
        u.sb = 1;
        u.big = 0;
        if ((kind != fwi::FWD_RBC && kind != fwi::BWD_RTM) || !fwi::step2_supported(kind, c->arith, u))
            return fail(c, FWI_ERR_ARG, "no two-step instance for this pass kind / tile");
        if (zchunks < 1 || zchunks > c->g.nzz - 8) return fail(c, FWI_ERR_ARG, "z chunks");
    }
    c->tune[kind] = t;
    // the receiver-plane maps of the current geometry
    if (enable && !c->geom_dirty && !c->rp_ok) {
        HIPCHK(c, hipSetDevice(c->device));
        CHKRC(build_rec_planes(c, c->cell_off_h));
    }
    return FWI_OK;
}

int fwi_get_two_step(fwi_ctx* c, int kind, int* active, int* by, int* rpt, int* zchunks) {
    if (!c || kind < 0 || kind > 3 || !active || !by || !rpt || !zchunks) return FWI_ERR_ARG;
    const Tuning& t = c->tune[kind];
    *active = two_step_eligible(c, kind) ? 1 : (t.tb == 2 ? -1 : 0);
    *by = t.tby;
    *rpt = t.trpt;
    *zchunks = t.tzc;
    return FWI_OK;
}

int fwi_set_physics(fwi_ctx* c, int flags) {
    if (!c) return FWI_ERR_ARG;
    c->gen++;  // launch arguments may change: recapture pass graphs
    if (flags & ~(FWI_PHYS_GRAD_DYY | FWI_PHYS_SUBTRACT_SOURCE | FWI_PHYS_ALIGNED_IMAGING))
        return fail(c, FWI_ERR_ARG, "unknown physics flag");
    c->phys = flags;
    return FWI_OK;
}

int fwi_set_deferred_imaging(fwi_ctx* c, int enable) {
    if (!c) return FWI_ERR_ARG;
    c->gen++;  // launch arguments may change: recapture pass graphs
    c->defer_imaging = enable != 0;
    return FWI_OK;
}

int fwi_set_shots_per_block(fwi_ctx* c, int kind, int sb) {
    if (!c) return FWI_ERR_ARG;
    c->gen++;  // launch arguments may change: recapture pass graphs
    if (kind < 0 || kind > 3) return fail(c, FWI_ERR_ARG, "kind");
    Tuning t = c->tune[kind];
    t.sb = sb;
    t.spt = 1;  // one shot per thread: sb groups of waves
    if (sb != 1) t.pd = 1;  // two-plane prefetch: one shot per block only
    if (sb < 1 || c->nshots % sb != 0) return fail(c, FWI_ERR_ARG, "unsupported shots-per-block");
    // keep the current tile if it exists for this sb, else take the first one that does
    if (!fwi::step_config_supported(kind, c->arith, t))
        for (const auto& tc : kTileChoices) {
            t.by = tc[0];
            t.rpt = tc[1];
            if (fwi::step_config_supported(kind, c->arith, t)) break;
        }
    if (!fwi::step_config_supported(kind, c->arith, t)) return fail(c, FWI_ERR_ARG, "unsupported shots-per-block");
    c->tune[kind] = t;
    return FWI_OK;
}

int fwi_get_shots_per_block(fwi_ctx* c, int kind, int* sb) {
    if (!c || kind < 0 || kind > 3 || !sb) return FWI_ERR_ARG;
    *sb = c->tune[kind].sb;
    return FWI_OK;
}

int fwi_set_shots_per_thread(fwi_ctx* c, int kind, int spt, int by, int zchunks) {
    if (!c) return FWI_ERR_ARG;
    c->gen++;  // launch arguments may change: recapture pass graphs
    if (kind < 0 || kind > 3 || (spt != 1 && spt != 2)) return fail(c, FWI_ERR_ARG, "kind / shots per thread");
    if (zchunks < 1 || zchunks > c->g.nzz - 8) return fail(c, FWI_ERR_ARG, "z chunks");
    Tuning t = c->tune[kind];
    t.spt = spt;
    t.by = by;
    t.zchunks = zchunks;
    t.pd = 1;
    if (spt == 2) {
        if (c->nshots % 2 != 0) return fail(c, FWI_ERR_ARG, "shot pairs need an even number of slots");
        if (!pairs_fit(c)) return fail(c, FWI_ERR_RANGE, "shot pairs need two volumes below 2 GiB");
        t.sb = 2;
        t.rpt = 1;
    }
    if (!fwi::step_config_supported(kind, c->arith, t)) return fail(c, FWI_ERR_ARG, "unsupported launch shape");
    c->tune[kind] = t;
    return FWI_OK;
}

int fwi_get_shots_per_thread(fwi_ctx* c, int kind, int* spt) {
    if (!c || kind < 0 || kind > 3 || !spt) return FWI_ERR_ARG;
    *spt = c->tune[kind].spt;
    return FWI_OK;
}

int fwi_autotune(fwi_ctx* c, int kind, int steps, const char* cache_path) {
    if (!c) return FWI_ERR_ARG;
    c->gen++;  // launch arguments may change: recapture pass graphs
    if (kind < 0 || kind > 3 || steps < 1) return fail(c, FWI_ERR_ARG, "kind/steps");
    if (kind == 1 && !c->b1d) return fail(c, FWI_ERR_STATE, "dampers not set");
    HIPCHK(c, hipSetDevice(c->device));
    CHKRC(sync_geometry(c));
    char key[200];
    std::snprintf(key, sizeof key, "step-spt kind=%d arith=%d nxx=%d nyy=%d nzz=%d nshots=%d", kind, c->arith,
                  c->g.nxx, c->g.nyy, c->g.nzz, c->nshots);
    if (cache_path) {
        std::ifstream in(cache_path);
        std::string line;
        while (std::getline(in, line)) {
            const size_t bar = line.find('|');
            if (bar == std::string::npos || line.substr(0, bar) != key) continue;
            std::istringstream ss(line.substr(bar + 1));
            Tuning t;
            t.big = c->tune[kind].big;
            if (ss >> t.by >> t.rpt >> t.zchunks >> t.sb >> t.pd >> t.tb >> t.tby >> t.trpt >> t.tzc >> t.spt &&
                c->nshots % t.sb == 0 && fwi::step_config_supported(kind, c->arith, t) &&
                (t.tb == 1 || fwi::step2_supported(kind, c->arith, two_step_tuning(t)))) {
                c->tune[kind] = t;
                if (t.tb == 2 && !c->geom_dirty) CHKRC(build_rec_planes(c, c->cell_off_h));
                return FWI_OK;
            }
        }
    }
    const bool bwd = kind >= 2;
    if (bwd) CHKRC(ensure_backward_buffers(c, kind - 2));
    // the trials run on noise, not on zero fields: a wavefield's changing
    // bits draw more power and lower the clock, and the shapes rank
    // differently on zeros (the buffers are cleared after the trials)
    {
        const long vols = c->L.vol * c->nshots;
        uint64_t k = 1;
        for (float* p : {c->f[0], c->f[1], c->f[2], c->f[3], c->r[0], c->r[1], c->r[2], c->r[3]})
            if (p) HIPCHK(c, fwi::launch_fill_noise(p, vols, 0x5EED0000ull + k++, c->st));
    }
    // backward trials time a multiple of six steps: with deferred imaging a
    // pass repeats the image update every other launch and the illumination
    // update every second or third, and another count would add updates at
    // the end that whole passes have only once
    if (bwd) steps = (steps + 5) / 6 * 6;
    steps += steps % 2;  // whole two-step launches
    const Tuning saved = c->tune[kind];
    const bool saved_graph = c->use_graph, saved_kt = c->ktime;
    c->use_graph = false;
    c->ktime = false;
    double best = 1e300;  // ms per step
    Tuning best_t = saved;
    const int zc_choices[] = {1, 2, 4, 8, 16};
    // shots per block: 1, 2, 4 groups of waves; 2 as one shot pair per thread (sb 3 below)
    const int sb_choices[] = {1, 2, 4, 3};
    int rc = FWI_OK;
    // trial steps past nt: no injection, no recording, no data reads
    auto body = [&](int nsteps) {
        if (bwd) return run_backward(c, kind - 2, c->g.nt, c->g.nt + nsteps);
        return run_forward(c, kind == 1 ? FWI_BOUNDARY_ABC : FWI_BOUNDARY_RBC, c->g.nt, c->g.nt + nsteps, false);
    };
    // a candidate's time per step over n steps (after one untimed step)
    auto trial = [&](const Tuning& t, int n, double* ms) {
        c->tune[kind] = t;
        int r = timed_pass(c, 1, [&] { return body(1); });
        if (!r) r = timed_pass(c, n, [&] { return body(n); });
        *ms = c->last_ms / n;
        return r;
    };
    // screen every candidate with a short trial, then time the four best
    // again with four times the steps, twice, and keep the fastest: the
    // short trials alone mis-rank shapes within a few per cent of each other
    std::vector<std::pair<double, Tuning>> screened;
    for (int sbc : sb_choices) {
        const int sb = sbc == 3 ? 2 : sbc;
        if (c->nshots % sb != 0) continue;
        for (auto& tc : kTileChoices) {
            for (int pd = 1; pd <= 2; ++pd)
            for (int zc : zc_choices) {
                Tuning t;
                t.by = tc[0];
                t.rpt = tc[1];
                t.zchunks = zc;
                t.sb = sb;
                t.spt = sbc == 3 ? 2 : 1;
                t.pd = pd;
                t.big = saved.big;
                t.tb = 1;
                if (!fwi::step_config_supported(kind, c->arith, t) || zc > c->g.nzz - 8) continue;
                if (t.spt == 2 && !pairs_fit(c)) continue;
                double ms = 0.0;
                if ((rc = trial(t, steps, &ms))) break;
                screened.push_back({ms, t});
            }
            if (rc) break;
        }
        if (rc) break;
    }
    if (!rc) {
        std::stable_sort(screened.begin(), screened.end(),
                         [](const std::pair<double, Tuning>& x, const std::pair<double, Tuning>& y) {
                             return x.first < y.first;
                         });
        for (size_t i = 0; i < screened.size() && i < 4 && !rc; ++i)
            for (int rep = 0; rep < 2 && !rc; ++rep) {
                double ms = 0.0;
                if ((rc = trial(screened[i].second, 4 * steps, &ms))) break;
                if (ms < best) {
                    best = ms;
                    best_t = screened[i].second;
                }
            }
    }
    // two-step launches (random-boundary forward, RTM backward): tiles and z
    // chunks, kept when faster than the best one-step configuration; the
    // leftover step of an odd segment runs that configuration
    if (!rc && (kind == fwi::FWD_RBC || kind == fwi::BWD_RTM) && best_t.sb == 1 && !best_t.big &&
        !(kind == fwi::BWD_RTM && c->phys)) {
        for (float** p : {&c->f[2], &c->f[3], &c->r[2], &c->r[3]})
            if (kind == fwi::BWD_RTM || p == &c->f[2] || p == &c->f[3])
                if ((rc = alloc_slots(c, p))) break;
        Tuning base = best_t;
        base.tb = 2;
        c->tune[kind] = base;
        if (!rc) rc = build_rec_planes(c, c->cell_off_h);
        const int tiles2[][2] = {{16, 2}, {12, 3}};
        const int zc2[] = {1, 2, 3, 4, 5, 6, 8};
        for (auto& tc : tiles2) {
            if (rc || !c->rp_ok) break;
            for (int zc : zc2) {
                Tuning t = base;
                t.tby = tc[0];
                t.trpt = tc[1];
                t.tzc = zc;
                if (!fwi::step2_supported(kind, c->arith, two_step_tuning(t)) || zc > c->g.nzz - 8) continue;
                double ms = 0.0;
                if ((rc = trial(t, 4 * steps, &ms))) break;
                if (ms < best) {
                    best = ms;
                    best_t = t;
                }
            }
        }
    }
    c->tune[kind] = rc ? saved : best_t;
    c->use_graph = saved_graph;
    c->ktime = saved_kt;
    // the trial steps ran on the live buffers: clear them
    const long vols = c->L.vol * c->nshots;
    for (float* p : {c->f[0], c->f[1], c->f[2], c->f[3], c->r[0], c->r[1], c->r[2], c->r[3], c->image, c->illum,
                     c->grad})
        if (p) HIPCHK(c, hipMemsetAsync(p, 0, vols * sizeof(float), c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->fwd_done = 0;
    if (rc) return rc;
    if (cache_path) {
        std::ofstream out(cache_path, std::ios::app);
        out << key << "|" << best_t.by << " " << best_t.rpt << " " << best_t.zchunks << " " << best_t.sb << " "
            << best_t.pd << " " << best_t.tb << " " << best_t.tby << " " << best_t.trpt << " " << best_t.tzc << " "
            << best_t.spt << " " << best << "\n";
    }
    return FWI_OK;
}

int fwi_set_graph(fwi_ctx* c, int enable) {
    if (!c) return FWI_ERR_ARG;
    c->use_graph = enable != 0;
    if (!c->use_graph) drop_graphs(c);
    return FWI_OK;
}

int fwi_graph_instantiations(fwi_ctx* c, int* count) {
    if (!c || !count) return FWI_ERR_ARG;
    *count = c->graph_instantiations;
    return FWI_OK;
}

int fwi_last_pass_timing(fwi_ctx* c, double* ms, int* launches) {
    if (!c || !ms || !launches) return FWI_ERR_ARG;
    *ms = c->last_ms;
    *launches = c->last_launches;
    return FWI_OK;
}

int fwi_set_kernel_timing(fwi_ctx* c, int enable) {
    if (!c) return FWI_ERR_ARG;
    c->ktime = enable != 0;
    return FWI_OK;
}

int fwi_last_kernel_timing(fwi_ctx* c, double* ms, int* launches) {
    if (!c || !ms || !launches) return FWI_ERR_ARG;
    *ms = c->kernel_ms;
    *launches = c->kev_used;
    return FWI_OK;
}

int fwi_copy_bandwidth(fwi_ctx* c, long bytes, int reps, double* gbps) {
    if (!c || !gbps || bytes < 1024 || reps < 1) return c ? fail(c, FWI_ERR_ARG, "bytes/reps") : FWI_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    const long n = bytes / 16 * 4;
    float *a = nullptr, *b = nullptr;
    HIPCHK(c, hipMalloc((void**)&a, n * sizeof(float)));
    hipError_t e = hipMalloc((void**)&b, n * sizeof(float));
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, n * sizeof(float), c->st);
    if (e == hipSuccess) e = fwi::launch_copy(a, b, n, c->st);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(c->ev0, c->st);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = fwi::launch_copy(a, b, n, c->st);
    if (e == hipSuccess) e = hipEventRecord(c->ev1, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
    (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) return fail(c, FWI_ERR_HIP, std::string("copy bandwidth: ") + hipGetErrorString(e));
    *gbps = 2.0 * n * sizeof(float) * reps / (ms * 1e-3) / 1e9;
    return FWI_OK;
}

int fwi_sync(fwi_ctx* c) {
    if (!c) return FWI_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return FWI_OK;
}

}  // extern "C"

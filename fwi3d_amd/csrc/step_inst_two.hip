// step_inst_two.hip -- instances and launcher of the two-step kernel
// (step2_kernel.hpp)
#include "step2_kernel.hpp"

namespace fwi {

namespace {

#define FWI_E2T(M, A, BY, RPT) {M, A, BY, RPT, &launch_two<M, A, BY, RPT>}

const Entry2* find2(int mode, int arith, const Tuning& t) {
    static const Entry2 table[] = {
        FWI_E2T(FWD_RBC, ARITH_REF, 16, 2), FWI_E2T(BWD_RTM, ARITH_REF, 16, 2),
        FWI_E2T(FWD_RBC, ARITH_FP32, 16, 2), FWI_E2T(BWD_RTM, ARITH_FP32, 16, 2),
        FWI_E2T(FWD_RBC, ARITH_REF, 12, 3), FWI_E2T(FWD_RBC, ARITH_FP32, 12, 3), FWI_E2T(BWD_RTM, ARITH_REF, 12, 3),
        FWI_E2T(BWD_RTM, ARITH_FP32, 12, 3),
    };
    for (const Entry2& e : table)
        if (e.mode == mode && e.arith == arith && e.by == t.by && e.rpt == t.rpt) return &e;
    return nullptr;
}

}  // namespace

bool step2_supported(int mode, int arith, const Tuning& t) {
    return t.sb == 1 && t.big == 0 && find2(mode, arith, t) != nullptr;
}

hipError_t launch_step2(int mode, int arith, const Tuning& t, const Step2Args& a, hipStream_t st) {
    const Entry2* e = find2(mode, arith, t);
    if (!e || t.sb != 1 || t.big) return hipErrorInvalidConfiguration;
    Step2Args args = a;
    const int TY = t.by * t.rpt;
    args.ntx = (a.L.nx - 8 + 55) / 56;
    args.nty = (a.L.ny - 8 + TY - 9) / (TY - 8);
    const int M = a.L.nz - 8;
    args.zlen = (M + t.zchunks - 1) / t.zchunks;
    args.nzc = (M + args.zlen - 1) / args.zlen;
    args.nitems = args.ntx * args.nty * args.nzc * a.nshots;
    const int blocks = (args.nitems + 7) / 8 * 8;
    return e->fn(args, dim3(blocks), st);
}

}  // namespace fwi

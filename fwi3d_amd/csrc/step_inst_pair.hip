// step_inst_pair.hip -- batched step instances with two shots per thread
// (pair_kernel.hpp): forward (random boundary and absorbing boundary) and
// backward (RTM, FWI) in the reference arithmetic, tiles of 64 x BY.
#include "pair_kernel.hpp"

namespace fwi {

// per-step imaging, deferred imaging (IMG 0 and 2), image every other launch
// and illumination every third (3 and 5: RTM only; 4), no corrections
#define FWI_EP(M, BY)                                                                                    \
    {M, ARITH_REF, BY, 1, 2, 1, 0, &launch_pair<M, ARITH_REF, BY, 1>, nullptr,                            \
     is_bwd(M) ? &launch_pair<M, ARITH_REF, BY, is_bwd(M) ? 0 : 1> : nullptr,                            \
     is_bwd(M) ? &launch_pair<M, ARITH_REF, BY, is_bwd(M) ? 2 : 1> : nullptr, nullptr, nullptr,           \
     M == BWD_RTM ? &launch_pair<M, ARITH_REF, BY, M == BWD_RTM ? 3 : 1> : nullptr,                      \
     is_bwd(M) ? &launch_pair<M, ARITH_REF, BY, is_bwd(M) ? 4 : 1> : nullptr,                            \
     M == BWD_RTM ? &launch_pair<M, ARITH_REF, BY, M == BWD_RTM ? 5 : 1> : nullptr,                      \
     nullptr, nullptr, nullptr, 2}
#define FWI_PAIRS(M) FWI_EP(M, 4), FWI_EP(M, 8), FWI_EP(M, 16)

// host function: the table holds host launch functions
const Entry* step_table_pair(int* count) {
    static const Entry table[] = {
        FWI_PAIRS(FWD_RBC),
        FWI_PAIRS(FWD_ABC),
        FWI_PAIRS(BWD_RTM),
        FWI_PAIRS(BWD_FWI),
    };
    *count = (int)(sizeof(table) / sizeof(Entry));
    return table;
}

}  // namespace fwi

// step_inst_prefetch.hip -- single-shot step instances with two planes of
// loads in flight (forward, RTM and FWI backward; reference and FP32
// arithmetic)
#include "step_kernel.hpp"

namespace fwi {

// host function: the table holds host launch functions
const Entry* step_table_prefetch(int* count) {
    static const Entry table[] = {
        FWI_TILES2(FWD_RBC, ARITH_REF), FWI_E2(FWD_RBC, ARITH_REF, 16, 2), FWI_TILES2(FWD_ABC, ARITH_REF),
        FWI_TILES2(BWD_RTM, ARITH_REF), FWI_TILES2(BWD_FWI, ARITH_REF),
        FWI_TILES2(FWD_RBC, ARITH_FP32), FWI_E2(FWD_RBC, ARITH_FP32, 16, 2), FWI_TILES2(FWD_ABC, ARITH_FP32),
        FWI_TILES2(BWD_RTM, ARITH_FP32), FWI_TILES2(BWD_FWI, ARITH_FP32),
    };
    *count = (int)(sizeof(table) / sizeof(Entry));
    return table;
}

}  // namespace fwi

// step_inst_big.hip -- step instances for padded volumes of 2 GiB and more
// (per-plane buffer addressing; one shot per block)
#include "step_kernel.hpp"

namespace fwi {

// host function: the table holds host launch functions
const Entry* step_table_big(int* count) {
    static const Entry table[] = {
        FWI_TILESB(FWD_RBC, ARITH_REF), FWI_TILESB(FWD_ABC, ARITH_REF), FWI_TILESB(BWD_RTM, ARITH_REF),
        FWI_TILESB(BWD_FWI, ARITH_REF), FWI_TILESB(FWD_RBC, ARITH_FP32), FWI_TILESB(FWD_ABC, ARITH_FP32),
        FWI_TILESB(BWD_RTM, ARITH_FP32), FWI_TILESB(BWD_FWI, ARITH_FP32),
    };
    *count = (int)(sizeof(table) / sizeof(Entry));
    return table;
}

}  // namespace fwi

// step_inst_fwd.hip -- single-shot forward step instances
#include "step_kernel.hpp"

namespace fwi {

// host function: the table holds host launch functions
const Entry* step_table_fwd(int* count) {
    static const Entry table[] = {
        FWI_TILES(FWD_RBC, ARITH_REF, 1),
        FWI_TILES(FWD_ABC, ARITH_REF, 1),
        FWI_TILES(FWD_RBC, ARITH_FP32, 1),
        FWI_TILES(FWD_ABC, ARITH_FP32, 1),
    };
    *count = (int)(sizeof(table) / sizeof(Entry));
    return table;
}

}  // namespace fwi

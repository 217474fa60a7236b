// step_inst_fwi.hip -- single-shot FWI backward step instances
#include "step_kernel.hpp"

namespace fwi {

// host function: the table holds host launch functions
const Entry* step_table_fwi(int* count) {
    static const Entry table[] = {
        FWI_TILES(BWD_FWI, ARITH_REF, 1),
        FWI_TILES(BWD_FWI, ARITH_FP32, 1),
    };
    *count = (int)(sizeof(table) / sizeof(Entry));
    return table;
}

}  // namespace fwi

// step_inst_rtm.hip -- single-shot RTM backward step instances
#include "step_kernel.hpp"

namespace fwi {

// host function: the table holds host launch functions
const Entry* step_table_rtm(int* count) {
    static const Entry table[] = {
        FWI_TILES(BWD_RTM, ARITH_REF, 1),
        FWI_TILES(BWD_RTM, ARITH_FP32, 1),
    };
    *count = (int)(sizeof(table) / sizeof(Entry));
    return table;
}

}  // namespace fwi

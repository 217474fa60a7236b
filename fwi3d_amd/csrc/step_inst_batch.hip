// step_inst_batch.hip -- batched (several shots per block) step instances
#include "step_kernel.hpp"

namespace fwi {

// host function: the table holds host launch functions
const Entry* step_table_batch(int* count) {
    static const Entry table[] = {
        FWI_BATCH(FWD_RBC),
        FWI_BATCH(FWD_ABC),
        FWI_BATCH(BWD_RTM),
        FWI_BATCH(BWD_FWI),
    };
    *count = (int)(sizeof(table) / sizeof(Entry));
    return table;
}

}  // namespace fwi

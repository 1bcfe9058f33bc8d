// Instantiations of the stacking kernels for f16 frames (stack_body.hpp).
#include "stack_inst_16.hpp"

namespace ftk {

SingleFn single_kernel_f16(bool ring, int pre) {
    return ring ? inst16::single16<kStoreF16, true>(pre) : inst16::single16<kStoreF16, false>(pre);
}

MultiFn multi_kernel_f16(bool ring, int pre) {
    return ring ? inst16::multi16<kStoreF16, true>(pre) : inst16::multi16<kStoreF16, false>(pre);
}

QueueFn queue_kernel_f16(int nw, int pre) { return inst16::queue16_nw<kStoreF16>(nw, pre); }

}  // namespace ftk

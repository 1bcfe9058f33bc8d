// Instantiations of the stacking kernels for fp32 frames (stack_body.hpp).
#include "stack_body.hpp"
#include "stack_inst.hpp"

namespace ftk {

SingleFn single_kernel_f32(int kind, bool padded) {
    if (kind == kArgsLat) return stack_kernel<4, false, false, kStoreF32, 0, true>;
    if (kind == kArgsGen) return stack_kernel<4, false, false, kStoreF32, 0>;
    return padded ? stack_kernel<4, true, true, kStoreF32, 0> : stack_kernel<4, true, false, kStoreF32, 0>;
}

MultiFn multi_kernel_f32(int kind, bool padded) {
    if (kind == kArgsLat) return stack_multi_kernel<4, false, false, kStoreF32, 0, true>;
    if (kind == kArgsGen) return stack_multi_kernel<4, false, false, kStoreF32, 0>;
    return padded ? stack_multi_kernel<4, true, true, kStoreF32, 0> : stack_multi_kernel<4, true, false, kStoreF32, 0>;
}

QueueFn queue_kernel_f32(int nw, bool padded) {
    switch (nw) {
        case 2: return padded ? stack_queue_kernel<2, true, kStoreF32, 0> : stack_queue_kernel<2, false, kStoreF32, 0>;
        case 4: return padded ? stack_queue_kernel<4, true, kStoreF32, 0> : stack_queue_kernel<4, false, kStoreF32, 0>;
        case 8: return padded ? stack_queue_kernel<8, true, kStoreF32, 0> : stack_queue_kernel<8, false, kStoreF32, 0>;
        default: return nullptr;
    }
}

}  // namespace ftk

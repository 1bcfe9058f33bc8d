// Instantiation helpers of the stacking kernels for 16-bit frames (stack_body.hpp),
// shared by stack_inst_f16.hip and stack_inst_bf16.hip: box rows staged per wavefront
// (pre) 5, 6, 8 or 12; 16-bit frames are always in a zero-bordered buffer.
#pragma once

#include "stack_body.hpp"
#include "stack_inst.hpp"

namespace ftk {

namespace inst16 {
template <int ST, bool RING>
SingleFn single16(int pre) {
    return pre <= 5   ? stack_kernel<4, RING, true, ST, 5>
           : pre <= 6 ? stack_kernel<4, RING, true, ST, 6>
           : pre <= 8 ? stack_kernel<4, RING, true, ST, 8>
                      : stack_kernel<4, RING, true, ST, 12>;
}
template <int ST, bool RING>
MultiFn multi16(int pre) {
    return pre <= 5   ? stack_multi_kernel<4, RING, true, ST, 5>
           : pre <= 6 ? stack_multi_kernel<4, RING, true, ST, 6>
           : pre <= 8 ? stack_multi_kernel<4, RING, true, ST, 8>
                      : stack_multi_kernel<4, RING, true, ST, 12>;
}
template <int ST, int NW>
QueueFn queue16(int pre) {
    return pre <= 5   ? stack_queue_kernel<NW, true, ST, 5>
           : pre <= 6 ? stack_queue_kernel<NW, true, ST, 6>
           : pre <= 8 ? stack_queue_kernel<NW, true, ST, 8>
                      : stack_queue_kernel<NW, true, ST, 12>;
}
template <int ST>
QueueFn queue16_nw(int nw, int pre) {
    switch (nw) {
        case 2: return queue16<ST, 2>(pre);
        case 4: return queue16<ST, 4>(pre);
        case 8: return queue16<ST, 8>(pre);
        default: return nullptr;
    }
}
}  // namespace inst16

}  // namespace ftk

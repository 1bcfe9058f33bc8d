// The stacking kernels' instantiations (stack_inst_f32.hip, stack_inst_f16.hip,
// stack_inst_bf16.hip): the
// launchers pick a kernel through these functions by work kind, frame storage, padded
// frames, 16-bit rows per wavefront (pre) and wavefronts per workgroup (nw: tiles of
// 4 nw rows).
#pragma once

#include "stack_launch.hpp"

namespace ftk {

using SingleFn = void (*)(StackArgs);
using MultiFn = void (*)(const StackArgs*, unsigned);
using QueueFn = void (*)(QueueDesc);

// one-workgroup-per-item kernels: 4 wavefronts (16-row tiles)
SingleFn single_kernel_f32(int kind, bool padded);
MultiFn multi_kernel_f32(int kind, bool padded);
SingleFn single_kernel_f16(bool ring, int pre);
MultiFn multi_kernel_f16(bool ring, int pre);
SingleFn single_kernel_bf16(bool ring, int pre);
MultiFn multi_kernel_bf16(bool ring, int pre);
// persistent kernels: nw = 2, 4 or 8 (null for another nw)
QueueFn queue_kernel_f32(int nw, bool padded);
QueueFn queue_kernel_f16(int nw, int pre);
QueueFn queue_kernel_bf16(int nw, int pre);

inline SingleFn single_kernel_16(int storage, bool ring, int pre) {
    return storage == kStoreF16 ? single_kernel_f16(ring, pre) : storage == kStoreBF16 ? single_kernel_bf16(ring, pre) : nullptr;
}
inline MultiFn multi_kernel_16(int storage, bool ring, int pre) {
    return storage == kStoreF16 ? multi_kernel_f16(ring, pre) : storage == kStoreBF16 ? multi_kernel_bf16(ring, pre) : nullptr;
}
inline QueueFn queue_kernel_16(int storage, int nw, int pre) {
    return storage == kStoreF16 ? queue_kernel_f16(nw, pre) : storage == kStoreBF16 ? queue_kernel_bf16(nw, pre) : nullptr;
}

}  // namespace ftk

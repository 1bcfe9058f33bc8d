// Launchers of the gfx950 stacking kernels (stack_kernel.hip), called by the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "stack_plan.hpp"

namespace ftk {

// Frame storage of a stacking pass: fp32, or 16-bit (fp16 / bf16) values q with a
// power-of-two scale per frame, decoded exactly as q * scale[i] before the fp32 sum.
constexpr int kStoreF32 = 0;
constexpr int kStoreF16 = 1;
constexpr int kStoreBF16 = 2;

// One launch of the stacking kernel over all tiles of output rows [y_begin, y_end)
// and the block segments [seg0, seg0 + nseg) (bps blocks each). The frames buffer
// holds frame rows [row0, row1) of every frame, frame i at frames + i * frame_stride.
//
// Frame layout: pixel (x, y) of frame i is frames[i * frame_stride + (y - row0) * pitch
// + x] (pitch 0 = w). A padded buffer (padded = true) holds zeros around the frame
// rows, laid out as pad_geometry says, so that every box a tile stages lies inside it.
struct StackLaunch {
    const void* frames = nullptr;  // float, or uint16_t for 16-bit storage (always padded)
    int storage = kStoreF32;
    const float* scale = nullptr;  // per frame, 16-bit storage only
    long long frame_stride = 0;
    int row0 = 0, row1 = 0;
    int pitch = 0;
    bool padded = false;
    const int* shift = nullptr;
    const int* box = nullptr;
    const int* ubox = nullptr;
    const int* lds_off = nullptr;
    const long long* cov_off = nullptr;  // per block: offsets into cov_x, cov_y
    const int* cov_x = nullptr;
    const int* cov_y = nullptr;
    const int* chain = nullptr;          // per block and motion: chain / anti-chain flags
    const int* lat = nullptr;            // lattice blocks' column and row offsets per frame (plan.lat)
    const int* motion = nullptr;
    const float* inv_sqrt = nullptr;
    int y_begin = 0, y_end = 0, tiles_x = 0, tiles_y = 0;
    int bps = 1, seg0 = 0, nseg = 1;
    unsigned long long* keys = nullptr;  // (y_end - y_begin) x w packed maxima
};

// Tile coordinates of a launch: interior tiles (every block's reads inside the frame
// rows present) form the rectangle [tx_lo, tx_hi) x [ty_lo, ty_hi); the ring lists
// the other tiles row by row.
struct TileRect {
    int tiles_x, tiles_y, tx_lo, tx_hi, ty_lo, ty_hi;
};

// Everything one workgroup of a pass reads: the kernel arguments of a single-pass
// launch, or one descriptor of a multi-pass launch (read with scalar loads).
struct StackArgs {
    const void* frames;
    const float* scale;
    long long frame_stride;
    const int* shift;
    const int* box;
    const int* ubox;
    const int* lds_off;
    const int* motion;
    const float* inv_sqrt;
    const long long* cov_off;
    const int* cov_x;
    const int* cov_y;
    const int* chain;
    const int* lat;
    unsigned long long* keys;
    TileRect rect;
    int row0, row1, n, w, h, pitch, padded, storage, min_count, nb, bps;
    int y_begin, y_end, seg0, nseg, ntiles, lds_rows, cov_ok;
    int pre;  // 16-bit frames: box rows staged per wavefront (5, 8 or 12)
    int blk_lo, blk_hi;  // the blocks this launch stacks (of each segment's range)
};

// The descriptors of one pass: its lattice blocks on interior tiles (the lattice
// kernel), its other blocks on interior tiles, and all blocks on the ring tiles. A kind
// with ntiles 0 or nseg 0 has no launch.
enum { kArgsLat = 0, kArgsGen = 1, kArgsRing = 2, kArgsKinds = 3 };

hipError_t launch_copy_rows(const float* src, long long src_stride, float* dst, long long dst_stride, int dst_pitch,
                            int w, int rows, int n, hipStream_t s);

// The three descriptors of one launch (see kArgsLat / kArgsGen / kArgsRing). Lattice
// blocks get their own launch for fp32 frames with plan.lat present (L.lat set).
void make_stack_args(const StackPlan& p, const StackLaunch& L, StackArgs out[kArgsKinds]);

// ring_s (optional): the stream of the edge tiles' launch and of the general blocks'
// interior launch, beside the lattice launch on s; ordered by the caller (every launch
// merges into the same keys by atomic maximum, so any order gives the same bits).
hipError_t launch_stack(const StackPlan& p, const StackLaunch& L, hipStream_t s, hipStream_t ring_s = nullptr);

// Several passes in up to three launches (lattice interior, general interior, ring).
// host[kind] / dev[kind]: count descriptors of that kind each, as make_stack_args writes
// them, on the host (for the grid sizes) and in device memory. The descriptors of one
// kind must have the same lds_rows. ring_s (optional): the ring tiles' stream, ordered by the
// caller (the launches merge into the same keys by atomic maximum, so any order gives
// the same bits).
hipError_t launch_stack_multi(const StackArgs* const host[kArgsKinds], const StackArgs* const dev[kArgsKinds], int count,
                              hipStream_t s, hipStream_t ring_s = nullptr);
hipError_t launch_fill_keys(unsigned long long* keys, long long count, hipStream_t s);
hipError_t launch_finalize(const unsigned long long* keys, long long count, float* max_snr, int* best_v,
                           hipStream_t s);
hipError_t launch_copy(const float* src, float* dst, long long count, hipStream_t s);

}  // namespace ftk

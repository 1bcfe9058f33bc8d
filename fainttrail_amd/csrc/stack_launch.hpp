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
    // persistent form (launch_stack with counters set): every kind of the pass in one
    // launch taking work items from 8 device counters (cleared by the launch); counters
    // points at kQueueWorkspaceBytes of device memory (counters, then descriptors)
    int* counters = nullptr;
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
    int pre;  // 16-bit frames: box rows staged per wavefront (5, 6, 8 or 12)
    int blk_lo, blk_hi;  // the blocks this launch stacks (of each segment's range)
    int tile_rows;       // 8, 16 or 32: 4 rows per wavefront of the kernel
};

// The descriptors of one pass: its lattice blocks on interior tiles (the lattice
// kernel), its other blocks on interior tiles, and all blocks on the ring tiles. A kind
// with ntiles 0 or nseg 0 has no launch.
enum { kArgsLat = 0, kArgsGen = 1, kArgsRing = 2, kArgsKinds = 3 };

// The kernel argument of a persistent launch (stack_queue_kernel): per kind, the tiles
// and segments of each of `count` passes (the passes of one launch share their plan);
// the descriptor of pass p and kind k is args[k * stride + p] in device memory;
// counters: the 8 queue counters.
struct QueueDesc {
    const StackArgs* args;
    int* counters;
    int count, stride;
    int ntiles[kArgsKinds], nseg[kArgsKinds];
};

// Device workspace of a persistent one-pass launch (StackLaunch::counters): the 8 queue
// counters, then the pass's three descriptors. 256-byte aligned.
constexpr size_t kQueueWorkspaceBytes = 4096;

hipError_t launch_copy_rows(const float* src, long long src_stride, float* dst, long long dst_stride, int dst_pitch,
                            int w, int rows, int n, hipStream_t s);

// The three descriptors of one launch (see kArgsLat / kArgsGen / kArgsRing). Lattice
// blocks get their own launch for fp32 frames with plan.lat present (L.lat set). The
// tiles are the plan's (p.tile_rows rows).
void make_stack_args(const StackPlan& p, const StackLaunch& L, StackArgs out[kArgsKinds]);
// The work shape of a stacking pass: tile height and launch form (queue 1: persistent
// launches; 0: one workgroup per work item, 16-row tiles only). No result bit depends
// on it.
struct StackShape {
    int tile_rows = 16;
    int queue = 1;
};
// The shape for asked values (tile_rows 0 = automatic, queue -1 = automatic) of a pass
// of out_rows output rows in this storage. Automatic: 16-row tiles; the persistent
// launch for 16-bit frames, and for fp32 frames of at least 1024 tiles (measured on
// MI355X: faster there, a few % slower on smaller frames, whose edge tiles run at one
// wavefront per SIMD more in their own launch). A tile height that the 16-bit staging
// cannot hold for this plan becomes 16. Throws on a tile height other than 16 with
// queue 0.
StackShape resolve_shape(const StackPlan& p, int storage, int tile_rows, int queue, int out_rows);
// 16-bit frames: box rows staged per wavefront for tiles of tile_rows rows (5, 6, 8 or
// 12), 0 when the plan's vertical spread needs more.
int rows16_per_wave(const StackPlan& p, int tile_rows);
// Whether the kernels can stack the plan with tiles of tile_rows rows in this storage.
bool tile_rows_fit(const StackPlan& p, int storage, int tile_rows);

// One pass. With L.counters set: one persistent launch on s (any tile height; see
// launch_stack_queue). Otherwise up to three launches (16-row tiles); ring_s
// (optional): the stream of the edge tiles' launch and of the general blocks' interior
// launch, beside the lattice launch on s; ordered by the caller (every launch merges
// into the same keys by atomic maximum, so any order gives the same bits).
hipError_t launch_stack(const StackPlan& p, const StackLaunch& L, hipStream_t s, hipStream_t ring_s = nullptr);

// Several passes in up to three launches (lattice interior, general interior, ring).
// host[kind] / dev[kind]: count descriptors of that kind each, as make_stack_args writes
// them, on the host (for the grid sizes) and in device memory. The descriptors of one
// kind must have the same lds_rows. ring_s (optional): the ring tiles' stream, ordered by the
// caller (the launches merge into the same keys by atomic maximum, so any order gives
// the same bits).
hipError_t launch_stack_multi(const StackArgs* const host[kArgsKinds], const StackArgs* const dev[kArgsKinds], int count,
                              hipStream_t s, hipStream_t ring_s = nullptr);
// Several passes in ONE persistent launch of about as many workgroups as the GPU holds
// at once, which take (pass, kind, tile, segment) work items from the 8 device counters
// (XCD-aware queues, see stack_queue_kernel); the counters are cleared on s first.
// host / dev: descriptors kind-major, stride apart (host for the shapes, dev for the
// kernel); every pass of the launch must have the same tiles and segments per kind.
hipError_t launch_stack_queue(const StackArgs* host, const StackArgs* dev, int count, int stride, int* counters,
                              hipStream_t s);
hipError_t launch_fill_keys(unsigned long long* keys, long long count, hipStream_t s);
hipError_t launch_finalize(const unsigned long long* keys, long long count, float* max_snr, int* best_v,
                           hipStream_t s);
hipError_t launch_copy(const float* src, float* dst, long long count, hipStream_t s);

}  // namespace ftk

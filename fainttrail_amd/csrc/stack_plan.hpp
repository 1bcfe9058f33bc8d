// Host planning for the gfx950 stacking kernel: velocity blocks, per-frame staging
// boxes and work segments. Plain C++17 (no HIP), shared by the device launcher and
// the host emulation.
//
// Work decomposition. The output is cut into tiles of kTileX x kTileY pixels (one
// wavefront spans a tile row of 64 pixels; each of the 4 wavefronts owns
// kRowsPerWave rows). The trial motions are grouped into blocks of kVel motions
// whose shifts lie close together in every frame, so one staged frame box
// (tile + the block's shift spread) in LDS serves all kVel motions. A workgroup
// takes one tile and one segment of consecutive blocks, keeps the best S/N per
// pixel in registers across the segment, and merges it into a packed 64-bit key
// per pixel with one atomic max (value, then lowest motion index).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace ftk {

constexpr int kTileX = 64;                     // pixels per tile row = wavefront width
constexpr int kWaves = 4;                      // wavefronts per workgroup of the default tile
constexpr int kRowsPerWave = 4;                // tile rows per wavefront
constexpr int kTileY = kWaves * kRowsPerWave;  // 16: the default tile height (StackPlan::tile_rows)
constexpr int kVel = 16;                       // motions per block
constexpr int kSpreadMax = 32;                 // largest per-frame shift spread of a block, each axis
constexpr int kPitch = 128;                     // LDS row pitch (floats): rows 0..3 of a lane are 2 x 64 dwords apart (ds_read2st64)
constexpr int kLdsRows = kTileY + kSpreadMax;  // LDS rows per staged frame
constexpr int kThreads = kTileX * kWaves;      // 256
// Lattice blocks: 4 x 4 motions whose shifts factor in every frame as dx = X[a], dy = Y[b]
// (a, b = 0..3; motion slot k = 4 a + b), with a vertical spread max Y - min Y of at most
// kLatSyMax in every frame.
constexpr int kLatSide = 4;
constexpr int kLatSyMax = 3;
constexpr int kLatMinFill = 12;                // fewer motions in a lattice group: general blocks
constexpr int kLeftSpreadMax = 8;              // spread limit of the general blocks beside lattice blocks

struct StackPlan {
    int n = 0, w = 0, h = 0, nv = 0, min_count = 1;
    int nb = 0;                      // velocity blocks
    std::vector<int32_t> motion;     // nb x kVel motion indices (general blocks: ascending); -1 = padding
    std::vector<int32_t> shift;      // nb x n x (kVel dx, kVel dy), clamped to +-w, +-h
    std::vector<int32_t> box;        // nb x n x 4: min dx, min dy, spread x, spread y
    std::vector<int32_t> ubox;       // nb x 4: min dx, min dy, max dx, max dy over all frames
    std::vector<int32_t> lds_off;    // nb x n x kVel: (dy - min dy) * kPitch + (dx - min dx)
    // coverage tables: for block b and motion k, cov_x[cov_off[2b] + k * Lx + (d - min dx)]
    // = number of frames with dx <= d (Lx = max dx - min dx + 1); cov_y likewise
    std::vector<long long> cov_off;
    std::vector<int32_t> cov_x, cov_y;
    // per block and motion (nb x kVel): 1 when the motion's (dx, dy) pairs form a chain
    // ((dx_i - dx_j)(dy_i - dy_j) >= 0 for all frames i, j), 2 when they form an
    // anti-chain (<= 0), else 0. Then the frames landing inside both ranges are an
    // intersection of two intervals of one order, counted from cov_x and cov_y alone.
    std::vector<int32_t> chain;
    // Blocks [0, nb_lat) are lattice blocks. lat: nb_lat x n x 2 words per block and frame:
    // word 0 holds the four column offsets X[a] - min X in bits 8a..8a+7, word 1 the four row
    // offsets Y[b] - min Y in bits 4b..4b+3 and, in bits 16..17, the row pairs a lane reads
    // beyond its first two: (spread + 1) / 2 (0, 1 or 2).
    int nb_lat = 0;
    std::vector<int32_t> lat;
    int max_sy_lat = 0;  // largest per-frame vertical spread of a lattice block (<= kLatSyMax)
    // motions not stacked because a lower-index motion has the same shifts in every frame
    // (equal sums and counts: the lower index always wins the tie)
    int shadowed = 0;
    bool cov_ok = true;  // false: the tables would be too large; the kernel walks the frames
    int all_box[4] = {0, 0, 0, 0};   // union of ubox over the blocks: min dx, min dy, max dx, max dy
    std::vector<float> inv_sqrt;     // c = 0..n: float(1 / sqrt(double c))
    int min_dy = 0, max_dy = 0;      // extremes of the clamped vertical shifts
    int max_sy = 0;                  // largest per-frame vertical spread of a block (<= kSpreadMax)
    // tile height: 4 rows per wavefront, 2, 4 or 8 wavefronts per workgroup (8, 16 or
    // 32 rows); set before plan_tiles
    int tile_rows = kTileY;
    int tiles_x = 0, tiles_y = 0;
    int blocks_per_seg = 1, nseg = 1;
    double units = 0.0;              // nv * n * w * h
};

// Validates like validate_pass (src/backend.cpp:89-108) and builds the plan.
// Throws std::runtime_error with the reference's messages.
StackPlan make_stack_plan(int n, int w, int h, int nv, const int32_t* dx, const int32_t* dy, int min_count);

// Tile rows for output rows [y0, y1) (row bands) and the segment size for a grid, at
// the plan's tile height.
void plan_tiles(StackPlan& p, int out_rows);
// Whether a tile height is one the kernels are built for (8, 16 or 32 rows).
inline bool tile_rows_ok(int rows) { return rows == 8 || rows == 16 || rows == 32; }
constexpr int kTileRowsMax = 32;

// Packed (value, lowest index) key: larger key = larger S/N, then lower index.
// Never called with NaN; 0 means "no defined motion".
inline uint32_t ordered_bits(float f) {
    uint32_t u;
    static_assert(sizeof(u) == sizeof(f), "float size");
    __builtin_memcpy(&u, &f, sizeof u);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
inline float from_ordered_bits(uint32_t o) {
    const uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    float f;
    __builtin_memcpy(&f, &u, sizeof f);
    return f;
}
inline uint64_t pack_key(float snr, int32_t j) {
    return (static_cast<uint64_t>(ordered_bits(snr)) << 32) | static_cast<uint32_t>(~static_cast<uint32_t>(j));
}

// Row bands whose frame rows fit a byte budget (the reference's plan_bands,
// src/cuda_stack_core.hpp:90-106, same rule).
struct Band {
    int y0 = 0, y1 = 0;  // output rows
    int r0 = 0, r1 = 0;  // frame rows needed
};
std::vector<Band> plan_row_bands(int w, int h, int n, int min_dy, int max_dy, double budget_bytes);

// Zero border of a padded frame buffer for frame rows [row0, row1) and the tiles of
// output rows from y_begin: `left` zero columns, rows of `pitch` floats, `top` zero
// rows above row0, `rows` rows in all; frames points at (x 0, row0), i.e. top * pitch +
// left floats into a frame's buffer. The rows hold the tiles of any tile height over
// the same output rows (the layout does not change with the tile height).
struct PadGeometry {
    int left = 0, pitch = 0, top = 0, rows = 0;
};
PadGeometry pad_geometry(const StackPlan& p, int row0, int row1, int y_begin, int tiles_x, int tiles_y);
// The same for 16-bit frames, which are fetched as pixel pairs from the even column at
// or left of each box start: one more zero column each side, and an even left border,
// pitch and frame size so every pair is 4-byte aligned.
PadGeometry pad_geometry16(const PadGeometry& g32);

}  // namespace ftk

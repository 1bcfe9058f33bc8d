// gfx950 shift-and-stack kernels with per-pixel max S/N and argmax: the workgroup
// body and the kernel templates, instantiated per frame storage in stack_inst_*.hip.
//
// Contract (include/fainttrail/backend.hpp:10-40): for output pixel (x, y) and trial
// motion j, S_j = (sum over covering frames i, ascending, of g_i(x+dx_ji, y+dy_ji))
// * float(1/sqrt(c)), summed in fp32 from +0.0f; undefined when c < min_count;
// max over j with ties to the lowest j. Bit-identical to the CPU backend
// (src/backend_cpu.cpp:77-204) because
//   * every (pixel, motion) sum adds the covered samples in ascending frame order,
//     and an uncovered frame adds a zero: a sum that starts at +0.0f is never -0.0f
//     (IEEE round-to-nearest gives +0 for exact cancellation), and x + (+-0) = x for
//     every x that is not -0, so the zero is an exact identity;
//   * the build uses -ffp-contract=off, so no fused multiply-add appears;
//   * c and 1/sqrt(c) come from the same integer count and the same host table;
//   * the maximum keeps (value, lowest index), merged across workgroups with a
//     64-bit atomic max of (ordered value bits, inverted index).
//
// One work item = one tile (64 x 4 NW pixels for NW wavefronts per workgroup; each
// wavefront owns 4 rows of 64) and a segment of velocity blocks. Per block and frame,
// the box the block's shifted tiles can reach is staged in LDS (double buffered, one
// barrier per frame) and each lane reads its 16 motions x 4 rows from it. No MFMA:
// this is a gather-sum.
//
// Three kinds of work: tiles inside the plan's union shift box (every read of every
// block lies in the frame rows present) skip all bounds and coverage work (c = n); the
// ring of remaining tiles tests each block, and for blocks that reach outside the
// frame zero-fills the box, loads only in-frame pixels and counts coverage.
//
// Lattice blocks (4 x 4 motions with dx = X[a], dy = Y[b] in every frame, slot 4 a + b)
// have an interior body of their own: per frame a lane reads each of the four
// shifted columns once, 4 + the frame's vertical spread rows (at most 8), and adds
// rows c_b .. c_b + 3 of column a to motion (a, b) for the frame's row offset c_b of
// row class b, the adds' source registers offset by c_b through GPR indexing (no
// branch). That is at most 32 LDS dwords per lane and frame for 64 adds instead of 64.
//
// Kernels: stack_kernel (one pass and kind, one workgroup per work item),
// stack_multi_kernel (several passes of one kind) and stack_queue_kernel (persistent:
// every kind of several passes in one launch, work items taken from device counters).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "stack_launch.hpp"
#include "stack_plan.hpp"

namespace ftk {

namespace stack_detail {

constexpr int V = kVel;
constexpr int RPW = kRowsPerWave;

// Plan tables are read at wave-uniform addresses; through the constant address space
// the compiler issues scalar loads (s_load), which do not wait on the vector memory
// counter that the in-flight frame prefetch holds.
using cint = const __attribute__((address_space(4))) int;

// Word 3 of a raw (untyped, stride 0) buffer descriptor on gfx9-family parts: 32-bit
// data format, as AMD's own gfx90a/gfx94x/gfx950 kernels set it.
constexpr int kBufferWord3 = 0x00020000;
__device__ inline cint* as_const(const int* p) { return (cint*)(p); }

__device__ inline uint32_t d_ordered(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float d_unordered(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

__device__ inline void tile_xy(const TileRect& R, bool ring, int t, int& tx, int& ty) {
    if (!ring) {
        const int wdt = R.tx_hi - R.tx_lo;
        tx = R.tx_lo + t % wdt;
        ty = R.ty_lo + t / wdt;
        return;
    }
    const int top = R.ty_lo * R.tiles_x;
    const int per = R.tx_lo + (R.tiles_x - R.tx_hi);
    const int mid = (R.ty_hi - R.ty_lo) * per;
    if (t < top) {
        ty = t / R.tiles_x;
        tx = t % R.tiles_x;
    } else if (t < top + mid) {
        const int u = t - top;
        ty = R.ty_lo + u / per;
        const int c = u % per;
        tx = c < R.tx_lo ? c : R.tx_hi + (c - R.tx_lo);
    } else {
        const int u = t - top - mid;
        ty = R.ty_hi + u / R.tiles_x;
        tx = u % R.tiles_x;
    }
}

}  // namespace stack_detail

using namespace stack_detail;

// acc + q s for the 16-bit value q in the low or high half of w and a power-of-two
// scale s: the product is exact, so the single rounding of the fused operation equals
// adding the decoded value q s. fp16: one v_fma_mix_f32 (widening inside the FMA);
// bf16: the value's fp32 bits are its 16 bits shifted up.
template <int ST>
__device__ __forceinline__ float fma_q16(uint32_t w, bool hi, float s, float acc) {
    if constexpr (ST == kStoreF16) {
        if (hi)
            asm("v_fma_mix_f32 %0, %1, %2, %0 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "+v"(acc) : "v"(w), "s"(s));
        else
            asm("v_fma_mix_f32 %0, %1, %2, %0 op_sel_hi:[1,0,0]" : "+v"(acc) : "v"(w), "s"(s));
        return acc;
    } else {
        return __builtin_fmaf(__uint_as_float(hi ? (w & 0xffff0000u) : (w << 16)), s, acc);
    }
}

// All four row classes of a lattice frame in one GPR-indexed sequence: class B adds rows
// c_B .. c_B + 3 of column a to motion (a, B). The column values sit in v40 .. v71 (column
// a, row q in v(40 + 8 a + q)); s_set_gpr_idx_on adds c_B to the source register number
// of the adds that follow, so one instruction sequence serves every row offset without a
// branch. The index is set once and changed with s_set_gpr_idx_idx between classes (one
// mode switch on and off per frame, not four). Each index setting writes M0 and is
// followed by one wait state before the first indexed add: without it an add can take
// its index from M0's previous contents (a direct-to-LDS load address), a register
// outside the wavefront's allocation.
__device__ __forceinline__ void lat_add_all(float (&acc)[kVel][kRowsPerWave], const float (&col)[kLatSide][8], int c0,
                                            int c1, int c2, int c3) {
    asm volatile(
        "s_set_gpr_idx_on %64, gpr_idx(SRC0)\n\t"
        "s_nop 0\n\t"
        "v_add_f32 %0, v40, %0\n\t"
        "v_add_f32 %1, v41, %1\n\t"
        "v_add_f32 %2, v42, %2\n\t"
        "v_add_f32 %3, v43, %3\n\t"
        "v_add_f32 %16, v48, %16\n\t"
        "v_add_f32 %17, v49, %17\n\t"
        "v_add_f32 %18, v50, %18\n\t"
        "v_add_f32 %19, v51, %19\n\t"
        "v_add_f32 %32, v56, %32\n\t"
        "v_add_f32 %33, v57, %33\n\t"
        "v_add_f32 %34, v58, %34\n\t"
        "v_add_f32 %35, v59, %35\n\t"
        "v_add_f32 %48, v64, %48\n\t"
        "v_add_f32 %49, v65, %49\n\t"
        "v_add_f32 %50, v66, %50\n\t"
        "v_add_f32 %51, v67, %51\n\t"
        "s_set_gpr_idx_idx %65\n\t"
        "s_nop 0\n\t"
        "v_add_f32 %4, v40, %4\n\t"
        "v_add_f32 %5, v41, %5\n\t"
        "v_add_f32 %6, v42, %6\n\t"
        "v_add_f32 %7, v43, %7\n\t"
        "v_add_f32 %20, v48, %20\n\t"
        "v_add_f32 %21, v49, %21\n\t"
        "v_add_f32 %22, v50, %22\n\t"
        "v_add_f32 %23, v51, %23\n\t"
        "v_add_f32 %36, v56, %36\n\t"
        "v_add_f32 %37, v57, %37\n\t"
        "v_add_f32 %38, v58, %38\n\t"
        "v_add_f32 %39, v59, %39\n\t"
        "v_add_f32 %52, v64, %52\n\t"
        "v_add_f32 %53, v65, %53\n\t"
        "v_add_f32 %54, v66, %54\n\t"
        "v_add_f32 %55, v67, %55\n\t"
        "s_set_gpr_idx_idx %66\n\t"
        "s_nop 0\n\t"
        "v_add_f32 %8, v40, %8\n\t"
        "v_add_f32 %9, v41, %9\n\t"
        "v_add_f32 %10, v42, %10\n\t"
        "v_add_f32 %11, v43, %11\n\t"
        "v_add_f32 %24, v48, %24\n\t"
        "v_add_f32 %25, v49, %25\n\t"
        "v_add_f32 %26, v50, %26\n\t"
        "v_add_f32 %27, v51, %27\n\t"
        "v_add_f32 %40, v56, %40\n\t"
        "v_add_f32 %41, v57, %41\n\t"
        "v_add_f32 %42, v58, %42\n\t"
        "v_add_f32 %43, v59, %43\n\t"
        "v_add_f32 %56, v64, %56\n\t"
        "v_add_f32 %57, v65, %57\n\t"
        "v_add_f32 %58, v66, %58\n\t"
        "v_add_f32 %59, v67, %59\n\t"
        "s_set_gpr_idx_idx %67\n\t"
        "s_nop 0\n\t"
        "v_add_f32 %12, v40, %12\n\t"
        "v_add_f32 %13, v41, %13\n\t"
        "v_add_f32 %14, v42, %14\n\t"
        "v_add_f32 %15, v43, %15\n\t"
        "v_add_f32 %28, v48, %28\n\t"
        "v_add_f32 %29, v49, %29\n\t"
        "v_add_f32 %30, v50, %30\n\t"
        "v_add_f32 %31, v51, %31\n\t"
        "v_add_f32 %44, v56, %44\n\t"
        "v_add_f32 %45, v57, %45\n\t"
        "v_add_f32 %46, v58, %46\n\t"
        "v_add_f32 %47, v59, %47\n\t"
        "v_add_f32 %60, v64, %60\n\t"
        "v_add_f32 %61, v65, %61\n\t"
        "v_add_f32 %62, v66, %62\n\t"
        "v_add_f32 %63, v67, %63\n\t"
        "s_set_gpr_idx_off"
        : "+v"(acc[0][0]), "+v"(acc[0][1]), "+v"(acc[0][2]), "+v"(acc[0][3]), "+v"(acc[1][0]), "+v"(acc[1][1]), "+v"(acc[1][2]), "+v"(acc[1][3]), "+v"(acc[2][0]), "+v"(acc[2][1]), "+v"(acc[2][2]), "+v"(acc[2][3]), "+v"(acc[3][0]), "+v"(acc[3][1]), "+v"(acc[3][2]), "+v"(acc[3][3]), "+v"(acc[4][0]), "+v"(acc[4][1]), "+v"(acc[4][2]), "+v"(acc[4][3]), "+v"(acc[5][0]), "+v"(acc[5][1]), "+v"(acc[5][2]), "+v"(acc[5][3]), "+v"(acc[6][0]), "+v"(acc[6][1]), "+v"(acc[6][2]), "+v"(acc[6][3]), "+v"(acc[7][0]), "+v"(acc[7][1]), "+v"(acc[7][2]), "+v"(acc[7][3]), "+v"(acc[8][0]), "+v"(acc[8][1]), "+v"(acc[8][2]), "+v"(acc[8][3]), "+v"(acc[9][0]), "+v"(acc[9][1]), "+v"(acc[9][2]), "+v"(acc[9][3]), "+v"(acc[10][0]), "+v"(acc[10][1]), "+v"(acc[10][2]), "+v"(acc[10][3]), "+v"(acc[11][0]), "+v"(acc[11][1]), "+v"(acc[11][2]), "+v"(acc[11][3]), "+v"(acc[12][0]), "+v"(acc[12][1]), "+v"(acc[12][2]), "+v"(acc[12][3]), "+v"(acc[13][0]), "+v"(acc[13][1]), "+v"(acc[13][2]), "+v"(acc[13][3]), "+v"(acc[14][0]), "+v"(acc[14][1]), "+v"(acc[14][2]), "+v"(acc[14][3]), "+v"(acc[15][0]), "+v"(acc[15][1]), "+v"(acc[15][2]), "+v"(acc[15][3])
        : "s"(c0), "s"(c1), "s"(c2), "s"(c3), "{v40}"(col[0][0]), "{v41}"(col[0][1]), "{v42}"(col[0][2]), "{v43}"(col[0][3]), "{v44}"(col[0][4]), "{v45}"(col[0][5]), "{v46}"(col[0][6]), "{v47}"(col[0][7]), "{v48}"(col[1][0]), "{v49}"(col[1][1]), "{v50}"(col[1][2]), "{v51}"(col[1][3]), "{v52}"(col[1][4]), "{v53}"(col[1][5]), "{v54}"(col[1][6]), "{v55}"(col[1][7]), "{v56}"(col[2][0]), "{v57}"(col[2][1]), "{v58}"(col[2][2]), "{v59}"(col[2][3]), "{v60}"(col[2][4]), "{v61}"(col[2][5]), "{v62}"(col[2][6]), "{v63}"(col[2][7]), "{v64}"(col[3][0]), "{v65}"(col[3][1]), "{v66}"(col[3][2]), "{v67}"(col[3][3]), "{v68}"(col[3][4]), "{v69}"(col[3][5]), "{v70}"(col[3][6]), "{v71}"(col[3][7])
        : "m0");
}

// PRE: for 16-bit frames, box rows staged per wavefront (the kernel is instantiated
// for the few values the planner picks from its largest vertical spread); 0 for fp32.
template <int NW, bool RING, bool PADDED, int ST, int PRE, bool LAT = false>
__device__ __forceinline__ void stack_body(const StackArgs& a, int tile, int seg) {
    // NW wavefronts per workgroup, each owning 4 rows of the tile: tiles of 64 x 4 NW
    constexpr int TY = NW * RPW;
    constexpr int kPreRows = (TY + kSpreadMax + NW - 1) / NW;  // staged rows per wavefront, at most
    static_assert(!LAT || (!RING && ST == kStoreF32), "lattice blocks: fp32 interior tiles");
    static_assert(ST == kStoreF32 || (ST == kStoreF16 || ST == kStoreBF16), "frame storage");
    static_assert(ST == kStoreF32 || !RING || PADDED, "16-bit frames are always padded");
    using elem_t = typename std::conditional<ST == kStoreF32, float, uint16_t>::type;
    const elem_t* __restrict__ frames = static_cast<const elem_t*>(a.frames);
    const float* __restrict__ scale = a.scale;
    const long long frame_stride = a.frame_stride;
    const int row0 = a.row0, row1 = a.row1, n = a.n, w = a.w, h = a.h, pitch = a.pitch;
    const int* __restrict__ shift = a.shift;
    const int* __restrict__ box = a.box;
    const int* __restrict__ ubox = a.ubox;
    const int* __restrict__ lds_off = a.lds_off;
    const int* __restrict__ motion = a.motion;
    const float* __restrict__ inv_sqrt = a.inv_sqrt;
    const int min_count = a.min_count, nb = a.nb, bps = a.bps;
    const TileRect rect = a.rect;
    const int y_begin = a.y_begin, y_end = a.y_end;
    const int lds_rows = a.lds_rows, cov_ok = a.cov_ok;
    const long long* __restrict__ cov_off = a.cov_off;
    const int* __restrict__ cov_x = a.cov_x;
    const int* __restrict__ cov_y = a.cov_y;
    const int* __restrict__ chain = a.chain;
    const int* __restrict__ lat = a.lat;
    unsigned long long* __restrict__ keys = a.keys;
    // two frame boxes of lds_rows x kPitch floats (lds_rows = 16 + the plan's largest
    // vertical spread), sized at launch
    extern __shared__ float lds_dyn[];
    float* const lds0 = lds_dyn;
    float* const lds1 = lds_dyn + lds_rows * kPitch;

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    __builtin_assume(wave >= 0 && wave < NW);
    int tx, ty;
    tile_xy(rect, RING, tile, tx, ty);
    const int x0 = tx * kTileX;
    const int y0 = y_begin + ty * TY;
    const int x = x0 + lane;
    const int yw = y0 + wave * RPW;
    const float inv_n = inv_sqrt[n];
    cint* const c_box = as_const(box);
    cint* const c_ubox = as_const(ubox);
    cint* const c_off = as_const(lds_off);
    cint* const c_motion = as_const(motion);
    cint* const c_shift = as_const(shift);
    cint* const c_covy = as_const(cov_y);
    cint* const c_chain = as_const(chain);
    cint* const c_lat = as_const(lat);
    const __attribute__((address_space(4))) long long* const c_covoff =
        (const __attribute__((address_space(4))) long long*)(cov_off);
    const int lane_base = wave * RPW * kPitch + lane;
    // the tile origin in a frame (rows from row0, the first row present)
    const long long tile_org = static_cast<long long>(y0 - row0) * pitch + x0;

    float best[RPW];
    int arg[RPW];
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        best[r] = -__builtin_inff();
        arg[r] = -1;
    }

    const int blk_begin = max(seg * bps, a.blk_lo);
    const int blk_end = min(min(nb, a.blk_hi), seg * bps + bps);
    for (int blk = blk_begin; blk < blk_end; ++blk) {
        bool interior = true;
        if (RING) {
            cint* ub = c_ubox + 4 * blk;
            interior = x0 + ub[0] >= 0 && x0 + kTileX - 1 + ub[2] < w && y0 + ub[1] >= row0 &&
                       y0 + TY - 1 + ub[3] < row1;
        }

        float acc[V][RPW];
#pragma unroll
        for (int k = 0; k < V; ++k)
#pragma unroll
            for (int r = 0; r < RPW; ++r) acc[k][r] = 0.0f;

        // ---- staging of frame i's box into LDS buffer `dst` with gfx950's direct
        // global -> LDS loads, one 64-float row chunk per wavefront instruction (lane l
        // writes the chunk's l-th float). In a padded frame buffer (a zero border wide
        // enough for every block's box) every box is loaded whole. Otherwise a block that
        // is not interior to this tile zeroes its box first and then loads only the lanes
        // whose pixel lies in the frame rows present and in [0, w). Either way the zeros
        // stand for frames that do not cover.
        // bx: the frame's box (x and y offset from the tile, extra columns, extra rows),
        // ft: the frame at the tile origin (frame i + (y0 - row0) pitch + x0)
        auto stage = [&](int i, const int (&bx)[4], const float* ft, float* dst) {
            // the plan's extra rows are max - min of the block's vertical shifts (at most
            // kLatSyMax for a lattice block): the first TY rows are always staged, so only
            // the rows past them need the count test
            __builtin_assume(bx[3] >= 0);
            if constexpr (LAT) __builtin_assume(bx[3] <= kLatSyMax);
            const int sx = bx[2], R = TY + bx[3];
            const int bx0 = x0 + bx[0], by0 = y0 + bx[1];
            if (RING && !PADDED && !interior) {
#pragma unroll
                for (int m = 0; m < kPreRows; ++m) {
                    const int rr = wave + NW * m;
                    if (rr >= R) break;
                    dst[rr * kPitch + lane] = 0.0f;
                    if (lane < kPitch - kTileX) dst[rr * kPitch + kTileX + lane] = 0.0f;
                }
                __builtin_amdgcn_s_waitcnt(0);  // the zeros land before the loads
            }
            if (RING && !PADDED && !interior) {
                const float* f = static_cast<const float*>(a.frames) + static_cast<long long>(i) * frame_stride;
#pragma unroll
                for (int m = 0; m < kPreRows; ++m) {
                    const int rr = wave + NW * m;
                    if (rr >= R) break;
                    const int gy = by0 + rr;
                    const float* src = f + static_cast<long long>(gy - row0) * pitch + bx0;
                    if (gy < row0 || gy >= row1) continue;
                    const int gx = bx0 + lane;
                    if (gx >= 0 && gx < w) __builtin_amdgcn_global_load_lds(src + lane, dst + rr * kPitch, 4, 0, 0);
                    if (lane < sx && gx + kTileX >= 0 && gx + kTileX < w)
                        __builtin_amdgcn_global_load_lds(src + kTileX + lane, dst + rr * kPitch + kTileX, 4, 0, 0);
                }
            } else {
                // uniform row pointers with a per-lane offset (the scalar-base load form);
                // for the box's columns past 64, lanes beyond the box width re-read its
                // last column into LDS columns that no motion reads, so every lane stays
                // inside the box and no exec mask is needed (a box exactly 64 wide re-reads
                // column 63 there: the second load needs no branch)
                // buffer form of the LDS load: descriptor at the box's first pixel, the
                // lane's byte offset in a VGPR (fixed for the frame), the row's byte
                // offset as a scalar, so a row costs no vector instruction
                const int b1 = 4 * lane;
                // second chunk: the instruction offset (4 kTileX bytes) applies to both the
                // frame and the LDS address, so both loads of a row share one M0 value
                const int b2 = 4 * min(lane, sx - 1);
                const float* boxp = ft + (bx[1] * pitch + bx[0]);
                const __amdgpu_buffer_rsrc_t rs =
                    __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(boxp), 0, 0x7fffffff, kBufferWord3);
                if constexpr (RING) {
                    // a rolled loop with running offsets: the edge kernel has no scalar
                    // registers to spare for the unrolled loop's per-row offsets
                    int row_b = wave * 4 * pitch;
                    float* d = dst + wave * kPitch;
#pragma unroll 1
                    for (int rr = wave; rr < R; rr += NW) {
                        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, d, 4, b1, row_b, 0, 0);
                        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, d, 4, b2, row_b, 4 * kTileX, 0);
                        row_b += NW * 4 * pitch;
                        d += NW * kPitch;
                    }
                } else {
                    // a lattice box has at most TY + kLatSyMax rows
                    constexpr int kStageRows = LAT ? (TY + kLatSyMax + NW - 1) / NW : kPreRows;
#pragma unroll
                    for (int m = 0; m < kStageRows; ++m) {
                        const int rr = wave + NW * m;
                        if (rr >= R) break;
                        const int row_b = rr * 4 * pitch;
                        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, dst + rr * kPitch, 4, b1, row_b, 0, 0);
                        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, dst + rr * kPitch, 4, b2, row_b, 4 * kTileX, 0);
                    }
                }
            }
        };
        // lw: a lattice frame's words (column offsets, row classes; unused otherwise)
        auto consume = [&](const float* L, int i, const int (&lw)[2]) {
            if constexpr (LAT) {
                // column a (dx = X[a]) of rows 0 .. 3 + spread, read once; motion (a, b) adds
                // rows c_b .. c_b + 3 of it
                const int wx = lw[0], wy = lw[1];
                constexpr int kRows = 2 * ((kLatSyMax + 1) / 2) + 4;  // row pairs
                // rows past 3 are read only when the frame's spread needs them; the
                // unread ones are never added (the plan keeps c_b <= 2 e), so they stay
                // unset instead of costing a move each
                float col[kLatSide][kRows];
                const float* q[kLatSide];
#pragma unroll
                for (int c = 0; c < kLatSide; ++c) {
                    q[c] = L + lane_base + ((wx >> (8 * c)) & 0xff);
#pragma unroll
                    for (int r = 0; r < 4; ++r) col[c][r] = q[c][r * kPitch];
                }
                const int extra = (wy >> 16) & 3;  // row pairs past the first four rows
                if (extra >= 1) {
#pragma unroll
                    for (int c = 0; c < kLatSide; ++c) {
                        col[c][4] = q[c][4 * kPitch];
                        col[c][5] = q[c][5 * kPitch];
                    }
                }
                if (extra >= 2) {
#pragma unroll
                    for (int c = 0; c < kLatSide; ++c) {
                        col[c][6] = q[c][6 * kPitch];
                        col[c][7] = q[c][7 * kPitch];
                    }
                }
                lat_add_all(acc, col, wy & 0xf, (wy >> 4) & 0xf, (wy >> 8) & 0xf, (wy >> 12) & 0xf);
            } else {
                cint* lo = c_off + (static_cast<long long>(blk) * n + i) * V;
#pragma unroll
                for (int k = 0; k < V; ++k) {
                    const float* p = L + lane_base + lo[k];
#pragma unroll
                    for (int r = 0; r < RPW; ++r) acc[k][r] += p[r * kPitch];
                }
            }
        };

        if constexpr (ST == kStoreF32) {
            // The scalar loads of a frame's box and lattice words are issued one frame
            // ahead (while the frame before is staged and consumed), so no frame waits
            // on its own descriptor; the frame pointer at the tile origin advances by the
            // frame stride.
            auto load_box = [&](int i, int (&d)[4]) {
                cint* bx = c_box + 4 * (static_cast<long long>(blk) * n + i);
                d[0] = bx[0];
                d[1] = bx[1];
                d[2] = bx[2];
                d[3] = bx[3];
            };
            auto load_lat = [&](int i, int (&d)[2]) {
                if constexpr (LAT) {
                    cint* lt = c_lat + 2 * (static_cast<long long>(blk) * n + i);
                    d[0] = lt[0];
                    d[1] = lt[1];
                }
            };
            const float* ft = frames + tile_org;
            int bn[4], ln[2] = {0, 0}, lc[2];
            load_box(0, bn);
            stage(0, bn, ft, lds0);
            if (n > 1) load_box(1, bn);
            load_lat(0, ln);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            for (int i = 0; i < n; i += 2) {
                lc[0] = ln[0];
                lc[1] = ln[1];
                if (i + 1 < n) {
                    ft += frame_stride;
                    stage(i + 1, bn, ft, lds1);
                    if (i + 2 < n) load_box(i + 2, bn);
                    load_lat(i + 1, ln);
                }
                consume(lds0, i, lc);
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
                if (i + 1 >= n) break;
                lc[0] = ln[0];
                lc[1] = ln[1];
                if (i + 2 < n) {
                    ft += frame_stride;
                    stage(i + 2, bn, ft, lds0);
                    if (i + 3 < n) load_box(i + 3, bn);
                    load_lat(i + 2, ln);
                }
                consume(lds1, i + 1, lc);
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
            }
        } else {
            // 16-bit frames, kept 16-bit in LDS as row pairs: LDS word (r, c) holds the
            // values of box rows r (low half) and r + 1 (high half) at box column c. Each
            // buffer starts one word into the LDS area, so a pair fetched from one column
            // left of the box (odd start) writes its first word to the word before the
            // buffer (the area's first word, or the previous buffer's last word, which no
            // motion reads). A lane's four rows of a motion are then two words, read
            // with one ds_read2st64, half the LDS traffic of fp32 rows; each value is
            // widened and multiplied by the frame's power-of-two scale inside a fused
            // multiply-add whose product is exact (q times a power of two), so every sum
            // is the fp32 sum of the decoded frames (q * s), added in frame order.
            //
            // Staging: wave w fetches box rows w PRE .. w PRE + PRE (clamped to the box)
            // as dwords of two pixels from the even column at or left of the box start,
            // while the current frame is consumed, then packs the pairs of rows into the
            // word rows w PRE .. w PRE + PRE - 1 of the other buffer (word rows past the
            // box are written but never read; the buffer has at least 4 PRE rows).
            uint32_t* const lw0 = reinterpret_cast<uint32_t*>(lds0) + 1;
            uint32_t* const lw1 = reinterpret_cast<uint32_t*>(lds1) + 1;
            uint32_t pre[PRE + 1];
            int pre_col = 0;
            float sc_next = 1.0f;
            auto fetch = [&](int i) {
                cint* bx = c_box + 4 * (static_cast<long long>(blk) * n + i);
                const int sx = bx[2], R = TY + bx[3];
                const int bxs = x0 + bx[0];
                const int off = bxs & 1;  // 1: the pair starts one column left of the box
                const int lmax = (kTileX - 1 + sx + off) >> 1;
                const int l = min(lane, lmax);  // lanes past the box repeat its last pair
                const elem_t* rowp =
                    frames + static_cast<long long>(i) * frame_stride + static_cast<long long>(y0 + bx[1] - row0) * pitch + (bxs - off);
                const __amdgpu_buffer_rsrc_t rs =
                    __builtin_amdgcn_make_buffer_rsrc(const_cast<elem_t*>(rowp), 0, 0x7fffffff, kBufferWord3);
                pre_col = 2 * l - off;  // word column of the pair's first pixel (-1 for the guard word)
                sc_next = reinterpret_cast<const __attribute__((address_space(4))) float*>(
                    (const __attribute__((address_space(4))) char*)(scale))[i];
#pragma unroll
                for (int m = 0; m <= PRE; ++m) {
                    const int rr = min(wave * PRE + m, R - 1);
                    pre[m] = __builtin_amdgcn_raw_buffer_load_b32(rs, 4 * l, rr * 2 * pitch, 0);
                }
            };
            auto commit = [&](uint32_t* dst) {
                uint32_t* d = dst + wave * PRE * kPitch + pre_col;
#pragma unroll
                for (int m = 0; m < PRE; ++m) {
                    // v_perm_b32 picks the bytes: low halves of rows m, m + 1, then high halves
                    d[m * kPitch] = __builtin_amdgcn_perm(pre[m + 1], pre[m], 0x05040100u);
                    d[m * kPitch + 1] = __builtin_amdgcn_perm(pre[m + 1], pre[m], 0x07060302u);
                }
            };
            // per-lane read bases of the two buffers, formed once (the empty asm keeps the
            // compiler from splitting the guard-word offset back out of them, which would
            // cost one more add per motion)
            int lb = lane_base + 1;
            asm volatile("" : "+v"(lb));
            const uint32_t* const rb0 = reinterpret_cast<const uint32_t*>(lds0) + lb;
            const uint32_t* const rb1 = reinterpret_cast<const uint32_t*>(lds1) + lb;
            auto consume16 = [&](const uint32_t* L, int i, float sc) {
                cint* lo = c_off + (static_cast<long long>(blk) * n + i) * V;
                // the words of motions k + 2, k + 3 are read while motions k, k + 1 are
                // added, so a wavefront keeps four LDS reads in flight
                uint32_t cur[2][2], nxt[2][2];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    const uint32_t* p = L + lo[t];
                    cur[t][0] = p[0];
                    cur[t][1] = p[2 * kPitch];
                }
#pragma unroll
                for (int k = 0; k < V; k += 2) {
                    if (k + 2 < V) {
#pragma unroll
                        for (int t = 0; t < 2; ++t) {
                            const uint32_t* p = L + lo[k + 2 + t];
                            nxt[t][0] = p[0];
                            nxt[t][1] = p[2 * kPitch];
                        }
                    }
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        acc[k + t][0] = fma_q16<ST>(cur[t][0], false, sc, acc[k + t][0]);
                        acc[k + t][1] = fma_q16<ST>(cur[t][0], true, sc, acc[k + t][1]);
                        acc[k + t][2] = fma_q16<ST>(cur[t][1], false, sc, acc[k + t][2]);
                        acc[k + t][3] = fma_q16<ST>(cur[t][1], true, sc, acc[k + t][3]);
                    }
                    if (k + 2 < V) {
#pragma unroll
                        for (int t = 0; t < 2; ++t) {
                            cur[t][0] = nxt[t][0];
                            cur[t][1] = nxt[t][1];
                        }
                    }
                }
            };
            fetch(0);
            commit(lw0);
            float sc = sc_next;
            __syncthreads();
            // (straight-line code: past the last frame the last one is staged again into
            // the buffer that is not read any more)
            for (int i = 0; i < n; i += 2) {
                fetch(min(i + 1, n - 1));
                consume16(rb0, i, sc);
                commit(lw1);
                sc = sc_next;
                __syncthreads();
                if (i + 1 >= n) break;
                fetch(min(i + 2, n - 1));
                consume16(rb1, i + 1, sc);
                commit(lw0);
                sc = sc_next;
                __syncthreads();
            }
        }

        // S/N of the block's motions (ascending index) and the running maximum
        cint* mo = c_motion + V * blk;
        const __attribute__((address_space(4))) float* const c_inv =
            (const __attribute__((address_space(4))) float*)(inv_sqrt);
        // coverage from the full-frame bounds. cov_x / cov_y hold, per block and motion,
        // the number of frames whose shift is <= d for every d of the block's range, so a
        // pixel's count of frames landing inside [0, w) is F(w-1-x) - F(-1-x). A block that
        // stays inside in y needs only the x table (one count per lane, the same for the
        // lane's rows); one that stays inside in x only the y table (one count per row,
        // the same for the wave: scalar loads). A block leaving both ranges uses both
        // tables when each motion's shifts form a chain or anti-chain (straight tracks
        // do), else walks the frames.
        int cover = 0;  // 0: c = n; 1: x table; 2: y table; 3: both tables or frame walk
        if (RING && !interior) {
            cint* ub = c_ubox + 4 * blk;
            const bool x_full = x0 + ub[0] >= 0 && x0 + kTileX - 1 + ub[2] < w;
            const bool y_full = y0 + ub[1] >= 0 && y0 + TY - 1 + ub[3] < h;
            cover = (x_full && y_full) ? 0 : (cov_ok && y_full) ? 1 : (cov_ok && x_full) ? 2 : 3;
        }
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const int j = mo[k];
            if (j < 0) continue;
            int cnt[RPW];
            float ix = inv_n;  // the factor of every row when the count is the same for them
#pragma unroll
            for (int r = 0; r < RPW; ++r) cnt[r] = n;
            if (RING && cover != 0) {
                cint* ub = c_ubox + 4 * blk;
                if (cover == 1) {
                    const int L = ub[2] - ub[0] + 1;
                    const int* F = cov_x + cov_off[2 * blk] + static_cast<long long>(k) * L;
                    auto cdf = [&](int q) { const int d = q - ub[0]; return d < 0 ? 0 : (d >= L ? n : F[d]); };
                    const int cx = cdf(w - 1 - x) - cdf(-1 - x);
                    ix = inv_sqrt[cx];
#pragma unroll
                    for (int r = 0; r < RPW; ++r) cnt[r] = cx;
                } else if (cover == 2) {
                    const int L = ub[3] - ub[1] + 1;
                    cint* G = c_covy + c_covoff[2 * blk + 1] + static_cast<long long>(k) * L;
                    auto cdf = [&](int q) { const int d = q - ub[1]; return d < 0 ? 0 : (d >= L ? n : G[d]); };
#pragma unroll
                    for (int r = 0; r < RPW; ++r) cnt[r] = cdf(h - 1 - (yw + r)) - cdf(-1 - (yw + r));
                } else if (cov_ok && c_chain[V * blk + k] != 0) {
                    // both ranges cut: for a chain (anti-chain) of (dx, dy) pairs the frames
                    // in the x range and those in the y range are two intervals of one
                    // order (by dx, then dy ascending, or descending), so the count is
                    // their overlap, from the two tables
                    const int Lx = ub[2] - ub[0] + 1, Ly = ub[3] - ub[1] + 1;
                    const int* F = cov_x + cov_off[2 * blk] + static_cast<long long>(k) * Lx;
                    cint* G = c_covy + c_covoff[2 * blk + 1] + static_cast<long long>(k) * Ly;
                    auto fx = [&](int q) { const int d = q - ub[0]; return d < 0 ? 0 : (d >= Lx ? n : F[d]); };
                    auto gy = [&](int q) { const int d = q - ub[1]; return d < 0 ? 0 : (d >= Ly ? n : G[d]); };
                    const int x_lo = fx(-1 - x), x_hi = fx(w - 1 - x);  // positions [x_lo, x_hi)
                    const bool co = c_chain[V * blk + k] == 1;
#pragma unroll
                    for (int r = 0; r < RPW; ++r) {
                        const int below = gy(-1 - (yw + r)), upto = gy(h - 1 - (yw + r));
                        // y range positions: [below, upto) ascending, [n - upto, n - below) descending
                        const int y_lo = co ? below : n - upto, y_hi = co ? upto : n - below;
                        cnt[r] = max(0, min(x_hi, y_hi) - max(x_lo, y_lo));
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < RPW; ++r) cnt[r] = 0;
                    for (int i = 0; i < n; ++i) {
                        cint* sh = c_shift + (static_cast<long long>(blk) * n + i) * (2 * V);
                        const int sdx = sh[k], sdy = sh[V + k];
                        const int xin = static_cast<unsigned>(x + sdx) < static_cast<unsigned>(w) ? 1 : 0;
#pragma unroll
                        for (int r = 0; r < RPW; ++r)
                            cnt[r] += static_cast<unsigned>(yw + r + sdy) < static_cast<unsigned>(h) ? xin : 0;
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                if (cnt[r] < min_count) continue;
                // x table or c = n: one factor; y table: the count is the wave's (scalar
                // load); frame walk: a per-lane count
                const float f = (!RING || cover <= 1) ? ix : cover == 2 ? c_inv[cnt[r]] : inv_sqrt[cnt[r]];
                const float snr = acc[k][r] * f;
                if (snr > best[r] || (snr == best[r] && j < arg[r])) {
                    best[r] = snr;
                    arg[r] = j;
                }
            }
        }
    }

#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        const int y = yw + r;
        if (arg[r] < 0 || x >= w || y >= y_end) continue;
        const unsigned long long key =
            (static_cast<unsigned long long>(d_ordered(best[r])) << 32) | static_cast<uint32_t>(~static_cast<uint32_t>(arg[r]));
        atomicMax(&keys[static_cast<long long>(y - y_begin) * w + x], key);
    }
}

// Workgroups go to the 8 XCDs round-robin (workgroup b to XCD b mod 8). The k-th
// workgroup an XCD receives works on tile 8 (k / nseg) + XCD, segment k mod nseg: the
// segments of a tile run together on one XCD (sharing its L2 copy of the tile's frame
// boxes) and the eight XCDs work on neighbouring tiles (sharing frame rows in the
// MALL). The grid is padded to whole groups of 8 tiles.
template <int NW, bool RING, bool PADDED, int ST, int PRE, bool LAT>
__device__ __forceinline__ void stack_grid_item(const StackArgs& a, unsigned b) {
    const unsigned xcd = b & 7u, kq = b >> 3;
    const int tile = static_cast<int>(8u * (kq / static_cast<unsigned>(a.nseg)) + xcd);
    const int seg = a.seg0 + static_cast<int>(kq % static_cast<unsigned>(a.nseg));
    if (tile >= a.ntiles) return;
    stack_body<NW, RING, PADDED, ST, PRE, LAT>(a, tile, seg);
}

template <int NW, bool RING, bool PADDED, int ST, int PRE, bool LAT = false>
__global__ __launch_bounds__(64 * NW) void stack_kernel(StackArgs a) {
    stack_grid_item<NW, RING, PADDED, ST, PRE, LAT>(a, blockIdx.x);
}

// Several passes in one launch (the real pass and null runs of a batched search, or
// regions of a frame): workgroup b works on pass b / per_pass. The descriptor address
// is uniform and the kernel never writes it, so the compiler reads it with scalar loads.
template <int NW, bool RING, bool PADDED, int ST, int PRE, bool LAT = false>
__global__ __launch_bounds__(64 * NW) void stack_multi_kernel(const StackArgs* __restrict__ passes, unsigned per_pass) {
    const unsigned p = blockIdx.x / per_pass;
    const StackArgs a = passes[p];
    stack_grid_item<NW, RING, PADDED, ST, PRE, LAT>(a, blockIdx.x - p * per_pass);
}

// Persistent form: every kind of work of d.count passes in one launch of about as many
// workgroups as the GPU holds at once. Work items sit in 8 queues, one per XCD: queue x
// holds, pass after pass, the ring items, then the lattice items, then the general
// items of the tiles t = x (mod 8), each tile's segments in a row, so the workgroups of
// XCD x (workgroup b runs on XCD b mod 8) take the items whose frame boxes its own L2
// holds, as the one-workgroup-per-item grid does. A workgroup takes the next item of
// its queue with one atomic add on d.counters[x]; when its queue is exhausted it takes
// items from the other queues in turn, and it exits once all 8 are exhausted: every
// workgroup ends after at most (its items + 8) counter reads, whatever the others do.
// The counters are zero before the launch.
template <int NW, bool PADDED, int ST, int PRE>
__global__ __launch_bounds__(64 * NW, ST == kStoreF32 ? 1 : 5) void stack_queue_kernel(QueueDesc d) {
    __shared__ int s_item;
    const int home = static_cast<int>(blockIdx.x & 7u);
    for (int q = 0; q < 8; ++q) {
        const int x = (home + q) & 7;
        int per[kArgsKinds];
        int items = 0;  // items of one pass in queue x
#pragma unroll
        for (int kind = 0; kind < kArgsKinds; ++kind) {
            const int t = d.ntiles[kind] > x ? (d.ntiles[kind] - 1 - x) / 8 + 1 : 0;
            per[kind] = t * d.nseg[kind];
            items += per[kind];
        }
        const int total = items * d.count;
        for (;;) {
            if (threadIdx.x == 0) s_item = atomicAdd(&d.counters[x], 1);
            __syncthreads();
            const int k = __builtin_amdgcn_readfirstlane(s_item);
            __syncthreads();  // every wavefront has read s_item before the next add writes it
            if (k >= total) break;
            const int p = k / items;
            int r = k - p * items;
            int kind = kArgsRing;
            if (r >= per[kArgsRing]) {
                r -= per[kArgsRing];
                kind = kArgsLat;
                if (r >= per[kArgsLat]) {
                    r -= per[kArgsLat];
                    kind = kArgsGen;
                }
            }
            const StackArgs a = d.args[kind * d.stride + p];
            const int ns = d.nseg[kind];
            const int tile = 8 * (r / ns) + x;
            const int seg = a.seg0 + r % ns;
            if (kind == kArgsRing) {
                stack_body<NW, true, PADDED, ST, PRE>(a, tile, seg);
            } else if (kind == kArgsLat) {
                if constexpr (ST == kStoreF32) stack_body<NW, false, false, ST, 0, true>(a, tile, seg);
            } else {
                stack_body<NW, false, ST != kStoreF32, ST, PRE>(a, tile, seg);
            }
        }
    }
}

}  // namespace ftk

// Launchers of the gfx950 stacking kernels (stack_body.hpp; instantiated per frame
// storage in stack_inst_f32.hip, stack_inst_f16.hip and stack_inst_bf16.hip), their
// descriptors and the small kernels around them (key clearing, finalisation, row
// copies).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <utility>
#include <vector>

#include "stack_inst.hpp"
#include "stack_launch.hpp"
#include "stack_plan.hpp"

namespace ftk {

namespace {

__global__ void zero_counters_kernel(int* __restrict__ c) {
    if (threadIdx.x < 8) c[threadIdx.x] = 0;
}

// The three descriptors of a one-pass persistent launch, from the kernel arguments
// into device memory (ordered on the launch's stream, no host synchronisation).
struct ArgsTriple {
    StackArgs a[kArgsKinds];
};
constexpr int kTripleWords = static_cast<int>(sizeof(ArgsTriple) / 4);
static_assert(sizeof(ArgsTriple) % 4 == 0, "descriptor words");
__global__ void put_args_kernel(ArgsTriple t, uint32_t* __restrict__ dst) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&t);
    for (int i = threadIdx.x; i < kTripleWords; i += blockDim.x) dst[i] = src[i];
}

__device__ inline float d_unordered(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

}  // namespace

__global__ void fill_keys_kernel(unsigned long long* __restrict__ keys, long long count) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) keys[i] = 0ull;
}

// Keys -> (max_snr, best_v): 0 / -1 where no motion is defined; a non-finite maximum
// becomes 0 with its motion kept (src/backend_cpu.cpp:191-192).
__global__ void finalize_kernel(const unsigned long long* __restrict__ keys, long long count,
                                float* __restrict__ max_snr, int* __restrict__ best_v) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const unsigned long long k = keys[i];
    if (k == 0ull) {
        max_snr[i] = 0.0f;
        best_v[i] = -1;
        return;
    }
    const float v = d_unordered(static_cast<uint32_t>(k >> 32));
    max_snr[i] = __builtin_isfinite(v) ? v : 0.0f;
    best_v[i] = static_cast<int>(~static_cast<uint32_t>(k & 0xffffffffull));
}

__global__ void copy_f32_kernel(const float* __restrict__ src, float* __restrict__ dst, long long count) {
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
        dst[i] = src[i];
}

// ------------------------------------------------------------------ launchers

hipError_t launch_fill_keys(unsigned long long* keys, long long count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    const unsigned blocks = static_cast<unsigned>((count + 255) / 256);
    hipLaunchKernelGGL(fill_keys_kernel, dim3(blocks), dim3(256), 0, s, keys, count);
    return hipGetLastError();
}

hipError_t launch_finalize(const unsigned long long* keys, long long count, float* max_snr, int* best_v,
                           hipStream_t s) {
    if (count <= 0) return hipSuccess;
    const unsigned blocks = static_cast<unsigned>((count + 255) / 256);
    hipLaunchKernelGGL(finalize_kernel, dim3(blocks), dim3(256), 0, s, keys, count, max_snr, best_v);
    return hipGetLastError();
}

hipError_t launch_copy(const float* src, float* dst, long long count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(copy_f32_kernel, dim3(2048), dim3(256), 0, s, src, dst, count);
    return hipGetLastError();
}

namespace {
int ceil_div_pos(int a, int b) { return a <= 0 ? 0 : (a + b - 1) / b; }
int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }
}  // namespace

// Interior tiles: x0 + min dx >= 0, x0 + 63 + max dx < w, y0 + min dy >= row0 and
// y0 + tile_rows - 1 + max dy < row1 for the union box of every block.
void interior_rect(const StackPlan& p, const StackLaunch& L, int& tx_lo, int& tx_hi, int& ty_lo, int& ty_hi) {
    const int* u = p.all_box;
    const int ty = p.tile_rows;
    tx_lo = std::max(0, ceil_div_pos(-u[0], kTileX));
    tx_hi = std::min(L.tiles_x, floor_div(p.w - kTileX - u[2], kTileX) + 1);
    ty_lo = std::max(0, ceil_div_pos(L.row0 - u[1] - L.y_begin, ty));
    ty_hi = std::min(L.tiles_y, floor_div(L.row1 - ty - u[3] - L.y_begin, ty) + 1);
    if (tx_hi <= tx_lo || ty_hi <= ty_lo) tx_lo = tx_hi = ty_lo = ty_hi = 0;
}

int rows16_per_wave(const StackPlan& p, int tile_rows) {
    const int nw = tile_rows / kRowsPerWave;
    // word rows: tile_rows - 1 + the largest vertical spread, pre per wavefront
    const int need = (tile_rows - 1 + p.max_sy + nw - 1) / nw;
    return need <= 5 ? 5 : need <= 6 ? 6 : need <= 8 ? 8 : need <= 12 ? 12 : 0;
}

bool tile_rows_fit(const StackPlan& p, int storage, int tile_rows) {
    if (!tile_rows_ok(tile_rows)) return false;
    return storage == kStoreF32 || rows16_per_wave(p, tile_rows) > 0;
}

StackShape resolve_shape(const StackPlan& p, int storage, int tile_rows, int queue, int out_rows) {
    if (tile_rows != 0 && !tile_rows_ok(tile_rows)) throw std::runtime_error("tile height must be 8, 16 or 32 rows");
    if (queue < -1 || queue > 1) throw std::runtime_error("launch form (queue) must be -1, 0 or 1");
    if (queue == 0 && tile_rows != 0 && tile_rows != kTileY)
        throw std::runtime_error("tile heights other than 16 rows need the persistent launch (queue 1)");
    StackShape sh;
    if (queue < 0) {
        const long long tiles = static_cast<long long>((p.w + kTileX - 1) / kTileX) * ((out_rows + kTileY - 1) / kTileY);
        queue = storage != kStoreF32 || tiles >= 1024 ? 1 : 0;
    }
    sh.queue = queue;
    sh.tile_rows = tile_rows > 0 ? tile_rows : kTileY;
    if (!tile_rows_fit(p, storage, sh.tile_rows)) sh.tile_rows = kTileY;
    if (!sh.queue) sh.tile_rows = kTileY;
    return sh;
}

void make_stack_args(const StackPlan& p, const StackLaunch& L, StackArgs out[kArgsKinds]) {
    TileRect R{L.tiles_x, L.tiles_y, 0, 0, 0, 0};
    interior_rect(p, L, R.tx_lo, R.tx_hi, R.ty_lo, R.ty_hi);
    const long long all = static_cast<long long>(L.tiles_x) * L.tiles_y;
    const long long in = static_cast<long long>(R.tx_hi - R.tx_lo) * (R.ty_hi - R.ty_lo);
    const int ty = p.tile_rows, nw = ty / kRowsPerWave;
    StackArgs a{};
    a.frames = L.frames;
    a.scale = L.scale;
    a.storage = L.storage;
    a.frame_stride = L.frame_stride;
    a.shift = L.shift;
    a.box = L.box;
    a.ubox = L.ubox;
    a.lds_off = L.lds_off;
    a.motion = L.motion;
    a.inv_sqrt = L.inv_sqrt;
    a.cov_off = L.cov_off;
    a.cov_x = L.cov_x;
    a.cov_y = L.cov_y;
    a.chain = L.chain;
    a.lat = L.lat;
    a.keys = L.keys;
    a.rect = R;
    a.row0 = L.row0;
    a.row1 = L.row1;
    a.n = p.n;
    a.w = p.w;
    a.h = p.h;
    a.pitch = L.pitch > 0 ? L.pitch : p.w;
    a.padded = L.padded ? 1 : 0;
    a.min_count = p.min_count;
    a.nb = p.nb;
    a.bps = L.bps;
    a.y_begin = L.y_begin;
    a.y_end = L.y_end;
    a.seg0 = L.seg0;
    a.nseg = L.nseg;
    a.blk_lo = 0;
    a.blk_hi = p.nb;
    a.tile_rows = ty;
    a.lds_rows = ty + p.max_sy;
    if (L.storage != kStoreF32) {
        // 0 when this tile height cannot hold the spread (the launchers refuse it)
        a.pre = rows16_per_wave(p, ty);
        a.lds_rows = std::max(a.lds_rows, nw * a.pre);
    }
    a.cov_ok = p.cov_ok ? 1 : 0;
    StackArgs& lat = out[kArgsLat];
    StackArgs& gen = out[kArgsGen];
    StackArgs& ring = out[kArgsRing];
    lat = gen = ring = a;
    ring.ntiles = static_cast<int>(all - in);
    lat.ntiles = gen.ntiles = static_cast<int>(in);
    const bool split = L.storage == kStoreF32 && L.lat != nullptr && p.nb_lat > 0;
    const int bps = std::max(1, L.bps), seg_end = L.seg0 + L.nseg;
    if (split) {
        // lattice blocks [0, nb_lat) and the others on interior tiles, each launch over
        // the segments that hold its blocks. A lane of the lattice kernel reads rows
        // 4 w .. 4 w + 3 + 2 e of a buffer (wavefront w, e = (spread + 1) / 2 row pairs), up
        // to row tile_rows - 1 + 2 e: the buffers hold tile_rows + the largest spread
        // rounded up to even.
        lat.lds_rows = ty + 2 * ((p.max_sy_lat + 1) / 2);
        lat.blk_hi = p.nb_lat;
        lat.nseg = std::max(0, std::min(seg_end, (p.nb_lat + bps - 1) / bps) - L.seg0);
        gen.blk_lo = p.nb_lat;
        gen.seg0 = std::max(L.seg0, p.nb_lat / bps);
        gen.nseg = std::max(0, seg_end - gen.seg0);
        if (p.nb_lat >= p.nb) gen.nseg = 0;
    } else {
        lat.nseg = 0;
    }
}

namespace {
// workgroups of one pass: tiles padded to whole groups of 8 (one per XCD), times the
// segments
unsigned pass_groups(const StackArgs& a) {
    if (a.ntiles <= 0 || a.nseg <= 0) return 0;
    return static_cast<unsigned>((static_cast<long long>(a.ntiles) + 7) / 8 * 8 * a.nseg);
}
// two buffers and the guard word of the 16-bit layout
size_t lds_bytes_of(int lds_rows) { return sizeof(float) * (2 * lds_rows * kPitch + 1); }

bool storage_ok(const StackArgs& a) {
    if (a.storage == kStoreF32) return true;
    return (a.storage == kStoreF16 || a.storage == kStoreBF16) && a.padded && a.scale != nullptr && a.pre > 0;
}
SingleFn single_kernel(int kind, const StackArgs& a) {
    return a.storage == kStoreF32 ? single_kernel_f32(kind, a.padded != 0) : single_kernel_16(a.storage, kind == kArgsRing, a.pre);
}
MultiFn multi_kernel(int kind, const StackArgs& a) {
    return a.storage == kStoreF32 ? multi_kernel_f32(kind, a.padded != 0) : multi_kernel_16(a.storage, kind == kArgsRing, a.pre);
}
QueueFn queue_kernel(const StackArgs& a) {
    const int nw = a.tile_rows / kRowsPerWave;
    return a.storage == kStoreF32 ? queue_kernel_f32(nw, a.padded != 0) : queue_kernel_16(a.storage, nw, a.pre);
}

// Workgroups of one persistent launch: as many as the GPU holds at once (occupancy per
// CU for the kernel and its LDS, times the CUs), a multiple of 8, at most what the
// items need. Occupancy is asked once per kernel and LDS size.
int resident_groups(QueueFn fn, int threads, size_t lds) {
    static std::mutex mu;
    static std::vector<std::pair<std::pair<const void*, size_t>, int>> cache;
    const void* key = reinterpret_cast<const void*>(fn);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    {
        std::lock_guard<std::mutex> g(mu);
        for (const auto& c : cache)
            if (c.first.first == key && c.first.second == (lds << 8 | static_cast<size_t>(dev))) return c.second;
    }
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, key, threads, lds) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    const int groups = std::max(1, per_cu) * std::max(1, cus);
    std::lock_guard<std::mutex> g(mu);
    cache.push_back({{key, lds << 8 | static_cast<size_t>(dev)}, groups});
    return groups;
}
}  // namespace

hipError_t launch_stack(const StackPlan& p, const StackLaunch& L, hipStream_t s, hipStream_t ring_s) {
    StackArgs A[kArgsKinds];
    make_stack_args(p, L, A);
    const StackArgs& ring = A[kArgsRing];
    if (!storage_ok(ring) || !tile_rows_ok(ring.tile_rows)) return hipErrorInvalidValue;
    if (L.counters) {
        // descriptors after the counters (256 bytes in) in the launch's workspace
        static_assert(256 + sizeof(ArgsTriple) <= kQueueWorkspaceBytes, "workspace size");
        ArgsTriple t;
        for (int k = 0; k < kArgsKinds; ++k) t.a[k] = A[k];
        uint32_t* dst = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(L.counters) + 256);
        hipLaunchKernelGGL(put_args_kernel, dim3(1), dim3(256), 0, s, t, dst);
        return launch_stack_queue(A, reinterpret_cast<const StackArgs*>(dst), 1, 1, L.counters, s);
    }
    if (ring.tile_rows != kTileY) return hipErrorInvalidValue;  // one-workgroup-per-item kernels: 16-row tiles
    auto go = [&](int kind, hipStream_t st) {
        const StackArgs& a = A[kind];
        const unsigned g = pass_groups(a);
        if (g > 0)
            hipLaunchKernelGGL(single_kernel(kind, a), dim3(g), dim3(kThreads), lds_bytes_of(a.lds_rows), st, a);
    };
    // the edge tiles first (on their own stream when given), the interior ones fill in;
    // the general blocks' interior launch (LDS-bound) goes beside the lattice launch
    // (VALU-bound) on the edge stream
    go(kArgsRing, ring_s ? ring_s : s);
    go(kArgsLat, s);
    go(kArgsGen, ring_s ? ring_s : s);
    return hipGetLastError();
}

hipError_t launch_stack_multi(const StackArgs* const host[kArgsKinds], const StackArgs* const dev[kArgsKinds], int count,
                              hipStream_t s, hipStream_t ring_s) {
    if (count <= 0) return hipSuccess;
    unsigned per[kArgsKinds] = {0, 0, 0};
    const StackArgs& r0 = host[kArgsRing][0];
    const int padded = r0.padded, storage = r0.storage, pre = r0.pre;
    if (r0.tile_rows != kTileY) return hipErrorInvalidValue;
    for (int kind = 0; kind < kArgsKinds; ++kind)
        for (int q = 0; q < count; ++q) {
            const StackArgs& a = host[kind][q];
            // one LDS size per launch
            if (a.lds_rows != host[kind][0].lds_rows || a.storage != storage || a.pre != pre || !storage_ok(a) ||
                a.tile_rows != kTileY || (kind == kArgsRing && a.padded != padded))
                return hipErrorInvalidValue;
            per[kind] = std::max(per[kind], pass_groups(a));
        }
    const dim3 block(kThreads);
    // the ring tiles first: their workgroups take longer, the interior ones fill in
    for (int kind : {static_cast<int>(kArgsRing), static_cast<int>(kArgsLat), static_cast<int>(kArgsGen)}) {
        if (per[kind] == 0) continue;
        // the lattice launch on s; the edge tiles and the general blocks' interior launch
        // beside it on ring_s (when given)
        hipLaunchKernelGGL(multi_kernel(kind, host[kind][0]), dim3(per[kind] * static_cast<unsigned>(count)), block,
                           lds_bytes_of(host[kind][0].lds_rows), kind != kArgsLat && ring_s ? ring_s : s, dev[kind],
                           per[kind]);
    }
    return hipGetLastError();
}

hipError_t launch_stack_queue(const StackArgs* host, const StackArgs* dev, int count, int stride, int* counters,
                              hipStream_t s) {
    if (count <= 0) return hipSuccess;
    if (!counters || !dev || stride < count) return hipErrorInvalidValue;
    QueueDesc d{};
    d.args = dev;
    d.counters = counters;
    d.count = count;
    d.stride = stride;
    const StackArgs& r0 = host[kArgsRing * stride];
    int lds_rows = 0;
    long long items = 0;
    for (int kind = 0; kind < kArgsKinds; ++kind) {
        const StackArgs& k0 = host[kind * stride];
        d.ntiles[kind] = k0.nseg > 0 ? k0.ntiles : 0;
        d.nseg[kind] = d.ntiles[kind] > 0 ? k0.nseg : 0;
        for (int q = 0; q < count; ++q) {
            const StackArgs& a = host[kind * stride + q];
            // one kernel and one work shape for every pass of the launch
            if (a.storage != r0.storage || a.pre != r0.pre || a.padded != r0.padded || a.tile_rows != r0.tile_rows ||
                !storage_ok(a) || (d.ntiles[kind] > 0 && (a.ntiles != k0.ntiles || a.nseg != k0.nseg)))
                return hipErrorInvalidValue;
            if (d.ntiles[kind] > 0) lds_rows = std::max(lds_rows, a.lds_rows);
        }
        items += static_cast<long long>(d.ntiles[kind]) * d.nseg[kind] * count;
    }
    if (items == 0) return hipSuccess;
    if (items > 0x7fffffffLL - 64 * 1024 * 1024) return hipErrorInvalidValue;  // int counters
    const QueueFn fn = queue_kernel(r0);
    if (!fn) return hipErrorInvalidValue;
    const int threads = 64 * (r0.tile_rows / kRowsPerWave);
    const size_t lds = lds_bytes_of(lds_rows);
    int groups = resident_groups(fn, threads, lds);
    if (groups <= 0) return hipErrorInvalidValue;
    groups = static_cast<int>(std::min<long long>(groups, (items + 7) / 8 * 8)) / 8 * 8;
    groups = std::max(groups, 8);
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, counters);
    hipLaunchKernelGGL(fn, dim3(static_cast<unsigned>(groups)), dim3(static_cast<unsigned>(threads)), lds, s, d);
    return hipGetLastError();
}

// Frame rows into a padded buffer: frame i's rows [0, rows) of w floats from
// src + i * src_stride (pitch w) to dst + i * dst_stride (pitch dst_pitch).
__global__ void copy_rows_kernel(const float* __restrict__ src, long long src_stride, float* __restrict__ dst,
                                 long long dst_stride, int dst_pitch, int w) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, i = blockIdx.z;
    if (x >= w) return;
    dst[i * dst_stride + static_cast<long long>(y) * dst_pitch + x] = src[i * src_stride + static_cast<long long>(y) * w + x];
}

hipError_t launch_copy_rows(const float* src, long long src_stride, float* dst, long long dst_stride, int dst_pitch,
                            int w, int rows, int n, hipStream_t s) {
    if (w <= 0 || rows <= 0 || n <= 0) return hipSuccess;
    for (int y0 = 0; y0 < rows; y0 += 32768) {
        const int ry = std::min(32768, rows - y0);
        hipLaunchKernelGGL(copy_rows_kernel, dim3((w + 255) / 256, ry, n), dim3(256), 0, s,
                           src + static_cast<long long>(y0) * w, src_stride, dst + static_cast<long long>(y0) * dst_pitch,
                           dst_stride, dst_pitch, w);
    }
    return hipGetLastError();
}

}  // namespace ftk

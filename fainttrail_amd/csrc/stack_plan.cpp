// Host planning for the gfx950 stacking kernel (see stack_plan.hpp).
#include "stack_plan.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>

namespace ftk {

namespace {

struct Shifts {
    const int32_t* dx;
    const int32_t* dy;
    int n, w, h;
    int x(int j, int i) const { return std::clamp(dx[static_cast<size_t>(j) * n + i], -w, w); }
    int y(int j, int i) const { return std::clamp(dy[static_cast<size_t>(j) * n + i], -h, h); }
};

// The frame whose shifts spread most over all motions (for a straight-line grid the
// frame furthest from the reference epoch): its shifts order the motions.
int widest_frame(const Shifts& s, int nv) {
    int best = 0;
    long best_spread = -1;
    for (int i = 0; i < s.n; ++i) {
        int x0 = std::numeric_limits<int>::max(), x1 = std::numeric_limits<int>::min();
        int y0 = x0, y1 = x1;
        for (int j = 0; j < nv; ++j) {
            x0 = std::min(x0, s.x(j, i));
            x1 = std::max(x1, s.x(j, i));
            y0 = std::min(y0, s.y(j, i));
            y1 = std::max(y1, s.y(j, i));
        }
        const long spread = static_cast<long>(x1 - x0) + (y1 - y0);
        if (spread > best_spread) {
            best_spread = spread;
            best = i;
        }
    }
    return best;
}

// Split the motions into groups of kVel that are compact in the shift space of one
// frame (a k-d split along the wider axis, cut at a multiple of kVel).
void kd_groups(std::vector<int>& idx, size_t lo, size_t hi, const Shifts& s, int fi,
               std::vector<std::vector<int>>& out) {
    const size_t cnt = hi - lo;
    if (cnt <= static_cast<size_t>(kVel)) {
        out.emplace_back(idx.begin() + static_cast<ptrdiff_t>(lo), idx.begin() + static_cast<ptrdiff_t>(hi));
        return;
    }
    int x0 = std::numeric_limits<int>::max(), x1 = std::numeric_limits<int>::min(), y0 = x0, y1 = x1;
    for (size_t q = lo; q < hi; ++q) {
        x0 = std::min(x0, s.x(idx[q], fi));
        x1 = std::max(x1, s.x(idx[q], fi));
        y0 = std::min(y0, s.y(idx[q], fi));
        y1 = std::max(y1, s.y(idx[q], fi));
    }
    const bool by_x = (x1 - x0) >= (y1 - y0);
    const size_t groups = (cnt + kVel - 1) / kVel;
    const size_t mid = lo + ((groups + 1) / 2) * kVel;
    auto key = [&](int j) { return by_x ? std::make_pair(s.x(j, fi), s.y(j, fi)) : std::make_pair(s.y(j, fi), s.x(j, fi)); };
    std::nth_element(idx.begin() + static_cast<ptrdiff_t>(lo), idx.begin() + static_cast<ptrdiff_t>(mid),
                     idx.begin() + static_cast<ptrdiff_t>(hi), [&](int a, int b) {
                         const auto ka = key(a), kb = key(b);
                         return ka != kb ? ka < kb : a < b;
                     });
    kd_groups(idx, lo, mid, s, fi, out);
    kd_groups(idx, mid, hi, s, fi, out);
}

// Emit a group as blocks whose per-frame spread fits the LDS box; groups that do not
// fit are halved along the worst frame and axis (a single motion always fits).
void emit_blocks(std::vector<int> g, const Shifts& s, StackPlan& p) {
    int worst_frame = -1, worst_axis = 0, worst = 0;
    for (int i = 0; i < s.n; ++i) {
        int x0 = std::numeric_limits<int>::max(), x1 = std::numeric_limits<int>::min(), y0 = x0, y1 = x1;
        for (int j : g) {
            x0 = std::min(x0, s.x(j, i));
            x1 = std::max(x1, s.x(j, i));
            y0 = std::min(y0, s.y(j, i));
            y1 = std::max(y1, s.y(j, i));
        }
        if (x1 - x0 > kSpreadMax && x1 - x0 - kSpreadMax > worst) {
            worst = x1 - x0 - kSpreadMax;
            worst_frame = i;
            worst_axis = 0;
        }
        if (y1 - y0 > kSpreadMax && y1 - y0 - kSpreadMax > worst) {
            worst = y1 - y0 - kSpreadMax;
            worst_frame = i;
            worst_axis = 1;
        }
    }
    if (worst_frame >= 0) {
        std::sort(g.begin(), g.end(), [&](int a, int b) {
            const int ka = worst_axis ? s.y(a, worst_frame) : s.x(a, worst_frame);
            const int kb = worst_axis ? s.y(b, worst_frame) : s.x(b, worst_frame);
            return ka != kb ? ka < kb : a < b;
        });
        const size_t half = g.size() / 2;
        emit_blocks(std::vector<int>(g.begin(), g.begin() + static_cast<ptrdiff_t>(half)), s, p);
        emit_blocks(std::vector<int>(g.begin() + static_cast<ptrdiff_t>(half), g.end()), s, p);
        return;
    }
    // ascending motion order inside a block: the kernel's strict '>' then keeps the
    // lowest index among equal S/N values of the block
    std::sort(g.begin(), g.end());
    const int b = p.nb++;
    const int n = s.n;
    p.motion.resize(static_cast<size_t>(p.nb) * kVel, -1);
    p.shift.resize(static_cast<size_t>(p.nb) * n * 2 * kVel, 0);
    p.box.resize(static_cast<size_t>(p.nb) * n * 4, 0);
    p.ubox.resize(static_cast<size_t>(p.nb) * 4, 0);
    for (int k = 0; k < kVel; ++k) p.motion[static_cast<size_t>(b) * kVel + k] = k < static_cast<int>(g.size()) ? g[static_cast<size_t>(k)] : -1;
    int ux0 = std::numeric_limits<int>::max(), uy0 = ux0, ux1 = std::numeric_limits<int>::min(), uy1 = ux1;
    for (int i = 0; i < n; ++i) {
        int32_t* sh = &p.shift[(static_cast<size_t>(b) * n + i) * 2 * kVel];
        int x0 = std::numeric_limits<int>::max(), x1 = std::numeric_limits<int>::min(), y0 = x0, y1 = x1;
        for (int k = 0; k < kVel; ++k) {
            // padding lanes repeat the block's first motion: always inside the box
            const int j = g[static_cast<size_t>(k < static_cast<int>(g.size()) ? k : 0)];
            sh[k] = s.x(j, i);
            sh[kVel + k] = s.y(j, i);
            x0 = std::min(x0, sh[k]);
            x1 = std::max(x1, sh[k]);
            y0 = std::min(y0, sh[kVel + k]);
            y1 = std::max(y1, sh[kVel + k]);
        }
        int32_t* bx = &p.box[(static_cast<size_t>(b) * n + i) * 4];
        bx[0] = x0;
        bx[1] = y0;
        bx[2] = x1 - x0;
        bx[3] = y1 - y0;
        p.max_sy = std::max(p.max_sy, y1 - y0);
        p.lds_off.resize(static_cast<size_t>(p.nb) * n * kVel, 0);
        int32_t* lo = &p.lds_off[(static_cast<size_t>(b) * n + i) * kVel];
        for (int k = 0; k < kVel; ++k) lo[k] = (sh[kVel + k] - y0) * kPitch + (sh[k] - x0);
        ux0 = std::min(ux0, x0);
        ux1 = std::max(ux1, x1);
        uy0 = std::min(uy0, y0);
        uy1 = std::max(uy1, y1);
    }
    int32_t* ub = &p.ubox[static_cast<size_t>(b) * 4];
    ub[0] = ux0;
    ub[1] = uy0;
    ub[2] = ux1;
    ub[3] = uy1;
    // cumulative shift counts per motion over the block's shift range (coverage of
    // edge tiles without walking the frames); skipped when the tables grow too large
    const long long Lx = ux1 - ux0 + 1, Ly = uy1 - uy0 + 1;
    const long long need = static_cast<long long>(p.cov_x.size() + p.cov_y.size()) + kVel * (Lx + Ly);
    if (!p.cov_ok || need > (1LL << 26)) {
        p.cov_ok = false;
        return;
    }
    const size_t ox = p.cov_x.size(), oy = p.cov_y.size();
    p.cov_off.push_back(static_cast<long long>(ox));
    p.cov_off.push_back(static_cast<long long>(oy));
    p.cov_x.resize(ox + static_cast<size_t>(kVel * Lx), 0);
    p.cov_y.resize(oy + static_cast<size_t>(kVel * Ly), 0);
    for (int k = 0; k < kVel; ++k) {
        int32_t* fx = &p.cov_x[ox + static_cast<size_t>(k * Lx)];
        int32_t* fy = &p.cov_y[oy + static_cast<size_t>(k * Ly)];
        for (int i = 0; i < n; ++i) {
            const int32_t* sh = &p.shift[(static_cast<size_t>(b) * n + i) * 2 * kVel];
            ++fx[sh[k] - ux0];
            ++fy[sh[kVel + k] - uy0];
        }
        for (long long d = 1; d < Lx; ++d) fx[d] += fx[d - 1];
        for (long long d = 1; d < Ly; ++d) fy[d] += fy[d - 1];
    }
}

}  // namespace

StackPlan make_stack_plan(int n, int w, int h, int nv, const int32_t* dx, const int32_t* dy, int min_count) {
    // the messages of validate_pass (src/backend.cpp:89-108)
    if (n < 1) throw std::runtime_error("stack pass: no frames");
    if (w < 1 || h < 1) throw std::runtime_error("stack pass: frames differ in size");
    if (nv < 1 || !dx || !dy) throw std::runtime_error("stack pass: malformed shift table");
    if (min_count < 1) throw std::runtime_error("stack pass: min_count must be at least 1");
    StackPlan p;
    p.n = n;
    p.w = w;
    p.h = h;
    p.nv = nv;
    p.min_count = min_count;
    const Shifts s{dx, dy, n, w, h};
    p.min_dy = std::numeric_limits<int>::max();
    p.max_dy = std::numeric_limits<int>::min();
    for (int j = 0; j < nv; ++j)
        for (int i = 0; i < n; ++i) {
            p.min_dy = std::min(p.min_dy, s.y(j, i));
            p.max_dy = std::max(p.max_dy, s.y(j, i));
        }
    const int fi = widest_frame(s, nv);
    std::vector<int> idx(static_cast<size_t>(nv));
    std::iota(idx.begin(), idx.end(), 0);
    std::vector<std::vector<int>> groups;
    kd_groups(idx, 0, idx.size(), s, fi, groups);
    for (auto& g : groups) emit_blocks(std::move(g), s, p);
    p.all_box[0] = p.all_box[1] = std::numeric_limits<int>::max();
    p.all_box[2] = p.all_box[3] = std::numeric_limits<int>::min();
    for (int b = 0; b < p.nb; ++b) {
        const int32_t* ub = &p.ubox[static_cast<size_t>(b) * 4];
        p.all_box[0] = std::min(p.all_box[0], ub[0]);
        p.all_box[1] = std::min(p.all_box[1], ub[1]);
        p.all_box[2] = std::max(p.all_box[2], ub[2]);
        p.all_box[3] = std::max(p.all_box[3], ub[3]);
    }
    p.inv_sqrt.assign(static_cast<size_t>(n) + 1, 0.0f);
    for (int c = 1; c <= n; ++c) p.inv_sqrt[static_cast<size_t>(c)] = static_cast<float>(1.0 / std::sqrt(static_cast<double>(c)));
    p.units = static_cast<double>(nv) * n * static_cast<double>(w) * h;
    plan_tiles(p, h);
    return p;
}

void plan_tiles(StackPlan& p, int out_rows) {
    p.tiles_x = (p.w + kTileX - 1) / kTileX;
    p.tiles_y = (out_rows + kTileY - 1) / kTileY;
    const long tiles = static_cast<long>(p.tiles_x) * p.tiles_y;
    // Enough workgroups to fill 256 CUs many times over, and at least 16 segments per
    // tile: the segments of one tile run at the same time on one XCD (the kernel's
    // workgroup order) and read the same frame neighbourhood, so the tile's frames are
    // fetched from HBM about once per 16 workgroups instead of once per workgroup.
    const long target = 8192;
    long nseg = std::max<long>({1, (target + tiles - 1) / tiles, 16});
    nseg = std::min<long>(nseg, p.nb);
    p.blocks_per_seg = static_cast<int>((p.nb + nseg - 1) / nseg);
    p.nseg = (p.nb + p.blocks_per_seg - 1) / p.blocks_per_seg;
}

std::vector<Band> plan_row_bands(int w, int h, int n, int min_dy, int max_dy, double budget_bytes) {
    std::vector<Band> bands;
    const double per_row = 4.0 * w * n;
    const int halo = std::max(max_dy, 0) + std::max(-min_dy, 0);
    int rows = static_cast<int>(budget_bytes / per_row) - halo;
    rows = std::max(1, std::min(rows, h));
    for (int y0 = 0; y0 < h; y0 += rows) {
        Band b;
        b.y0 = y0;
        b.y1 = std::min(h, y0 + rows);
        b.r0 = std::max(0, std::min(h, b.y0 + min_dy));
        b.r1 = std::min(h, std::max(0, b.y1 - 1 + max_dy + 1));
        if (b.r1 <= b.r0) b.r1 = std::min(h, b.r0 + 1);
        bands.push_back(b);
    }
    return bands;
}

PadGeometry pad_geometry(const StackPlan& p, int row0, int row1, int y_begin, int tiles_x, int tiles_y) {
    const int* u = p.all_box;
    PadGeometry g;
    // columns read: [min dx, 64 tiles_x - 1 + max dx] (a tile's staged box ends 63 + its
    // largest dx past the tile start); rows: [y_begin + min dy, y_begin + 16 tiles_y - 1 + max dy]
    g.left = std::max(0, -u[0]);
    g.pitch = g.left + std::max(p.w, kTileX * tiles_x + std::max(0, u[2]));
    const int first = std::min(row0, y_begin + std::min(0, u[1]));
    const int last = std::max(row1, y_begin + kTileY * tiles_y + std::max(0, u[3]));
    g.top = row0 - first;
    g.rows = last - first;
    return g;
}

PadGeometry pad_geometry16(const PadGeometry& g32) {
    PadGeometry g = g32;
    g.left = (g32.left + 1 + 1) & ~1;                 // >= left + 1, even
    const int right = g32.pitch - g32.left + 1;       // columns from x = 0, one more
    g.pitch = (g.left + right + 1) & ~1;              // even
    return g;
}

}  // namespace ftk

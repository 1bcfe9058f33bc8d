// Host planning for the gfx950 stacking kernel (see stack_plan.hpp).
#include "stack_plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <stdexcept>
#include <tuple>
#include <unordered_map>

namespace ftk {

namespace {

// Clamped shifts, motion-major like the input: x(j, i) = clamp(dx[j n + i], -w, w) at
// X[j n + i], so the n frames of one motion are contiguous (a block of 16 motions
// reads 16 short rows).
struct Shifts {
    std::vector<int32_t> X, Y;
    int n, nv;
    Shifts(const int32_t* dx, const int32_t* dy, int n_, int nv_, int w, int h) : n(n_), nv(nv_) {
        const size_t m = static_cast<size_t>(n) * nv;
        X.resize(m);
        Y.resize(m);
        for (size_t q = 0; q < m; ++q) {
            X[q] = std::clamp(dx[q], -w, w);
            Y[q] = std::clamp(dy[q], -h, h);
        }
    }
    const int32_t* xm(int j) const { return &X[static_cast<size_t>(j) * n]; }
    const int32_t* ym(int j) const { return &Y[static_cast<size_t>(j) * n]; }
};

// The frame whose shifts spread most over all motions (for a straight-line grid the
// frame furthest from the reference epoch): its shifts order the motions.
int widest_frame(const Shifts& s) {
    std::vector<int> x0(static_cast<size_t>(s.n), std::numeric_limits<int>::max()), y0 = x0;
    std::vector<int> x1(static_cast<size_t>(s.n), std::numeric_limits<int>::min()), y1 = x1;
    for (int j = 0; j < s.nv; ++j) {
        const int32_t* xr = s.xm(j);
        const int32_t* yr = s.ym(j);
        for (int i = 0; i < s.n; ++i) {
            x0[static_cast<size_t>(i)] = std::min(x0[static_cast<size_t>(i)], xr[i]);
            x1[static_cast<size_t>(i)] = std::max(x1[static_cast<size_t>(i)], xr[i]);
            y0[static_cast<size_t>(i)] = std::min(y0[static_cast<size_t>(i)], yr[i]);
            y1[static_cast<size_t>(i)] = std::max(y1[static_cast<size_t>(i)], yr[i]);
        }
    }
    int best = 0;
    long best_spread = -1;
    for (int i = 0; i < s.n; ++i) {
        const size_t u = static_cast<size_t>(i);
        const long spread = static_cast<long>(x1[u] - x0[u]) + (y1[u] - y0[u]);
        if (spread > best_spread) {
            best_spread = spread;
            best = i;
        }
    }
    return best;
}

// Split the motions into groups of kVel that are compact in the shift space of one
// frame (a k-d split along the wider axis, cut at a multiple of kVel; ties ordered by
// the other axis, then the motion index).
void kd_groups(std::vector<int>& idx, size_t lo, size_t hi, const int32_t* fx, const int32_t* fy,
               std::vector<std::vector<int>>& out) {
    const size_t cnt = hi - lo;
    if (cnt <= static_cast<size_t>(kVel)) {
        out.emplace_back(idx.begin() + static_cast<ptrdiff_t>(lo), idx.begin() + static_cast<ptrdiff_t>(hi));
        return;
    }
    int x0 = std::numeric_limits<int>::max(), x1 = std::numeric_limits<int>::min(), y0 = x0, y1 = x1;
    for (size_t q = lo; q < hi; ++q) {
        x0 = std::min(x0, fx[idx[q]]);
        x1 = std::max(x1, fx[idx[q]]);
        y0 = std::min(y0, fy[idx[q]]);
        y1 = std::max(y1, fy[idx[q]]);
    }
    const bool by_x = (x1 - x0) >= (y1 - y0);
    const int32_t* a1 = by_x ? fx : fy;
    const int32_t* a2 = by_x ? fy : fx;
    const size_t groups = (cnt + kVel - 1) / kVel;
    const size_t mid = lo + ((groups + 1) / 2) * kVel;
    std::nth_element(idx.begin() + static_cast<ptrdiff_t>(lo), idx.begin() + static_cast<ptrdiff_t>(mid),
                     idx.begin() + static_cast<ptrdiff_t>(hi), [&](int a, int b) {
                         if (a1[a] != a1[b]) return a1[a] < a1[b];
                         if (a2[a] != a2[b]) return a2[a] < a2[b];
                         return a < b;
                     });
    kd_groups(idx, lo, mid, fx, fy, out);
    kd_groups(idx, mid, hi, fx, fy, out);
}

// Emit a group as blocks whose per-frame spread is at most spread_max (<= kSpreadMax, so
// it fits the LDS box); groups that do not fit are halved along the worst frame and
// axis (a single motion always fits).
void emit_blocks(std::vector<int> g, const Shifts& s, StackPlan& p, int spread_max) {
    const int n = s.n;
    // per-frame extremes over the group (frames in the inner loop: contiguous rows)
    thread_local std::vector<int> ex;
    ex.assign(static_cast<size_t>(4) * n, 0);
    int* x0 = ex.data();
    int* x1 = x0 + n;
    int* y0 = x1 + n;
    int* y1 = y0 + n;
    std::fill(x0, x0 + n, std::numeric_limits<int>::max());
    std::fill(x1, x1 + n, std::numeric_limits<int>::min());
    std::fill(y0, y0 + n, std::numeric_limits<int>::max());
    std::fill(y1, y1 + n, std::numeric_limits<int>::min());
    for (int j : g) {
        const int32_t* xr = s.xm(j);
        const int32_t* yr = s.ym(j);
        for (int i = 0; i < n; ++i) {
            x0[i] = std::min(x0[i], xr[i]);
            x1[i] = std::max(x1[i], xr[i]);
            y0[i] = std::min(y0[i], yr[i]);
            y1[i] = std::max(y1[i], yr[i]);
        }
    }
    int worst_frame = -1, worst_axis = 0, worst = 0;
    for (int i = 0; i < n; ++i) {
        if (x1[i] - x0[i] > spread_max && x1[i] - x0[i] - spread_max > worst) {
            worst = x1[i] - x0[i] - spread_max;
            worst_frame = i;
            worst_axis = 0;
        }
        if (y1[i] - y0[i] > spread_max && y1[i] - y0[i] - spread_max > worst) {
            worst = y1[i] - y0[i] - spread_max;
            worst_frame = i;
            worst_axis = 1;
        }
    }
    if (worst_frame >= 0) {
        const int32_t* tab = worst_axis ? s.Y.data() : s.X.data();
        auto key = [&](int j) { return tab[static_cast<size_t>(j) * n + worst_frame]; };
        std::sort(g.begin(), g.end(), [&](int a, int b) { return key(a) != key(b) ? key(a) < key(b) : a < b; });
        const size_t half = g.size() / 2;
        emit_blocks(std::vector<int>(g.begin(), g.begin() + static_cast<ptrdiff_t>(half)), s, p, spread_max);
        emit_blocks(std::vector<int>(g.begin() + static_cast<ptrdiff_t>(half), g.end()), s, p, spread_max);
        return;
    }
    // ascending motion order inside a block: the kernel's strict '>' then keeps the
    // lowest index among equal S/N values of the block
    std::sort(g.begin(), g.end());
    const int b = p.nb++;
    p.motion.resize(static_cast<size_t>(p.nb) * kVel, -1);
    p.shift.resize(static_cast<size_t>(p.nb) * n * 2 * kVel, 0);
    p.box.resize(static_cast<size_t>(p.nb) * n * 4, 0);
    p.ubox.resize(static_cast<size_t>(p.nb) * 4, 0);
    p.lds_off.resize(static_cast<size_t>(p.nb) * n * kVel, 0);
    int gj[kVel];
    for (int k = 0; k < kVel; ++k) {
        const bool real = k < static_cast<int>(g.size());
        p.motion[static_cast<size_t>(b) * kVel + k] = real ? g[static_cast<size_t>(k)] : -1;
        gj[k] = g[static_cast<size_t>(real ? k : 0)];  // padding lanes repeat the first motion: inside the box
    }
    int ux0 = std::numeric_limits<int>::max(), uy0 = ux0, ux1 = std::numeric_limits<int>::min(), uy1 = ux1;
    int32_t* shb = &p.shift[static_cast<size_t>(b) * n * 2 * kVel];
    for (int k = 0; k < kVel; ++k) {
        const int32_t* xr = s.xm(gj[k]);
        const int32_t* yr = s.ym(gj[k]);
        for (int i = 0; i < n; ++i) {
            shb[static_cast<size_t>(i) * 2 * kVel + k] = xr[i];
            shb[static_cast<size_t>(i) * 2 * kVel + kVel + k] = yr[i];
        }
    }
    for (int i = 0; i < n; ++i) {
        // the padding lanes repeat a motion of the group, so the group's extremes are
        // the block's
        const int32_t* sh = shb + static_cast<size_t>(i) * 2 * kVel;
        int32_t* bx = &p.box[(static_cast<size_t>(b) * n + i) * 4];
        bx[0] = x0[i];
        bx[1] = y0[i];
        bx[2] = x1[i] - x0[i];
        bx[3] = y1[i] - y0[i];
        p.max_sy = std::max(p.max_sy, y1[i] - y0[i]);
        int32_t* lo = &p.lds_off[(static_cast<size_t>(b) * n + i) * kVel];
        for (int k = 0; k < kVel; ++k) lo[k] = (sh[kVel + k] - y0[i]) * kPitch + (sh[k] - x0[i]);
        ux0 = std::min(ux0, x0[i]);
        ux1 = std::max(ux1, x1[i]);
        uy0 = std::min(uy0, y0[i]);
        uy1 = std::max(uy1, y1[i]);
    }
    int32_t* ub = &p.ubox[static_cast<size_t>(b) * 4];
    ub[0] = ux0;
    ub[1] = uy0;
    ub[2] = ux1;
    ub[3] = uy1;
}

// ------------------------------------------------------------------ lattice blocks
// Motions with the same dx in every frame form an x-class, those with the same dy an
// y-class (for a straight-line velocity grid: one class per vx, and per vy, or fewer
// where neighbouring velocities round alike in every frame). Classes are ordered by their
// shift in the widest frame, then by the whole row. Four consecutive x-classes and four
// consecutive y-classes make a lattice group; motion (x-class 4 ga + a, y-class 4 gb + b)
// takes slot 4 a + b of the group (motions of the same classes have the same shifts in
// every frame: only the lowest index of them is stacked). A layer becomes a lattice block when it holds at least kLatMinFill motions and
// its four y-classes spread at most kLatSyMax rows in every frame; in a lattice block a
// lane reads each of the four shifted columns once for all four y-classes.

namespace {

// Class ids of the rows of a motion-major table (n values per motion): equal rows, equal
// ids. rep[c] is a motion of class c.
void row_classes(const std::vector<int32_t>& T, int n, int nv, std::vector<int>& cls, std::vector<int>& rep) {
    std::unordered_map<uint64_t, int> first;  // hash -> class (collisions probe the next hash)
    first.reserve(static_cast<size_t>(nv) * 2);
    cls.assign(static_cast<size_t>(nv), 0);
    rep.clear();
    for (int j = 0; j < nv; ++j) {
        const int32_t* r = &T[static_cast<size_t>(j) * n];
        uint64_t hsh = 1469598103934665603ull;
        for (int i = 0; i < n; ++i) {
            hsh ^= static_cast<uint32_t>(r[i]);
            hsh *= 1099511628211ull;
        }
        for (;; ++hsh) {
            auto it = first.find(hsh);
            if (it == first.end()) {
                first.emplace(hsh, static_cast<int>(rep.size()));
                cls[static_cast<size_t>(j)] = static_cast<int>(rep.size());
                rep.push_back(j);
                break;
            }
            const int32_t* q = &T[static_cast<size_t>(rep[static_cast<size_t>(it->second)]) * n];
            if (std::memcmp(q, r, sizeof(int32_t) * static_cast<size_t>(n)) == 0) {
                cls[static_cast<size_t>(j)] = it->second;
                break;
            }
        }
    }
}

// Ranks of the classes: by the shift in frame fi, then by the whole row.
std::vector<int> class_ranks(const std::vector<int32_t>& T, int n, int fi, const std::vector<int>& rep) {
    std::vector<int> order(rep.size());
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        const int32_t* ra = &T[static_cast<size_t>(rep[static_cast<size_t>(a)]) * n];
        const int32_t* rb = &T[static_cast<size_t>(rep[static_cast<size_t>(b)]) * n];
        if (ra[fi] != rb[fi]) return ra[fi] < rb[fi];
        return std::lexicographical_compare(ra, ra + n, rb, rb + n);
    });
    std::vector<int> rank(rep.size());
    for (size_t r = 0; r < order.size(); ++r) rank[static_cast<size_t>(order[r])] = static_cast<int>(r);
    return rank;
}

}  // namespace

// Emits the lattice blocks of the plan (as blocks 0, 1, ...) and returns the motions left
// for the general grouping, ascending.
std::vector<int> lattice_blocks(const Shifts& s, int fi, StackPlan& p) {
    const int n = s.n, nv = s.nv;
    std::vector<int> cx, cy, repx, repy;
    row_classes(s.X, n, nv, cx, repx);
    row_classes(s.Y, n, nv, cy, repy);
    const std::vector<int> rx = class_ranks(s.X, n, fi, repx), ry = class_ranks(s.Y, n, fi, repy);
    std::vector<int> by_rank_x(repx.size()), by_rank_y(repy.size());  // rank -> a motion of that class
    for (size_t c = 0; c < repx.size(); ++c) by_rank_x[static_cast<size_t>(rx[c])] = repx[c];
    for (size_t c = 0; c < repy.size(); ++c) by_rank_y[static_cast<size_t>(ry[c])] = repy[c];
    // a motion whose shifts equal a lower-index motion's in every frame (same x- and
    // y-class) is never stacked: its sums and counts are those of the lower-index motion
    // bit for bit, so it never beats it, and ties go to the lower index
    std::map<std::pair<int, int>, int> seen;
    std::vector<char> stack(static_cast<size_t>(nv), 1);
    for (int j = 0; j < nv; ++j)
        if (!seen.emplace(std::make_pair(cx[static_cast<size_t>(j)], cy[static_cast<size_t>(j)]), j).second) {
            stack[static_cast<size_t>(j)] = 0;
            ++p.shadowed;
        }
    // (ga, gb, layer) -> 16 slots
    std::map<std::tuple<int, int, int>, std::vector<int>> groups;
    std::map<std::pair<int, int>, int> layers;
    for (int j = 0; j < nv; ++j) {
        if (!stack[static_cast<size_t>(j)]) continue;
        const int ax = rx[static_cast<size_t>(cx[static_cast<size_t>(j)])];
        const int by = ry[static_cast<size_t>(cy[static_cast<size_t>(j)])];
        const int ga = ax / kLatSide, gb = by / kLatSide, slot = (ax % kLatSide) * kLatSide + by % kLatSide;
        int& nl = layers[{ga, gb}];
        int l = 0;
        for (; l < nl; ++l) {
            auto& g = groups[{ga, gb, l}];
            if (g[static_cast<size_t>(slot)] < 0) {
                g[static_cast<size_t>(slot)] = j;
                break;
            }
        }
        if (l == nl) {
            auto& g = groups[{ga, gb, nl}];
            g.assign(kVel, -1);
            g[static_cast<size_t>(slot)] = j;
            ++nl;
        }
    }
    std::vector<int> left;
    const int nxr = static_cast<int>(repx.size()), nyr = static_cast<int>(repy.size());
    for (const auto& kv : groups) {
        const int ga = std::get<0>(kv.first), gb = std::get<1>(kv.first);
        const std::vector<int>& g = kv.second;
        const int fill = static_cast<int>(std::count_if(g.begin(), g.end(), [](int j) { return j >= 0; }));
        // the classes of the group's columns and rows (past the last class: the first one)
        const int32_t* X[kLatSide];
        const int32_t* Y[kLatSide];
        for (int a = 0; a < kLatSide; ++a) {
            const int r = kLatSide * ga + a < nxr ? kLatSide * ga + a : kLatSide * ga;
            X[a] = s.xm(by_rank_x[static_cast<size_t>(r)]);
        }
        for (int b = 0; b < kLatSide; ++b) {
            const int r = kLatSide * gb + b < nyr ? kLatSide * gb + b : kLatSide * gb;
            Y[b] = s.ym(by_rank_y[static_cast<size_t>(r)]);
        }
        bool ok = fill >= kLatMinFill;
        for (int i = 0; i < n && ok; ++i) {
            int x0 = X[0][i], x1 = X[0][i], y0 = Y[0][i], y1 = Y[0][i];
            for (int a = 1; a < kLatSide; ++a) {
                x0 = std::min(x0, X[a][i]);
                x1 = std::max(x1, X[a][i]);
            }
            for (int b = 1; b < kLatSide; ++b) {
                y0 = std::min(y0, Y[b][i]);
                y1 = std::max(y1, Y[b][i]);
            }
            ok = x1 - x0 <= kSpreadMax && y1 - y0 <= kLatSyMax;
        }
        if (!ok) {
            for (int j : g)
                if (j >= 0) left.push_back(j);
            continue;
        }
        const int b = p.nb++;
        p.motion.resize(static_cast<size_t>(p.nb) * kVel, -1);
        p.shift.resize(static_cast<size_t>(p.nb) * n * 2 * kVel, 0);
        p.box.resize(static_cast<size_t>(p.nb) * n * 4, 0);
        p.ubox.resize(static_cast<size_t>(p.nb) * 4, 0);
        p.lds_off.resize(static_cast<size_t>(p.nb) * n * kVel, 0);
        p.lat.resize(static_cast<size_t>(p.nb) * n * 2, 0);
        for (int k = 0; k < kVel; ++k) p.motion[static_cast<size_t>(b) * kVel + k] = g[static_cast<size_t>(k)];
        int ux0 = std::numeric_limits<int>::max(), uy0 = ux0, ux1 = std::numeric_limits<int>::min(), uy1 = ux1;
        for (int i = 0; i < n; ++i) {
            int x0 = X[0][i], x1 = X[0][i], y0 = Y[0][i], y1 = Y[0][i];
            for (int a = 1; a < kLatSide; ++a) {
                x0 = std::min(x0, X[a][i]);
                x1 = std::max(x1, X[a][i]);
            }
            for (int q = 1; q < kLatSide; ++q) {
                y0 = std::min(y0, Y[q][i]);
                y1 = std::max(y1, Y[q][i]);
            }
            int32_t* sh = &p.shift[(static_cast<size_t>(b) * n + i) * 2 * kVel];
            int32_t* lo = &p.lds_off[(static_cast<size_t>(b) * n + i) * kVel];
            uint32_t wx = 0, wy = 0;
            for (int a = 0; a < kLatSide; ++a) wx |= static_cast<uint32_t>(X[a][i] - x0) << (8 * a);
            for (int q = 0; q < kLatSide; ++q) wy |= static_cast<uint32_t>(Y[q][i] - y0) << (4 * q);
            wy |= static_cast<uint32_t>((y1 - y0 + 1) / 2) << 16;
            for (int a = 0; a < kLatSide; ++a)
                for (int q = 0; q < kLatSide; ++q) {
                    // padding slots take their column's and row's shifts: inside the box
                    const int k = kLatSide * a + q;
                    sh[k] = X[a][i];
                    sh[kVel + k] = Y[q][i];
                    lo[k] = (Y[q][i] - y0) * kPitch + (X[a][i] - x0);
                }
            int32_t* bx = &p.box[(static_cast<size_t>(b) * n + i) * 4];
            bx[0] = x0;
            bx[1] = y0;
            bx[2] = x1 - x0;
            bx[3] = y1 - y0;
            p.max_sy = std::max(p.max_sy, y1 - y0);
            p.max_sy_lat = std::max(p.max_sy_lat, y1 - y0);
            p.lat[(static_cast<size_t>(b) * n + i) * 2] = static_cast<int32_t>(wx);
            p.lat[(static_cast<size_t>(b) * n + i) * 2 + 1] = static_cast<int32_t>(wy);
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
    }
    p.nb_lat = p.nb;
    std::sort(left.begin(), left.end());
    return left;
}

// Cumulative shift counts per block and motion over the block's shift range: coverage
// of edge tiles without walking the frames. One allocation for all blocks; skipped
// (cov_ok = false) when the tables would exceed 2^26 entries.
void build_coverage(StackPlan& p) {
    const int n = p.n;
    long long tx = 0, ty = 0;
    for (int b = 0; b < p.nb; ++b) {
        const int32_t* ub = &p.ubox[static_cast<size_t>(b) * 4];
        tx += kVel * static_cast<long long>(ub[2] - ub[0] + 1);
        ty += kVel * static_cast<long long>(ub[3] - ub[1] + 1);
    }
    if (tx + ty > (1LL << 26)) {
        p.cov_ok = false;
        return;
    }
    p.cov_off.resize(static_cast<size_t>(p.nb) * 2);
    p.cov_x.assign(static_cast<size_t>(tx), 0);
    p.cov_y.assign(static_cast<size_t>(ty), 0);
    p.chain.assign(static_cast<size_t>(p.nb) * kVel, 0);
    std::vector<int> lo_y, hi_y;  // per dx value of one motion: least and largest dy
    long long ox = 0, oy = 0;
    for (int b = 0; b < p.nb; ++b) {
        const int32_t* ub = &p.ubox[static_cast<size_t>(b) * 4];
        const int ux0 = ub[0], uy0 = ub[1];
        const long long Lx = ub[2] - ux0 + 1, Ly = ub[3] - uy0 + 1;
        p.cov_off[2 * static_cast<size_t>(b)] = ox;
        p.cov_off[2 * static_cast<size_t>(b) + 1] = oy;
        int32_t* cx = &p.cov_x[static_cast<size_t>(ox)];
        int32_t* cy = &p.cov_y[static_cast<size_t>(oy)];
        for (int i = 0; i < n; ++i) {
            const int32_t* sh = &p.shift[(static_cast<size_t>(b) * n + i) * 2 * kVel];
            for (int k = 0; k < kVel; ++k) {
                ++cx[k * Lx + sh[k] - ux0];
                ++cy[k * Ly + sh[kVel + k] - uy0];
            }
        }
        for (int k = 0; k < kVel; ++k) {
            for (long long d = 1; d < Lx; ++d) cx[k * Lx + d] += cx[k * Lx + d - 1];
            for (long long d = 1; d < Ly; ++d) cy[k * Ly + d] += cy[k * Ly + d - 1];
        }
        // chains: ordered by dx, the dy ranges of successive dx values never step down
        // (chain) or never step up (anti-chain); a bucket pass over the dx range
        lo_y.assign(static_cast<size_t>(Lx), 0);
        hi_y.assign(static_cast<size_t>(Lx), 0);
        for (int k = 0; k < kVel; ++k) {
            std::fill(lo_y.begin(), lo_y.end(), std::numeric_limits<int>::max());
            std::fill(hi_y.begin(), hi_y.end(), std::numeric_limits<int>::min());
            for (int i = 0; i < n; ++i) {
                const int32_t* sh = &p.shift[(static_cast<size_t>(b) * n + i) * 2 * kVel];
                const size_t d = static_cast<size_t>(sh[k] - ux0);
                lo_y[d] = std::min(lo_y[d], static_cast<int>(sh[kVel + k]));
                hi_y[d] = std::max(hi_y[d], static_cast<int>(sh[kVel + k]));
            }
            bool co = true, anti = true, any = false;
            int plo = 0, phi = 0;
            for (long long d = 0; d < Lx; ++d) {
                if (lo_y[static_cast<size_t>(d)] > hi_y[static_cast<size_t>(d)]) continue;  // no frame at this dx
                const int clo = lo_y[static_cast<size_t>(d)], chi = hi_y[static_cast<size_t>(d)];
                if (any) {
                    co = co && phi <= clo;     // every earlier dy <= every later dy
                    anti = anti && plo >= chi;  // every earlier dy >= every later dy
                }
                // within one dx value the dy order is free in both cases
                plo = clo;
                phi = chi;
                any = true;
            }
            p.chain[static_cast<size_t>(b) * kVel + k] = co ? 1 : anti ? 2 : 0;
        }
        ox += kVel * Lx;
        oy += kVel * Ly;
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
    const Shifts s(dx, dy, n, nv, w, h);
    const auto [ylo, yhi] = std::minmax_element(s.Y.begin(), s.Y.end());
    p.min_dy = *ylo;
    p.max_dy = *yhi;
    const int fi = widest_frame(s);
    std::vector<int32_t> fx(static_cast<size_t>(nv)), fy(static_cast<size_t>(nv));
    for (int j = 0; j < nv; ++j) {
        fx[static_cast<size_t>(j)] = s.xm(j)[fi];
        fy[static_cast<size_t>(j)] = s.ym(j)[fi];
    }
    // lattice blocks first, then the remaining motions in compact groups
    std::vector<int> idx = lattice_blocks(s, fi, p);
    std::vector<std::vector<int>> groups;
    groups.reserve(static_cast<size_t>(nv) / kVel + 1);
    if (!idx.empty()) kd_groups(idx, 0, idx.size(), fx.data(), fy.data(), groups);
    const size_t est = static_cast<size_t>(p.nb) + groups.size() + groups.size() / 4 + 1;  // blocks (some groups split)
    p.motion.reserve(est * kVel);
    p.shift.reserve(est * n * 2 * kVel);
    p.box.reserve(est * n * 4);
    p.ubox.reserve(est * 4);
    p.lds_off.reserve(est * n * kVel);
    // the motions left beside lattice blocks are sparse (the rim of a velocity grid): their
    // blocks are cut smaller so their boxes stay narrow
    const int spread = p.nb_lat > 0 ? kLeftSpreadMax : kSpreadMax;
    for (auto& g : groups) emit_blocks(std::move(g), s, p, spread);
    build_coverage(p);
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
    p.tiles_y = (out_rows + p.tile_rows - 1) / p.tile_rows;
    const long tiles = static_cast<long>(p.tiles_x) * p.tiles_y;
    // Enough work items to fill 256 CUs many times over, and at least 64 segments per
    // tile: the segments of one tile run at the same time on one XCD (the kernel's
    // workgroup order) and read the same frame neighbourhood, so the tile's frames are
    // fetched from HBM about once per 64 workgroups instead of once per workgroup; 64 is
    // about the workgroups an XCD holds at once, so they share few tiles' boxes in its L2
    // (6144 x 4096 x 60 frames x 19 952 motions, tuned shape: 16 segments 1483 ms, 32
    // 1446, 64 1434, 128 1431). (8192 items of 16-row tiles; the same work per item for
    // other tile heights.)
    const long target = 8192L * kTileY / p.tile_rows;
    long nseg = std::max<long>({1, (target + tiles - 1) / tiles, 64});
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
    // largest dx past the tile start); rows: [y_begin + min dy, y_begin + rows tiles_y - 1 + max dy]
    // for tiles of `rows` = tile_rows rows
    g.left = std::max(0, -u[0]);
    g.pitch = g.left + std::max(p.w, kTileX * tiles_x + std::max(0, u[2]));
    // the tiles' rows rounded up to whole tiles of the tallest height (32 rows), so the
    // same buffer holds the reads of every tile height
    const int tiled = (p.tile_rows * tiles_y + kTileRowsMax - 1) / kTileRowsMax * kTileRowsMax;
    const int first = std::min(row0, y_begin + std::min(0, u[1]));
    const int last = std::max(row1, y_begin + tiled + std::max(0, u[3]));
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

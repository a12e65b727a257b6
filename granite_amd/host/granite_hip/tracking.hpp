// Level geometry of a Berger-Oliger hierarchy built from tracking spheres, as
// amr::AMRHierarchy::regrid lays it out (src/amr/amr.cpp:80-301): per level and sphere a
// candidate box of ncells fine cells centred on the sphere, lower corner snapped down onto
// the parent grid, clamped to the domain; overlapping (within one fine cell) candidates merged
// into their union until none overlap; each box one block of round((hi - lo) / dx) cells.
// The arithmetic follows the reference operation for operation, so the boxes come out
// bit-identical (granite_amd/hierarchy.py is the same restatement in Python).
//
// Header-only C++17, host only.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

namespace granite_hip {

struct TrackingSphere {
    std::array<double, 3> center;
    double radius;
    int min_level;
};

struct BoxGeom {
    std::array<int, 3> n;
    std::array<double, 3> lo, hi;
    std::array<double, 3> dx() const {
        return {(hi[0] - lo[0]) / n[0], (hi[1] - lo[1]) / n[1], (hi[2] - lo[2]) / n[2]};
    }
};

using LevelGeoms = std::vector<std::vector<BoxGeom>>;

namespace tracking_detail {

inline void regrid(LevelGeoms& levels, int level, const std::array<int, 3>& nc, const std::array<double, 3>& dlo,
                   const std::array<double, 3>& dhi, const std::vector<TrackingSphere>& spheres, int max_levels,
                   int ratio) {
    if (level < 0 || level >= (int)levels.size() || max_levels <= 1) return;
    const int next = level + 1;
    if (next >= max_levels || levels[level].empty()) return;
    const std::array<double, 3> pdx = levels[level][0].dx();
    std::array<double, 3> cdx, half;
    for (int d = 0; d < 3; ++d) {
        cdx[d] = pdx[d] / (double)ratio;
        half[d] = 0.5 * (double)nc[d] * cdx[d];
    }
    bool active = false;
    for (const auto& s : spheres) active = active || next <= s.min_level;
    if (!active) return;
    if (next >= (int)levels.size()) levels.emplace_back();
    struct Box {
        std::array<double, 3> lo, hi;
    };
    std::vector<Box> c;
    for (const auto& s : spheres) {
        if (next > s.min_level) continue;
        Box b;
        for (int d = 0; d < 3; ++d) {
            const double snapped = std::floor((s.center[d] - half[d] - dlo[d]) / pdx[d]) * pdx[d] + dlo[d];
            b.lo[d] = snapped;
            b.hi[d] = snapped + (double)nc[d] * cdx[d];
            if (b.lo[d] < dlo[d]) {
                b.lo[d] = dlo[d];
                b.hi[d] = b.lo[d] + (double)nc[d] * cdx[d];
            }
            if (b.hi[d] > dhi[d]) {
                b.hi[d] = dhi[d];
                b.lo[d] = b.hi[d] - (double)nc[d] * cdx[d];
                if (b.lo[d] < dlo[d]) b.lo[d] = dlo[d];
            }
        }
        c.push_back(b);
    }
    if (c.empty()) return;
    const double tol = cdx[0];
    for (bool merged = true; merged;) {
        merged = false;
        for (size_t a = 0; a < c.size() && !merged; ++a)
            for (size_t b = a + 1; b < c.size(); ++b) {
                bool ov = true;
                for (int d = 0; d < 3 && ov; ++d)
                    if (c[a].lo[d] > c[b].hi[d] + tol || c[b].lo[d] > c[a].hi[d] + tol) ov = false;
                if (!ov) continue;
                for (int d = 0; d < 3; ++d) {
                    c[a].lo[d] = std::min(c[a].lo[d], c[b].lo[d]);
                    c[a].hi[d] = std::max(c[a].hi[d], c[b].hi[d]);
                    c[a].lo[d] = std::max(c[a].lo[d], dlo[d]);
                    c[a].hi[d] = std::min(c[a].hi[d], dhi[d]);
                }
                c[b] = c.back();
                c.pop_back();
                merged = true;
                break;
            }
    }
    for (const Box& b : c) {
        BoxGeom g;
        for (int d = 0; d < 3; ++d) g.n[d] = std::max((int)std::round((b.hi[d] - b.lo[d]) / cdx[d]), 1);
        g.lo = b.lo;
        g.hi = b.hi;
        bool dup = false;
        for (const BoxGeom& e : levels[next]) {
            const std::array<double, 3> edx = e.dx();
            bool same = true;
            for (int d = 0; d < 3 && same; ++d) {
                const double exp_first = b.lo[d] + 0.5 * cdx[d];
                const double got_first = e.lo[d] + (0.0 + 0.5) * edx[d];
                if (std::abs(edx[d] - cdx[d]) > 1e-10 * cdx[d] || std::abs(got_first - exp_first) > 0.1 * cdx[d])
                    same = false;
            }
            dup = dup || same;
        }
        if (!dup) levels[next].push_back(g);
    }
}

} // namespace tracking_detail

// Blocks per level, coarsest first: the domain block, then one regrid per level that a
// sphere's min_level asks for (AMRHierarchy::initialize, src/amr/amr.cpp:34-78).
inline LevelGeoms tracking_levels(const std::array<int, 3>& ncells, const std::array<double, 3>& dlo,
                                  const std::array<double, 3>& dhi, const std::vector<TrackingSphere>& spheres,
                                  int max_levels, int ratio = 2) {
    LevelGeoms levels(1);
    levels[0].push_back(BoxGeom{ncells, dlo, dhi});
    int deepest = 0;
    for (const auto& s : spheres) deepest = std::max(deepest, s.min_level);
    deepest = std::min(deepest, max_levels - 1);
    for (int l = 0; l < deepest; ++l) {
        tracking_detail::regrid(levels, l, ncells, dlo, dhi, spheres, max_levels, ratio);
        if ((int)levels.size() <= l + 1) break;
    }
    return levels;
}

} // namespace granite_hip

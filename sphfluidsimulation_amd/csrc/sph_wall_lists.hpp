// Wall lists for the density pass of the cell-wave walk: for every cell within
// one cell of a wall and each of its eight octants, the static particles that a
// dynamic particle in that octant can keep, in the order the wall walk visits
// them (the 27 neighbour cells in visit order n = 9a + 3b + c, each cell's static
// bucket in table order, a bucket shared by two neighbour cells once per cell;
// SF:455-496). A wall is listed when it lies within h of the octant's box
// (enlarged by a rounding margin), or when a coordinate is not finite (its
// distance is NaN, which the test keeps). Every wall a particle's exact test can
// keep is therefore listed, and in the walk's order, so the sums are the bucket
// walk's; walls farther than h add nothing in either. A cell with no entry has no
// wall within h. Static particles never move, so the lists are built once at
// initialization, on the host.
//
// Layout of a scene's part of the cell array, from word S.wlBegin: a power-of-two
// open-addressing table of S.wlMask + 1 slots of 4 words (cell x, y, z as
// cell_coord gives them, then entry index + 1; 0 marks an empty slot), probed
// linearly from wall_cell_hash; then 9 words per entry: offsets into the list
// array, octant o's list being [off[o], off[o + 1]). Octant o has bit 0 set for
// the upper half of the cell in x, bit 1 in y, bit 2 in z (wall_octant).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "sph_kernels.hpp"
#include "sph_math.hpp"

namespace sph {

__host__ __device__ inline uint32_t wall_cell_hash(uint32_t cx, uint32_t cy, uint32_t cz) {
    uint32_t v = cx * 0x9E3779B1u ^ cy * 0x85EBCA77u ^ cz * 0xC2B2AE3Du;
    v ^= v >> 15;
    v *= 0x2C1B3C6Du;
    v ^= v >> 12;
    return v;
}

// Octant of a particle inside cell (cx, cy, cz): the half of the cell it lies in
// on each axis, from its offset to the cell corner in FP32. The lists' boxes are
// enlarged enough that this rounding (and that of cell_coord) never puts a
// particle outside its octant's box.
__host__ __device__ inline uint32_t wall_octant(float x, float y, float z, uint32_t cx, uint32_t cy, uint32_t cz,
                                                float h) {
    const float half = 0.5f * h;
    return (x - (float)(int32_t)cx * h >= half ? 1u : 0u) | (y - (float)(int32_t)cy * h >= half ? 2u : 0u) |
           (z - (float)(int32_t)cz * h >= half ? 4u : 0u);
}

// This lane's wall list [*b, *e) for its particle in cell (cx, cy, cz), octant
// oct; empty when the cell has no entry (or the scene no table).
__host__ __device__ inline void find_wall_list(const uint32_t* __restrict__ cells, const SceneDev& S, uint32_t cx,
                                      uint32_t cy, uint32_t cz, uint32_t oct, uint32_t& b, uint32_t& e) {
    b = e = 0u;
    if (S.wlMask == 0u) return;
    const uint4* slots = reinterpret_cast<const uint4*>(cells + S.wlBegin);
    uint32_t k = wall_cell_hash(cx, cy, cz) & S.wlMask;
    for (;;) {
        const uint4 w = slots[k];
        if (w.w == 0u) return;
        if (w.x == cx && w.y == cy && w.z == cz) {
            const uint32_t* off = cells + S.wlBegin + 4u * (S.wlMask + 1u) + 9u * (w.w - 1u) + oct;
            b = off[0];
            e = off[1];
            return;
        }
        k = (k + 1u) & S.wlMask;
    }
}

// Host: one scene's table and lists. spos: every scene's static positions in
// bucket order (global indices); stab: this scene's M_s + 1 bucket starts (global
// indices). Appends the scene's part to `cells` (S.wlBegin receives its first
// word, S.wlMask its slot count - 1; 0 when the scene has no table) and its lists
// to `list` (offsets are indices into the whole list array).
inline void build_wall_lists(SceneDev& S, const float4* spos, const uint32_t* stab, std::vector<uint32_t>& cells,
                             std::vector<uint32_t>& list) {
    S.wlBegin = (uint32_t)cells.size();
    S.wlMask = 0u;
    if (S.sCount == 0u) return;
    const float h = S.h;
    const uint32_t s0 = S.sBegin, s1 = S.sBegin + S.sCount;
    auto finite = [](const float4& p) { return std::isfinite(p.x) && std::isfinite(p.y) && std::isfinite(p.z); };
    // Cells within one cell of a wall (cell_coord values, wrapping as they do).
    struct Cell {
        uint32_t x, y, z;
        bool operator<(const Cell& o) const { return x != o.x ? x < o.x : y != o.y ? y < o.y : z < o.z; }
        bool operator==(const Cell& o) const { return x == o.x && y == o.y && z == o.z; }
    };
    std::vector<Cell> wallCells;
    wallCells.reserve(S.sCount);
    for (uint32_t j = s0; j < s1; ++j)
        if (finite(spos[j]))
            wallCells.push_back({cell_coord(spos[j].x, h), cell_coord(spos[j].y, h), cell_coord(spos[j].z, h)});
    std::sort(wallCells.begin(), wallCells.end());
    wallCells.erase(std::unique(wallCells.begin(), wallCells.end()), wallCells.end());
    std::vector<Cell> near;
    near.reserve(wallCells.size() * 27);
    for (const Cell& w : wallCells)
        for (uint32_t n = 0; n < 27; ++n)
            near.push_back({w.x + n / 9u - 1u, w.y + (n / 3u) % 3u - 1u, w.z + n % 3u - 1u});
    std::sort(near.begin(), near.end());
    near.erase(std::unique(near.begin(), near.end()), near.end());
    if (near.empty()) return;

    // A wall is listed for an octant when within `reach` of the octant's box grown
    // by `grow`: the box covers every particle the kernel assigns to the octant
    // (cell_coord and wall_octant err by a few units in the last place of the
    // coordinate), and a computed r2 <= h2 means a true distance below h (1 +
    // 2^-22), so nothing the exact test keeps is left out.
    const double reach = std::sqrt((double)S.h2) * (1.0 + 0x1p-16);
    const uint32_t nc = (uint32_t)near.size();
    std::vector<std::vector<uint32_t>> per(nc);  // cell c's 8 lists, concatenated
    std::vector<uint32_t> counts((size_t)nc * 8u);
    auto work = [&](uint32_t c0, uint32_t c1) {
        std::vector<uint32_t> oct[8];
        const double hd = (double)h;
        for (uint32_t c = c0; c < c1; ++c) {
            const Cell C = near[c];
            const double lo[3] = {(double)(int32_t)C.x * hd, (double)(int32_t)C.y * hd, (double)(int32_t)C.z * hd};
            double blo[8][3], bhi[8][3];
            for (uint32_t o = 0; o < 8; ++o)
                for (int a = 0; a < 3; ++a) {
                    const double grow = 0x1p-17 * (std::fabs(lo[a]) + 2.0 * hd);
                    const bool up = (o >> a) & 1u;
                    blo[o][a] = lo[a] + (up ? 0.5 * hd : 0.0) - grow;
                    bhi[o][a] = lo[a] + (up ? hd : 0.5 * hd) + grow;
                }
            for (auto& v : oct) v.clear();
            for (uint32_t n = 0; n < 27; ++n) {
                const uint32_t sb = morton3(C.x + n / 9u - 1u, C.y + (n / 3u) % 3u - 1u, C.z + n % 3u - 1u) % S.Ms;
                const uint32_t a = std::max(stab[sb], s0), b = std::min(stab[sb + 1], s1);
                for (uint32_t j = a; j < b; ++j) {
                    const float4 p = spos[j];
                    const bool fin = finite(p);
                    const double q[3] = {p.x, p.y, p.z};
                    for (uint32_t o = 0; o < 8; ++o) {
                        double d2 = 0.0;
                        for (int ax = 0; ax < 3; ++ax) {
                            const double d = q[ax] < blo[o][ax]   ? blo[o][ax] - q[ax]
                                             : q[ax] > bhi[o][ax] ? q[ax] - bhi[o][ax]
                                                                  : 0.0;
                            d2 += d * d;
                        }
                        if (!fin || d2 <= reach * reach) oct[o].push_back(j);
                    }
                }
            }
            std::vector<uint32_t>& out = per[c];
            out.clear();
            for (uint32_t o = 0; o < 8; ++o) {
                counts[(size_t)c * 8u + o] = (uint32_t)oct[o].size();
                out.insert(out.end(), oct[o].begin(), oct[o].end());
            }
        }
    };
    const uint32_t threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (threads == 1u || nc < 4096u) {
        work(0u, nc);
    } else {
        std::vector<std::thread> pool;
        const uint32_t chunk = (nc + threads - 1u) / threads;
        for (uint32_t t = 0; t < threads; ++t) {
            const uint32_t c0 = std::min(nc, t * chunk), c1 = std::min(nc, c0 + chunk);
            if (c0 < c1) pool.emplace_back(work, c0, c1);
        }
        for (auto& th : pool) th.join();
    }

    // Slots: a power of two, at least four per entry (short probe sequences for
    // the cells without an entry, which most lookups are).
    uint32_t slots = 4u;
    while (slots < 4u * nc) slots *= 2u;
    S.wlMask = slots - 1u;
    const size_t base = cells.size();
    cells.resize(base + 4u * (size_t)slots + 9u * (size_t)nc, 0u);
    for (uint32_t c = 0; c < nc; ++c) {
        const Cell C = near[c];
        uint32_t k = wall_cell_hash(C.x, C.y, C.z) & S.wlMask;
        while (cells[base + 4u * (size_t)k + 3u] != 0u) k = (k + 1u) & S.wlMask;
        uint32_t* w = &cells[base + 4u * (size_t)k];
        w[0] = C.x;
        w[1] = C.y;
        w[2] = C.z;
        w[3] = c + 1u;
        uint32_t* off = &cells[base + 4u * (size_t)slots + 9u * (size_t)c];
        uint32_t at = (uint32_t)list.size();
        for (uint32_t o = 0; o < 8; ++o) {
            off[o] = at;
            at += counts[(size_t)c * 8u + o];
        }
        off[8] = at;
        list.insert(list.end(), per[c].begin(), per[c].end());
    }
}

}  // namespace sph

// Cell-group neighbour passes ("cell-wave walk", the default).
//
// Each wave owns 64 consecutive particles in bucket order. Consecutive
// particles with the same cell code form a group; all particles of a group
// share the same 27 neighbour buckets. For each group the wave:
//   1. computes the 27 neighbour buckets on lanes 0..26, in the reference's
//      x-y-z visit order (SF:397-404), with their particle ranges;
//   2. lays the candidates out as four rounds of 64 lanes, up to kSlots (9)
//      particles of each bucket in consecutive positions (position 64 r + lane
//      of round r is slot l % kSlots of bucket l / kSlots), so every lane loads
//      one candidate and the loads are contiguous per bucket;
//   3. tests every candidate against every target of the group, the target's
//      position being wave-uniform: a handful of VALU operations per pair and
//      no per-lane loop bookkeeping. Each (target, round) test is a 64-bit
//      ballot, written into the target's lane.
// After the groups, every lane holds its own particle's neighbour mask (4 x 64
// bits over the candidate layout) and walks only the set bits. The density pass
// stores the mask or a list of the kept candidates; the force pass reads it
// instead of testing again. Buckets with more than kSlots particles (crowded
// cells, buckets shared by two cells) leave their extra particles out of
// the layout; each lane of such a group tests them directly, so every bucket is
// visited exactly as often as in the reference, colliding buckets included.
// Static (wall) neighbours: each group reads the wall-cell bitmap bits of its 27
// neighbour cells with its tables, and only lanes of groups near walls walk the
// static buckets (wall_visits).
// SF = SPHFluidSimulation/source/SPH/Kernels/SPHFunctions.cpp.
#pragma once

#include "sph_kernels.hpp"
#include "sph_wall_lists.hpp"

namespace sph {

constexpr uint32_t kWaves = kBlock / 64;  // 64-particle slices ("waves") of a 256-particle block
constexpr uint32_t kGroupBatch = 16;      // groups whose bucket tables a wave keeps at once
constexpr uint32_t kMaxWG = 512;          // largest workgroup of the launch shapes (8 waves)
constexpr uint32_t kCwMaskWords = 8;      // 4 rounds x 64 bits per particle
// Candidates per bucket in the round layout: position l = 64 r + lane of round r
// holds slot l % kSlots of bucket l / kSlots. Nine slots (243 positions) still
// fit four rounds and take the ninth particle of a compressed cell without the
// per-lane walk of crowded buckets.
#ifndef SPH_CW_SLOTS
#define SPH_CW_SLOTS 9
#endif
constexpr uint32_t kSlots = SPH_CW_SLOTS;
static_assert(kSlots == 8 || kSlots == 9, "27 buckets of kSlots positions must fit four rounds of 64");
// Bucket of layout position l < 27 kSlots (l / 9 as (57 l) >> 9, exact below 243).
__device__ inline uint32_t slot_bucket(uint32_t l) { return kSlots == 8 ? l >> 3 : (l * 57u) >> 9; }

// Per-wave scratch: bucket start and count of the 27 neighbour buckets of up to
// kGroupBatch groups. A workgroup of W waves holds W of them in dynamic LDS
// (wave_tables), so the workgroup size is a launch parameter.
struct WaveTables {
    uint32_t start[kGroupBatch][27];
    uint32_t cnt[kGroupBatch][27];
};

__device__ inline WaveTables& wave_tables(uint32_t wave) {
    extern __shared__ __align__(16) unsigned char sph_dyn_lds[];
    return reinterpret_cast<WaveTables*>(sph_dyn_lds)[wave];
}

// Launch shape of the cell-wave passes. Workgroups hold W = blockDim.x / 64 waves,
// launched in XCD-aware order (xcd_tile over the workgroups); wave gw of the launch
// takes 64-particle slice gw % 4 of 256-particle block gw / 4 (the block in launch
// order: `order`, if given). Every slice keeps its place (block, slice), so the
// results do not depend on the shape. The wave index goes through readfirstlane:
// the compiler then knows it, the block and the scene are wave-uniform and keeps
// them (and every load through them) in scalar registers. Returns false for waves
// past the last block. run > 0: the workgroups take tiles by xcd_tile_runs.
struct WaveSlot {
    uint32_t tile;   // block index in launch order
    uint32_t slice;  // the block's 64-particle slice this wave takes
    uint32_t wave;   // the wave's index in its workgroup
};
__device__ inline bool wave_slot(uint32_t numBlocks, uint32_t numWG, WaveSlot& ws, uint32_t run = 0u) {
    const uint32_t W = blockDim.x >> 6;
    ws.wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t gw = (run ? xcd_tile_runs(blockIdx.x, run) : xcd_tile(blockIdx.x, numWG)) * W + ws.wave;
    ws.tile = gw / kWaves;
    ws.slice = gw % kWaves;
    return ws.tile < numBlocks;
}

__device__ inline uint64_t ballot64(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// A neighbour mask as it is built: 4 rounds x 64 bits held as eight 32-bit halves
// (round r in w[2r] low, w[2r + 1] high), so that writing a lane never has to
// re-pair 64-bit registers.
struct Mask8 {
    uint32_t w[8];
};

// Lane t of round r takes the wave-uniform ballot b[r], r = 0..3 (v_writelane_b32
// with the lane select in m0, since the data operand already uses the one scalar
// operand a VALU instruction may read; the s_nop covers the wait states between
// writing m0 and reading it as a lane select).
// With 8 slots, round 3 holds buckets 24..26 on lanes 0..23 only, so the high
// word of its ballot is always zero and w[7] is never written (it stays zero).
__device__ inline void writelane4(Mask8& m, const uint64_t (&b)[4], uint32_t t) {
    if constexpr (kSlots == 8) {
        asm volatile(
            "s_nop 4\n\t"
            "v_writelane_b32 %0, %7, m0\n\t"
            "v_writelane_b32 %1, %8, m0\n\t"
            "v_writelane_b32 %2, %9, m0\n\t"
            "v_writelane_b32 %3, %10, m0\n\t"
            "v_writelane_b32 %4, %11, m0\n\t"
            "v_writelane_b32 %5, %12, m0\n\t"
            "v_writelane_b32 %6, %13, m0"
            : "+v"(m.w[0]), "+v"(m.w[1]), "+v"(m.w[2]), "+v"(m.w[3]), "+v"(m.w[4]), "+v"(m.w[5]), "+v"(m.w[6])
            : "s"((uint32_t)b[0]), "s"((uint32_t)(b[0] >> 32)), "s"((uint32_t)b[1]), "s"((uint32_t)(b[1] >> 32)),
              "s"((uint32_t)b[2]), "s"((uint32_t)(b[2] >> 32)), "s"((uint32_t)b[3]), "{m0}"(t));
    } else {
        asm volatile(
            "s_nop 4\n\t"
            "v_writelane_b32 %0, %8, m0\n\t"
            "v_writelane_b32 %1, %9, m0\n\t"
            "v_writelane_b32 %2, %10, m0\n\t"
            "v_writelane_b32 %3, %11, m0\n\t"
            "v_writelane_b32 %4, %12, m0\n\t"
            "v_writelane_b32 %5, %13, m0\n\t"
            "v_writelane_b32 %6, %14, m0\n\t"
            "v_writelane_b32 %7, %15, m0"
            : "+v"(m.w[0]), "+v"(m.w[1]), "+v"(m.w[2]), "+v"(m.w[3]), "+v"(m.w[4]), "+v"(m.w[5]), "+v"(m.w[6]),
              "+v"(m.w[7])
            : "s"((uint32_t)b[0]), "s"((uint32_t)(b[0] >> 32)), "s"((uint32_t)b[1]), "s"((uint32_t)(b[1] >> 32)),
              "s"((uint32_t)b[2]), "s"((uint32_t)(b[2] >> 32)), "s"((uint32_t)b[3]), "s"((uint32_t)(b[3] >> 32)),
              "{m0}"(t));
    }
}

// Mixed precision: for each position of the candidate layout (round pair p, lane
// l: rounds 2p and 2p + 1), the cell offset of its bucket from the group's cell
// as FP16 (x: dx of both rounds in the two halves, y: dy, z: dz) and, in w, each
// round's expected cell_tag difference (low half round 2p); positions past the
// 27 buckets take 0xFFFF, which no difference matches. One table per workgroup,
// filled at the start of the density pass.
template <bool kOn>
__device__ inline uint4* slot_cells() {
    if constexpr (kOn) {
        __shared__ uint4 table[2 * 64];
        return table;
    } else {
        return nullptr;
    }
}
__device__ inline void fill_slot_cells(uint4* table) {
    for (uint32_t t = threadIdx.x; t < 128u; t += blockDim.x) {
        const uint32_t p = t >> 6, ln = t & 63u;
        _Float16 d[3][2];
        uint32_t e[2];
        for (uint32_t u = 0; u < 2; ++u) {
            const uint32_t b = (64u * (2u * p + u) + ln) / kSlots;
            const int dx = (int)(b / 9u) - 1, dy = (int)((b / 3u) % 3u) - 1, dz = (int)(b % 3u) - 1;
            const bool in = b < 27u;
            d[0][u] = (_Float16)(in ? dx : 0);
            d[1][u] = (_Float16)(in ? dy : 0);
            d[2][u] = (_Float16)(in ? dz : 0);
            e[u] = in ? cell_tag((uint32_t)dx, (uint32_t)dy, (uint32_t)dz) : 0xFFFFu;
        }
        table[t] = make_uint4(h2_bits(half2v{d[0][0], d[0][1]}), h2_bits(half2v{d[1][0], d[1][1]}),
                              h2_bits(half2v{d[2][0], d[2][1]}), e[0] | (e[1] << 16));
    }
    __syncthreads();
}

// Stored filter: what a candidate's FP16 offsets become (one half of each
// register, for one round). Its cell's tag difference d from the group's equals
// its bucket's (expected): its stored offset plus the bucket's cell offset.
// Else the candidate is a particle of another cell sharing the bucket. Within two
// cells on every axis (each 4-bit field of d + 2 stays below 5 only for a
// difference in -2..2 modulo 16) it becomes NaN, which passes; the walk's exact
// test decides. Farther, its x offset gets 64 instead of the bucket's, which
// fails the test, unless the stored offset is NaN (a NaN coordinate): that
// candidate passes, as the reference keeps NaN distances.
struct HalfMark {
    uint32_t keep, set, far;  // 16-bit masks / bits
};
__device__ inline HalfMark mismatch_mark(uint32_t d, uint32_t expected) {
    // d has clear guard bits: a field's carry stops in its guard bit.
    const uint32_t plus2 = (d + 0x0842u) & kTagFields;  // 2 in every field
    // (field | 16) - 5 keeps its guard bit only for a field >= 5 (5 in every field: 0x14A5).
    const bool near = (((plus2 | kTagGuards) - 0x14A5u) & kTagGuards) == 0u;
    if (d == expected) return {0xFFFFu, 0u, 0u};
    return near ? HalfMark{0u, 0x7E00u, 0u} : HalfMark{0xFFFFu, 0u, 0xFFFFu};
}
constexpr uint32_t kFarOffsetX = 0x54005400u;  // 64.0 in both FP16 halves

// Record j of an array of 8-byte records through a 32-bit byte offset.
__device__ inline uint2 ld8(const uint2* __restrict__ base, uint32_t j) {
    return *reinterpret_cast<const uint2*>(reinterpret_cast<const char*>(base) + (j << 3));
}

// Record j of an array of 16-byte records through a 32-bit byte offset (the
// engine holds fewer than 2^28 particles of each kind), so the load takes a
// scalar base and one 32-bit vector offset instead of 64-bit address arithmetic.
__device__ inline float4 ld16(const float4* __restrict__ base, uint32_t j) {
    return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(base) + (j << 4));
}

// m[r] without dynamic register indexing.
__device__ inline uint64_t pick4(const uint64_t (&m)[4], uint32_t r) {
    return r == 0u ? m[0] : r == 1u ? m[1] : r == 2u ? m[2] : m[3];
}

// An empty asm statement that takes the values as read-write register operands:
// the loads feeding them are issued and complete before this point, which keeps
// the compiler from sinking a batch of loads into the predicated branches that
// use them (one wait for four loads would become four waits).
__device__ inline void settle(uint32_t (&j)[2]) { asm volatile("" : "+v"(j[0]), "+v"(j[1])); }
__device__ inline void settle(uint32_t (&j)[3]) { asm volatile("" : "+v"(j[0]), "+v"(j[1]), "+v"(j[2])); }
__device__ inline void settle(uint32_t (&j)[4]) {
    asm volatile("" : "+v"(j[0]), "+v"(j[1]), "+v"(j[2]), "+v"(j[3]));
}
__device__ inline void settle(uint32_t (&j)[8]) {
    asm volatile("" : "+v"(j[0]), "+v"(j[1]), "+v"(j[2]), "+v"(j[3]), "+v"(j[4]), "+v"(j[5]), "+v"(j[6]), "+v"(j[7]));
}

// Candidates per batch of the mask walk: the density pass loads one float4 per
// candidate, the force pass two, so the force pass holds fewer in registers.
#ifndef SPH_CW_BATCH_DENSITY
#define SPH_CW_BATCH_DENSITY 3
#endif
#ifndef SPH_CW_BATCH_FORCE
#define SPH_CW_BATCH_FORCE 4
#endif
constexpr uint32_t kBatchDensity = SPH_CW_BATCH_DENSITY;
// Density pass, crowded buckets' extra particles, one at a time: a crowded bucket
// mostly holds one to three extras, so larger batches are mostly padding
// (dam-break-1m, 1000 steps: 0.3247 ms a step against 0.3290 with batches of 2).
#ifndef SPH_CW_BATCH_CROWDED
#define SPH_CW_BATCH_CROWDED 1
#endif
constexpr uint32_t kBatchCrowded = SPH_CW_BATCH_CROWDED;
constexpr uint32_t kBatchForce = SPH_CW_BATCH_FORCE;
// Force pass over masks: it also holds the cell groups and the masks in
// registers, so it takes two candidates at a time.
#ifndef SPH_CW_BATCH_FORCE_MASK
#define SPH_CW_BATCH_FORCE_MASK 2
#endif
constexpr uint32_t kBatchForceMask = SPH_CW_BATCH_FORCE_MASK;
// Crowded buckets, mask mode: the density pass records which of a particle's
// first kExtraBits extra candidates are neighbours, for the force pass.
constexpr uint32_t kExtraBits = 64;
// Density pass, walls: walked after the cell groups, when the groups' registers
// are free, so more loads can be in flight.
#ifndef SPH_CW_BATCH_WALL
#define SPH_CW_BATCH_WALL 8
#endif
constexpr uint32_t kBatchWall = SPH_CW_BATCH_WALL;
// Neighbour lists. The density pass records, for the force pass, every
// candidate it keeps (within_h2, j != i) in visit order: up to kListCap per
// particle, slot n of the particle on lane l of wave w (workgroup b, wave
// index q: w = b * kWaves + q) at nbr[(w * kListCap + n) * 64 + l], so a wave
// reads or writes one slot of its 64 particles as one 256-byte row. The count
// goes to nbrCount; kListPoison marks a poisoned particle (NaN results), and a
// count above kListCap makes the force pass walk the 27 buckets itself.
constexpr uint32_t kListCap = 64;
constexpr uint32_t kListPoison = 0xFFFFFFFFu;

// Walks the set bits of a neighbour mask B at a time: f(j) gets B candidate
// indices (bucket start of the bit's bucket + slot), `self` standing in for
// missing ones, so the B loads of a call are issued together. Each lane walks
// its rounds in order and the bits of a round in ascending order, but moves on
// to its next nonzero round as soon as the current one is spent, whatever the
// other lanes do: the wave iterates as often as its busiest lane needs, not
// once per round for the busiest lane of each round. With kFlat a batch also
// crosses from one round into the next (fewer iterations, more work to pick
// each bit); without it a batch stays in one round.
template <uint32_t B, bool kFlat, class F>
__device__ inline void walk_mask(const uint64_t (&m)[4], const uint32_t* start, uint32_t self, F&& f) {
    // Bit r: round r still to walk (its word is nonzero).
    uint32_t left = (m[0] != 0ull ? 1u : 0u) | (m[1] != 0ull ? 2u : 0u) | (m[2] != 0ull ? 4u : 0u) |
                    (m[3] != 0ull ? 8u : 0u);
    uint64_t mm = 0ull;  // the rest of the round being walked
    uint32_t base = 0u;  // its first layout position
    auto refill = [&]() {
        if (mm == 0ull && left != 0u) {
            const uint32_t r = (uint32_t)__builtin_ctz(left);
            left &= left - 1u;
            mm = pick4(m, r);
            base = 64u * r;
        }
    };
    while (true) {
        refill();
        if (mm == 0ull) break;
        uint32_t l[B], b[B], c[B], j[B];
        bool ok[B];
#pragma unroll
        for (uint32_t u = 0; u < B; ++u) {
            if (kFlat && u > 0) refill();
            ok[u] = mm != 0ull;
            l[u] = ok[u] ? base + (uint32_t)__builtin_ctzll(mm) : 0u;
            mm &= mm - 1ull;
        }
#pragma unroll
        for (uint32_t u = 0; u < B; ++u) {
            c[u] = slot_bucket(l[u]);
            b[u] = start[c[u]];  // valid slots: read unconditionally
        }
        settle(b);
#pragma unroll
        for (uint32_t u = 0; u < B; ++u) j[u] = ok[u] ? b[u] + (l[u] - kSlots * c[u]) : self;
        f(j);
    }
}

__device__ inline float readlanef(float v, uint32_t lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)lane));
}

// A neighbour sum: plain FP32, or Kahan-compensated FP32 (mixed-precision mode).
template <bool kKahan>
struct Acc {
    float s = 0.f, c = 0.f;
    __device__ void add(float v) {
        if (kKahan) {
            const float y = v - c;
            const float t = s + y;
            c = (t - s) - y;
            s = t;
        } else {
            s += v;
        }
    }
};


// Mixed-precision candidate filter. Offsets are u = (x - c) / h from the corner c
// of the group's cell, rounded to FP16 (corner and target share the formula, so
// x - c is exact and the corner's own rounding cancels in differences). A target
// in that cell has |u_t| <= 1 (a hair above with rounding) and a true neighbour
// |u_c| < 2 (a hair above), so the rounded offsets are within 2^-11 and 2^-10 of
// their FP32 values and the FP16 difference adds at most 2^-11: 2^-9 per axis in
// all. For a pair with r <= h the computed |u_c - u_t|^2 is then at most
// 1 + 2 sqrt(3) 2^-9 + 3 2^-18 plus five FP16 roundings (<= 1.7e-3): below
// 1.0085. The test against kFilterR2 = 1.02 (1.0195 in FP16) keeps every
// neighbour. Candidates far from the cell get large or infinite offsets and fail
// it, as they should; the exact FP32 test follows in the mask walk.
//
// Stored offsets (SPH_FLAG_FP16_STORED). A particle's offset record
// (offset_record) holds u = (x - c) / h from the corner c = cell * h of its own
// cell, rounded to FP16: |u| <= 1 (a hair above), within 2^-11. A target's offset
// from the group's corner is its own (same cell). A candidate's is its stored u
// plus its bucket's cell offset d in {-1, 0, 1} (slot_cells), added in FP16:
// |u + d| < 2, another 2^-11. The corners are FP32 products, so c_own - c_group
// differs from d h by their two roundings: below 2^-23 4097 h (4.9e-4 h) for
// cells within kStoredFilterReach of the origin, to which the stored filter is
// kept (a farther group passes every candidate). With the FP16 difference (2^-11) the
// error is 2^-9 + 4.9e-4 per axis, and the computed |u_c - u_t|^2 of a pair with
// r <= h is at most 1 + 2 sqrt(3) 2.44e-3 + 3 2.44e-3^2 plus five FP16 roundings:
// below 1.0102, so the same kFilterR2 keeps every neighbour.
constexpr float kFilterR2 = 1.02f;
constexpr uint32_t kStoredFilterReach = 4096u;

// Candidate of round r on this lane: layout position 64 r + lane.
struct RoundCand {
    uint32_t j;
    bool valid;
};

__device__ inline RoundCand round_candidate(uint32_t r, uint32_t lane, uint32_t start27, uint32_t cnt27) {
    const uint32_t l = 64u * r + lane, bi = slot_bucket(l), k = l - kSlots * bi;
    const int src = (int)(bi < 27u ? bi : 0u);
    const uint32_t bs = (uint32_t)__shfl((int)start27, src);
    const uint32_t bc = (uint32_t)__shfl((int)cnt27, src);
    RoundCand c;
    c.valid = bi < 27u && k < bc;
    c.j = bs + k;
    return c;
}

// Table reads of one cell group: its lanes g0..g1-1 and, on lanes 0..26, the
// raw table entries of neighbour cell `lane` in the reference's x-y-z visit
// order (SF:397-404).
struct GroupFetch {
    uint32_t g0, g1, tb, te;  // lanes g0..g1-1; dynamic table entries of neighbour cell `lane`
    uint32_t wword;           // the word of the wall-cell bitmap holding that cell's bit
    bool poison;              // the group's particles visit a bucket holding a NaN particle
    bool nanGroup;            // one of the group's own particles has a NaN coordinate
};

// Bit of myWalls (for_each_group) that marks a poisoned particle.
constexpr uint32_t kPoisonBit = 0x80000000u;

// Neighbour cell `lane` of a group is cell `lane` of the reference's x-y-z
// visit order (SF:397-404) around the group's cell; lanes past 26 repeat cell
// 26. `noff` holds the lane's dilated offsets (neighbour_offsets), and the
// cell's Morton code is neighbour_code of the group's code.
__device__ inline uint32_t lane_offsets(uint32_t lane) { return neighbour_offsets(lane < 27u ? lane : 26u); }

// NaN particles. A particle with a NaN coordinate has the NaN distance to every
// candidate, which the reference keeps (within_h2), so every particle that
// visits its bucket ends the step with NaN density, pressure, velocity and
// position. Fully NaN positions all fall in cell 0, bucket 0 (cell_coord of NaN
// is 0), where they pile up. `nan0` counts the particles with a NaN coordinate
// in bucket 0 of the scene (the force pass counts them for the next step):
// when a group visits bucket 0 and one of them is not the group's own particle
// (nan0 >= 2, or no particle of the group has a NaN coordinate), every particle
// of the group is poisoned, so the walk drops bucket 0 (its range reads as
// empty) instead of walking the pile, and the passes write NaN results.
// NaN particles outside bucket 0 are walked like any other.
template <bool kStatic>
__device__ inline GroupFetch fetch_group(uint64_t& heads, uint64_t stops, uint32_t lane, uint32_t code,
                                         uint32_t noff, const SceneDev& S, const uint32_t* __restrict__ table,
                                         const uint32_t* __restrict__ wallMap, uint32_t nan0, uint64_t nanLanes) {
    GroupFetch f;
    f.g0 = (uint32_t)__builtin_ctzll(heads);
    heads &= heads - 1ull;
    const uint64_t above = f.g0 < 63u ? (stops & ~((2ull << f.g0) - 1ull)) : 0ull;
    f.g1 = above ? (uint32_t)__builtin_ctzll(above) : 64u;
    const uint32_t ncode = neighbour_code((uint32_t)__builtin_amdgcn_readlane((int)code, (int)f.g0), noff);
    const uint32_t bk = fastmod(ncode, S.magicM, S.M);
    const uint64_t groupLanes = (f.g1 < 64u ? (1ull << f.g1) : 0ull) - (1ull << f.g0);
    f.nanGroup = (nanLanes & groupLanes) != 0ull;
    f.poison = nan0 != 0u && (nan0 >= 2u || !f.nanGroup) && ballot64(lane < 27u && bk == 0u) != 0ull;
    // Every lane reads (lanes past 26 repeat cell 26; the reader ignores them), so
    // the reads are not in a branch and a wait for later loads need not cover them.
    // A dropped bucket 0 reads entry 1 twice: an empty range.
    f.tb = table[f.poison && bk == 0u ? 1u : bk];
    f.te = table[bk + 1];
    f.wword = kStatic ? wallMap[S.wmBegin + ((wall_hash(ncode) >> 5) & S.wmMask)] : 0u;
    return f;
}

// Neighbour cells (bits 0..26, visit order) whose static bucket must be walked
// for the group whose first lane is g0: every cell whose wall-bitmap bit is set
// (the bitmap has the bit of every cell holding walls, so a clear bit means no
// walls there), and every cell sharing a static bucket with one of those, since
// the reference walks a bucket once per neighbour cell mapping to it
// (SF:455-496). A skipped bucket holds only walls of cells outside the 27,
// farther than h, which the reference tests and drops: skipping them changes no
// sum. The exception is a particle with a NaN coordinate, whose distances are
// NaN and kept; a group holding one (`all`) walks the static buckets of all 27
// cells.
__device__ inline uint32_t wall_visits(uint32_t wword, uint32_t lane, uint32_t g0, uint32_t code, uint32_t noff,
                                       const SceneDev& S, bool all) {
    if (S.sCount == 0u) return 0u;
    const uint32_t ncode = neighbour_code((uint32_t)__builtin_amdgcn_readlane((int)code, (int)g0), noff);
    const bool hit = lane < 27u && (all || ((wword >> (wall_hash(ncode) & 31u)) & 1u));
    uint32_t hits = (uint32_t)ballot64(hit);
    if (hits == 0u) return 0u;
    const uint32_t sb = fastmod(ncode, S.magicMs, S.Ms);
    uint32_t visits = 0u;
    for (; hits; hits &= hits - 1u) {
        const uint32_t sbn = (uint32_t)__builtin_amdgcn_readlane((int)sb, (int)__builtin_ctz(hits));
        visits |= (uint32_t)ballot64(lane < 27u && sb == sbn);
    }
    return visits;
}

// Iterates the groups of a wave in batches:
// phase1(slot, g0, g1, start27, cnt27, between) runs wave-uniformly per group
// (lanes g0..g1-1 are its particles; lanes 0..26 hold the clamped start and
// count of its neighbour buckets) and calls between() once, after issuing its
// own loads: between() issues the next group's table reads (unconditionally, so
// the wait for the group's own loads never has to include them), so they
// overlap the group's work. Then phase2(slot, crowded) runs on the lanes of the
// batch's groups with `slot` their group's table slot and `crowded` the bit set
// of its buckets holding more than kSlots particles. myWalls receives, on every
// lane of a group, the static buckets its particles walk (wall_visits), or
// kPoisonBit alone for a poisoned group (fetch_group), which walks no walls.
// With kWallLists the caller walks the wall lists instead, and myWalls is set
// (to all 27 cells' buckets) only for a group holding a NaN particle.
// nanLanes: the lanes whose particle has a NaN coordinate.
template <bool kStatic, bool kWallLists, class P1, class P2>
__device__ inline void for_each_group(uint32_t lane, bool valid, uint32_t code, uint32_t noff,
                                      const SceneDev& S, const uint32_t* __restrict__ table,
                                      const uint32_t* __restrict__ wallMap, uint32_t nan0, uint64_t nanLanes,
                                      uint32_t lo, uint32_t hi, WaveTables& gt, uint32_t& myWalls, P1&& phase1,
                                      P2&& phase2) {
    const uint32_t prevCode = (uint32_t)__shfl_up((int)code, 1);
    uint64_t heads = ballot64(valid && (lane == 0u || prevCode != code));
    const uint64_t stops = heads | ~ballot64(valid);  // a group ends at the next head or invalid lane
    while (heads) {
        uint32_t mySlot = 0xFFFFFFFFu, myCrowded = 0u;
        GroupFetch cur =
            fetch_group<kStatic && !kWallLists>(heads, stops, lane, code, noff, S, table, wallMap, nan0, nanLanes);
        for (uint32_t slot = 0;; ++slot) {
            const uint32_t s0 = cur.tb > lo ? cur.tb : lo, e0 = cur.te < hi ? cur.te : hi;
            const uint32_t start27 = s0, cnt27 = lane < 27u && e0 > s0 ? e0 - s0 : 0u;  // empty past lane 26
            if (lane < 27u) {
                gt.start[slot][lane] = start27;
                gt.cnt[slot][lane] = cnt27;
            }
            const uint32_t crowded = (uint32_t)ballot64(cnt27 > kSlots);  // buckets 0..26
            const uint32_t walls = kStatic && !cur.poison && (!kWallLists || cur.nanGroup)
                                       ? wall_visits(cur.wword, lane, cur.g0, code, noff, S, cur.nanGroup)
                                       : 0u;
            if (lane >= cur.g0 && lane < cur.g1) {
                mySlot = slot;
                myCrowded = crowded;
                myWalls = cur.poison ? kPoisonBit : walls;
            }
            const bool more = heads != 0ull && slot + 1u < kGroupBatch;
            GroupFetch nxt;
            uint64_t rest = more ? heads : 1ull;  // at the batch end: a harmless read of group 0's table
            phase1(slot, cur.g0, cur.g1, start27, cnt27,
                   [&]() {
                       nxt = fetch_group<kStatic && !kWallLists>(rest, stops, lane, code, noff, S, table, wallMap,
                                                                 nan0, nanLanes);
                   });
            if (!more) break;
            heads = rest;
            cur = nxt;
        }
        __builtin_amdgcn_wave_barrier();
        if (mySlot != 0xFFFFFFFFu) phase2(mySlot, myCrowded);
        __builtin_amdgcn_wave_barrier();
    }
}

// Walks the static particles of the neighbour cells whose bits are set in
// `walls`, cell by cell in the reference's visit order (SF:455-496 density,
// SF:590-623 forces), each bucket's particles in table order, a bucket shared by
// two neighbour cells once per cell. f(j0, e) takes the batch j0 .. min(j0 + B,
// e) - 1 of one bucket, so its B loads are in flight together; the next
// bucket's table entries are read before the current bucket's particles.
template <uint32_t B, class F>
__device__ inline void walk_walls(uint32_t walls, uint32_t cx, uint32_t cy, uint32_t cz, const SceneDev& S,
                                  const uint32_t* __restrict__ stable, F&& f) {
    if (!walls) return;
    const uint32_t slo = S.sBegin, shi = S.sBegin + S.sCount;
    auto range = [&](uint32_t n, uint32_t& s0, uint32_t& e0) {
        const uint32_t sk = fastmod(morton3(cx + n / 9u - 1u, cy + (n / 3u) % 3u - 1u, cz + n % 3u - 1u), S.magicMs,
                                    S.Ms);
        s0 = stable[sk];
        e0 = stable[sk + 1];
    };
    uint32_t n = (uint32_t)__builtin_ctz(walls);
    walls &= walls - 1u;
    uint32_t s0, e0;
    range(n, s0, e0);
    for (;;) {
        const bool more = walls != 0u;
        uint32_t s1, e1;
        range(more ? (uint32_t)__builtin_ctz(walls) : n, s1, e1);  // the last round re-reads a valid entry
        walls &= walls - 1u;
        const uint32_t a = s0 > slo ? s0 : slo, b = e0 < shi ? e0 : shi;
        for (uint32_t j = a; j < b; j += B) f(j, b);
        if (!more) break;
        s0 = s1;
        e0 = e1;
    }
}

// Launch order of the cell-wave passes' workgroups. A workgroup with particles
// near walls (marked in `heavy` by the density pass) takes several times as long
// as one in the interior, and with few workgroups per SIMD the last of them set
// the pass's duration. So the next step launches them first: tile t of XCD
// chunk k (xcd_tile) runs block order[t], and within each chunk the marked
// blocks come first, in block order, then the rest. One workgroup per chunk;
// clears the marks. Any permutation within the chunks gives the same results;
// this one only changes when each block runs.
__device__ inline void tile_order_chunk(uint32_t k, uint32_t* __restrict__ heavy, uint32_t numBlocks,
                                        uint32_t* __restrict__ order) {
    __shared__ uint32_t sWave[4];
    __shared__ uint32_t sTotal;
    const uint32_t chunk = (numBlocks + 7u) >> 3;
    const uint32_t base = k * chunk;
    const uint32_t valid = base < numBlocks ? min(chunk, numBlocks - base) : 0u;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    // Pass 1: how many marked blocks the chunk holds.
    uint32_t count = 0;
    for (uint32_t off = 0; off < valid; off += 256u) {
        const bool marked = off + tid < valid && heavy[base + off + tid] != 0u;
        count += (uint32_t)__popcll(__ballot(marked));
    }
    if (lane == 0) sWave[wave] = count;
    __syncthreads();
    if (tid == 0) sTotal = sWave[0] + sWave[1] + sWave[2] + sWave[3];
    __syncthreads();
    const uint32_t marks = sTotal;
    // Pass 2: stable partition, marked blocks first.
    uint32_t markedBefore = 0;
    for (uint32_t off = 0; off < valid; off += 256u) {
        const uint32_t b = base + off + tid;
        const bool in = off + tid < valid;
        const bool marked = in && heavy[b] != 0u;
        const uint64_t ball = __ballot(marked);
        const uint32_t below = (uint32_t)__popcll(ball & ((1ull << lane) - 1ull));
        __syncthreads();
        if (lane == 0) sWave[wave] = (uint32_t)__popcll(ball);
        __syncthreads();
        uint32_t waveOff = 0;
        for (uint32_t w = 0; w < wave; ++w) waveOff += sWave[w];
        const uint32_t tileMarks = sWave[0] + sWave[1] + sWave[2] + sWave[3];
        if (in) {
            const uint32_t mk = markedBefore + waveOff + below;  // marked blocks before this one
            const uint32_t pos = marked ? mk : marks + (off + tid - mk);
            order[base + pos] = b;
            heavy[b] = 0u;
        }
        markedBefore += tileMarks;
    }
}

// The scatter (workgroups 0..numBlocks-1) and, on eight more workgroups, the
// next step's launch order (tile_order_chunk k on workgroup numBlocks + k), in
// one launch: both only need this step's density and force passes done.
static_assert(kBlock == 256, "the launch order's workgroups take the scatter's size");
__global__ __launch_bounds__(kBlock) void k_scatter_order(const BlockDesc* __restrict__ blocks, uint32_t numBlocks,
                                                          const SceneDev* __restrict__ scenes,
                                                          const float4* __restrict__ posT,
                                                          const float4* __restrict__ velT,
                                                          const uint32_t* __restrict__ idIn,
                                                          uint32_t* __restrict__ cursor, float4* __restrict__ pos,
                                                          float4* __restrict__ vel, uint32_t* __restrict__ idOut,
                                                          uint32_t* __restrict__ heavy, uint32_t* __restrict__ order,
                                                          uint2* __restrict__ off16) {
    if (blockIdx.x < numBlocks)
        scatter_block(blockIdx.x, blocks, scenes, posT, velT, idIn, cursor, pos, vel, idOut, nullptr, off16);
    else
        tile_order_chunk(blockIdx.x - numBlocks, heavy, numBlocks, order);
}

// Where the density pass writes its per-particle results (read once, at the
// end of the pass, so these pointers hold no registers during the walk).
struct DensityOut {
    uint32_t* heavy;      // workgroups with walls in reach
    uint32_t* nbrCount;   // neighbour list lengths (list mode)
    uint32_t* mask;       // neighbour masks (mask mode)
    uint64_t* extraMask;  // crowded buckets' neighbours (mask mode)
    float4* wallSum;      // walls' geometric sums
    float4* wallZ;        // coincident walls' w1
    float* dens;          // density (debug recording)
    float4* visc;         // viscosity sums over the dynamic neighbours (list mode)
    const uint32_t* wallCells;  // wall lists (sph_wall_lists.hpp): cell tables
    const uint32_t* wallList;   // and octant lists
};

// List mode: the density pass, which visits every kept neighbour, also sums the
// viscosity terms (v_j - v_i) w2(r) (SF:557-589; they need no pressure) in the
// order the force pass would, and stores them; the force pass over the lists
// then gathers only each neighbour's position and pressure (one 16-byte record
// per list entry instead of two). Mask mode keeps the sums in the force pass:
// with them the mask-mode density pass needs 68 instead of 58 VGPRs (7 waves per
// SIMD instead of 8), and dam-break-8m ran 2.17 against 2.02 ms a step, the
// 256-scene batch 3.42 against 2.98 ms.
#ifndef SPH_CW_VISC_LISTS
#define SPH_CW_VISC_LISTS 1
#endif
constexpr bool kViscLists = SPH_CW_VISC_LISTS != 0;
// Density pass, list mode with the viscosity sums: two candidates per batch (the
// velocities take the registers of a third). One per batch fits 64 VGPRs and 8
// waves per SIMD, but loses more to the shorter batches: 1M density 237 against
// 226 us a step over 1000 steps. Three per batch take 80 VGPRs with 2 spilled at 6
// waves (232 us), or 82 at 5 waves (231 us).
#ifndef SPH_CW_BATCH_DENSITY_LIST
#define SPH_CW_BATCH_DENSITY_LIST 2
#endif
constexpr uint32_t kBatchDensityList = kViscLists ? SPH_CW_BATCH_DENSITY_LIST : kBatchDensity;
template <bool kMixed>
struct ViscSums {
    Acc<kMixed> x, y, z;
    float4 vi;
};

// kList: record neighbour lists for k_force_cw; else the neighbour masks and
// the crowded buckets' record for k_force_cw_mask.
// kStored (with kMixed): the FP16 filter reads the stored offset records.
template <bool kStatic, bool kDebug, bool kMixed, bool kList, bool kStored>
#ifndef SPH_CW_DENSITY_WAVES
#define SPH_CW_DENSITY_WAVES 7
#endif
// List mode with the viscosity sums: 6 waves per SIMD, so the velocities fit
// registers without spilling (74 VGPRs; at 7 waves 72 with 2 spilled, and 1M
// took 0.3252 against 0.3195 ms a step over 1000 steps, density 230 against 226 us).
#ifndef SPH_CW_DENSITY_LIST_WAVES
#define SPH_CW_DENSITY_LIST_WAVES 6
#endif
// Mixed precision (Kahan sums): 6 waves per SIMD. With stored offsets the list
// mode takes 5 (at 6 it spills 17 registers: 1M density 232 against 220 us) and
// the mask mode keeps 6 (5 spills outside the loops; the 64-scene batch's density
// 3.25 against 3.38 ms at 5).
#ifndef SPH_CW_DENSITY_MIXED_WAVES
#define SPH_CW_DENSITY_MIXED_WAVES 6
#endif
#ifndef SPH_CW_DENSITY_STORED_LIST_WAVES
#define SPH_CW_DENSITY_STORED_LIST_WAVES 5
#endif
__global__ __launch_bounds__(kMaxWG, kMixed ? (kStored && kList ? SPH_CW_DENSITY_STORED_LIST_WAVES : SPH_CW_DENSITY_MIXED_WAVES) : (kList && kViscLists ? SPH_CW_DENSITY_LIST_WAVES : SPH_CW_DENSITY_WAVES)) void k_density_cw(const BlockDesc* __restrict__ blocks, uint32_t numBlocks,
                                                       const SceneDev* __restrict__ scenes, float4* pos,
                                                       const float4* __restrict__ vel,
                                                       const uint32_t* __restrict__ tab,
                                                       const float4* __restrict__ spos,
                                                       const uint32_t* __restrict__ stab,
                                                       const uint32_t* __restrict__ wallMap,
                                                       const uint32_t* __restrict__ order,
                                                       uint32_t* __restrict__ nbr,
                                                       const uint32_t* __restrict__ nanCnt,
                                                       const DensityOut* __restrict__ out,
                                                       const uint2* __restrict__ off16, uint32_t numWG) {
    static_assert(kMixed || !kStored, "stored offsets are a mixed-precision filter");
    uint4* const slotCell = slot_cells<kStored>();
    if constexpr (kStored) fill_slot_cells(slotCell);  // every wave, before any leaves
    WaveSlot ws;
    if (!wave_slot(numBlocks, numWG, ws)) return;  // only wave-level sync below
    WaveTables& gt = wave_tables(ws.wave);
    const uint32_t blk = order[ws.tile];
    const BlockDesc bd = blocks[blk];
    const SceneDev& S = scenes[bd.scene];
    const uint32_t lane = __lane_id(), wave = ws.slice;  // the block's 64-particle slice
    const uint32_t wbegin = bd.begin + 64u * wave;
    if (wbegin >= bd.end) return;  // whole wave leaves
    const uint32_t wcount = min(64u, bd.end - wbegin);
    const uint32_t i = wbegin + lane;
    const bool valid = lane < wcount;
    const float4 xi = ld16(pos, valid ? i : wbegin);
    const float h = S.h, h2 = S.h2;
    const float h2wide = S.h2wide;
    const uint32_t cx = cell_coord(xi.x, h), cy = cell_coord(xi.y, h), cz = cell_coord(xi.z, h);
    const uint32_t code = morton3(cx, cy, cz);
    const uint32_t* table = tab + S.tBegin;
    const uint32_t* stable = stab + S.stBegin;
    const uint32_t lo = S.pBegin, hi = S.pBegin + S.pCount;

    Acc<kMixed> sumDyn, sumSt;
    // For the force pass, the walls' geometric sums (SF:590-623 without the
    // particle's own |p|, which the force pass multiplies in): sum of w1(r) dir,
    // sum of w2(r), and the w1 of coincident pairs.
    Acc<kMixed> wgx, wgy, wgz;
    float wW2 = 0.f, wZero = 0.f;
    bool wAny = false;
    Mask8 m8 = {{0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u}};  // this lane's neighbour mask, rounds 0..3
    // Mixed precision: this lane's cell corner and its FP16 offset from it, each
    // coordinate twice (both halves), for the candidate filter.
    const float invh = 1.0f / h;
    uint32_t tux = 0u, tuy = 0u, tuz = 0u;
    if (kMixed) {
        const float ux = corner_offset(xi.x, cx, h) * invh, uy = corner_offset(xi.y, cy, h) * invh,
                    uz = corner_offset(xi.z, cz, h) * invh;
        tux = h2_bits(pack_h2(ux, ux));
        tuy = h2_bits(pack_h2(uy, uy));
        tuz = h2_bits(pack_h2(uz, uz));
    }
    // Exact contributions of a batch of candidates j (SF:426-453). The kept ones
    // (r2 <= h2, j != i) are appended to this particle's list one batch later,
    // after the next batch's loads are issued, so the wait for those loads does
    // not include the stores.
    uint32_t* const list = nbr + (size_t)(blk * kWaves + wave) * (kListCap * 64u) + lane;
    uint32_t nc = 0;                      // kept candidates so far
    uint32_t pend[4] = {0u, 0u, 0u, 0u};  // the last batch, not yet appended
    uint32_t pendHits = 0u;               // its kept candidates (bit u)
    auto flush = [&]() {
        if constexpr (!kList) return;
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            const bool hit = (pendHits >> u) & 1u;
            if (hit && nc < kListCap) list[nc << 6] = pend[u];
            nc += hit ? 1u : 0u;
        }
        pendHits = 0u;
    };
    constexpr bool kVisc = kList && kViscLists;
    auto dn = [&](const auto& j, ViscSums<kMixed>& vs) -> uint32_t {
        constexpr uint32_t B = sizeof(j) / sizeof(j[0]);
        static_assert(B <= 4, "at most four pending candidates");
        float4 xj[B], vj[B];
#pragma unroll
        for (uint32_t u = 0; u < B; ++u) {
            xj[u] = ld16(pos, j[u]);
            if constexpr (kVisc) vj[u] = ld16(vel, j[u]);
        }
        flush();
        uint32_t hits = 0u;
#pragma unroll
        for (uint32_t u = 0; u < B; ++u) {
            const float dx = xj[u].x - xi.x, dy = xj[u].y - xi.y, dz = xj[u].z - xi.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            const float r = __builtin_amdgcn_sqrtf(r2);
            const float q = h - r;
            const bool ok = within_h2(r2, h2) && j[u] != i;
            const bool in = inside_h(r, h);
            sumDyn.add((ok && in) ? q * q * q : 0.f);
            if constexpr (kVisc) {
                // ForceSums::pair's viscosity terms, unchanged.
                const float w2 = ok && in ? 6.0f * q : 0.f;
                vs.x.add((vj[u].x - vs.vi.x) * w2);
                vs.y.add((vj[u].y - vs.vi.y) * w2);
                vs.z.add((vj[u].z - vs.vi.z) * w2);
            }
            hits |= ok ? 1u << u : 0u;
            if constexpr (kList) pend[u] = j[u];
        }
        if constexpr (kList) pendHits = hits;
        return hits;
    };

    // Static neighbours (SF:455-496): a batch of kBatchWall walls j, of which the
    // first `count` are real (the rest repeat one of them).
    auto dWallBatch = [&](const uint32_t (&j)[kBatchWall], uint32_t count) {
        float4 xj[kBatchWall];
#pragma unroll
        for (uint32_t u = 0; u < kBatchWall; ++u) xj[u] = ld16(spos, j[u]);
#pragma unroll
        for (uint32_t u = 0; u < kBatchWall; ++u) {
            const float dx = xj[u].x - xi.x, dy = xj[u].y - xi.y, dz = xj[u].z - xi.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            const float r = __builtin_amdgcn_sqrtf(r2);
            const float q = h - r;
            if (u < count && within_h2(r2, h2)) {
                const bool in = inside_h(r, h);
                sumSt.add(in ? q * q * q : 0.f);
                const bool zero = r == 0.f;  // r2 == 0 implies r == 0
                const float inv = zero ? 0.f : __builtin_amdgcn_rcpf(r);
                const float w1 = in ? -3.0f * q * q : 0.f;
                wgx.add((dx * inv) * w1);
                wgy.add((dy * inv) * w1);
                wgz.add((dz * inv) * w1);
                wZero += zero ? w1 : 0.f;
                wAny |= zero;
                wW2 += in ? 6.0f * q : 0.f;
            }
        }
    };
    // A static bucket's walls [j0, e) (walk_walls).
    auto dWalls = [&](uint32_t j0, uint32_t e) {
        uint32_t j[kBatchWall];
#pragma unroll
        for (uint32_t u = 0; u < kBatchWall; ++u) j[u] = j0 + u < e ? j0 + u : j0;
        dWallBatch(j, e - j0);
    };
    uint32_t walls = 0u;  // static buckets this lane walks (groups with a NaN particle), or kPoisonBit
    const uint64_t nanLanes = ballot64(valid && (xi.x != xi.x || xi.y != xi.y || xi.z != xi.z));
    for_each_group<kStatic, true>(
        lane, valid, code, lane_offsets(lane), S, table, wallMap, nanCnt[bd.scene], nanLanes, lo, hi, gt, walls,
        [&](uint32_t, uint32_t g0, uint32_t g1, uint32_t start27, uint32_t cnt27, auto&& between) {
            // The particle itself passes its own test (r = 0); dn skips it. Empty
            // slots are masked out with the (wave-uniform) slot masks.
            // The candidates' positions; with stored offsets their offset records
            // (8 bytes instead of 16), and positions only for the FP32 test.
            float4 xc[4];
            uint2 rc[4];
            uint64_t occupied[4];
#pragma unroll
            for (uint32_t r = 0; r < 4; ++r) {
                const RoundCand c = round_candidate(r, lane, start27, cnt27);
                if constexpr (kStored)
                    rc[r] = ld8(off16, c.valid ? c.j : lo);
                else
                    xc[r] = ld16(pos, c.valid ? c.j : lo);
                occupied[r] = ballot64(c.valid);
            }
            between();
            if (kMixed) {
                // FP16 filter, unless the group's particles span aliased cells (same
                // Morton code, cells 1024 apart) or, with stored offsets, lie too far
                // from the origin for their error bound (kFilterR2).
                const uint32_t gx = (uint32_t)__builtin_amdgcn_readlane((int)cx, (int)g0);
                const uint32_t gy = (uint32_t)__builtin_amdgcn_readlane((int)cy, (int)g0);
                const uint32_t gz = (uint32_t)__builtin_amdgcn_readlane((int)cz, (int)g0);
                const bool inGroup = lane >= g0 && lane < g1;
                const bool nearOrigin = !kStored || ((gx + kStoredFilterReach) | (gy + kStoredFilterReach) |
                                                     (gz + kStoredFilterReach)) < 2u * kStoredFilterReach;
                if (nearOrigin && ballot64(inGroup && (cx != gx || cy != gy || cz != gz)) == 0ull) {
                    half2v X[2], Y[2], Z[2];  // rounds (0, 1) and (2, 3)
                    if constexpr (kStored) {
                        // A candidate's offset from the group's cell corner is its stored
                        // offset plus its bucket's cell offset (slotCell), when its tag says
                        // it lies in that cell (modulo 16 per axis). A particle of another
                        // cell sharing the bucket is made NaN, which passes (the walk's exact
                        // test decides), when its tag puts it within two cells of the group
                        // on every axis, and infinite, which fails, when farther: more than
                        // one cell from the group's cell is farther than h from its targets
                        // except for cells met by rounding at distance h, which the two-cell
                        // reach still passes.
                        const uint32_t gtag = cell_tag(gx, gy, gz);
#pragma unroll
                        for (uint32_t p = 0; p < 2; ++p) {
                            // Read per group: the same for every group, but hoisted out of
                            // the group loop these registers would stay live through it.
                            uint32_t slotIdx = 64u * p + lane;
                            asm volatile("" : "+v"(slotIdx));
                            const uint4 k = slotCell[slotIdx];
                            const uint2 a = rc[2 * p], b = rc[2 * p + 1];
                            const uint32_t x = __builtin_amdgcn_perm(b.x, a.x, 0x05040100u);
                            const uint32_t y = __builtin_amdgcn_perm(b.x, a.x, 0x07060302u);
                            const uint32_t z = __builtin_amdgcn_perm(b.y, a.y, 0x05040100u);
                            const uint32_t d0 = ((a.y >> 16) - gtag) & kTagFields;
                            const uint32_t d1 = ((b.y >> 16) - gtag) & kTagFields;
                            const HalfMark m0 = mismatch_mark(d0, k.w & 0xFFFFu), m1 = mismatch_mark(d1, k.w >> 16);
                            const uint32_t keep = m0.keep | (m1.keep << 16), set = m0.set | (m1.set << 16);
                            const uint32_t far = m0.far | (m1.far << 16);
                            const uint32_t dx = (kFarOffsetX & far) | (k.x & ~far);
                            X[p] = h2_from((h2_bits(h2_from(x) + h2_from(dx)) & keep) | set);
                            Y[p] = h2_from((h2_bits(h2_from(y) + h2_from(k.y)) & keep) | set);
                            Z[p] = h2_from((h2_bits(h2_from(z) + h2_from(k.z)) & keep) | set);
                        }
                    } else {
                        // The same corner as the lanes' own offsets above (corner_offset).
#pragma unroll
                        for (uint32_t p = 0; p < 2; ++p) {
                            X[p] = pack_h2(corner_offset(xc[2 * p].x, gx, h) * invh,
                                           corner_offset(xc[2 * p + 1].x, gx, h) * invh);
                            Y[p] = pack_h2(corner_offset(xc[2 * p].y, gy, h) * invh,
                                           corner_offset(xc[2 * p + 1].y, gy, h) * invh);
                            Z[p] = pack_h2(corner_offset(xc[2 * p].z, gz, h) * invh,
                                           corner_offset(xc[2 * p + 1].z, gz, h) * invh);
                        }
                    }
                    const _Float16 R = (_Float16)kFilterR2;
                    for (uint32_t t = g0; t < g1; ++t) {
                        const half2v TX = h2_from((uint32_t)__builtin_amdgcn_readlane((int)tux, (int)t));
                        const half2v TY = h2_from((uint32_t)__builtin_amdgcn_readlane((int)tuy, (int)t));
                        const half2v TZ = h2_from((uint32_t)__builtin_amdgcn_readlane((int)tuz, (int)t));
                        uint64_t b[4];
#pragma unroll
                        for (uint32_t p = 0; p < 2; ++p) {
                            const half2v dx = X[p] - TX, dy = Y[p] - TY, dz = Z[p] - TZ;
                            const half2v d2 = dx * dx + dy * dy + dz * dz;
                            b[2 * p] = ballot64(!(d2.x > R)) & occupied[2 * p];  // NaN passes
                            b[2 * p + 1] = ballot64(!(d2.y > R)) & occupied[2 * p + 1];
                        }
                        writelane4(m8, b, t);
                    }
                    return;
                }
                if constexpr (kStored) {
                    // Aliased cells or far from the origin: every candidate passes (a
                    // superset; the walk tests each exactly), so this kernel never
                    // holds the candidates' positions.
                    for (uint32_t t = g0; t < g1; ++t) writelane4(m8, occupied, t);
                    return;
                }
            }
            for (uint32_t t = g0; t < g1; ++t) {
                const float px = readlanef(xi.x, t), py = readlanef(xi.y, t), pz = readlanef(xi.z, t);
                uint64_t b[4];
#pragma unroll
                for (uint32_t r = 0; r < 4; ++r) {
                    const float dx = xc[r].x - px, dy = xc[r].y - py, dz = xc[r].z - pz;
                    b[r] = ballot64(within_h2(dx * dx + dy * dy + dz * dz, h2wide)) & occupied[r];
                    // Two rounds at a time: fewer temporaries live at once.
                    if (r == 1u) __builtin_amdgcn_sched_barrier(0);
                }
                writelane4(m8, b, t);
            }
        },
        [&](uint32_t slot, uint32_t crowded) {
            uint64_t m[4];
#pragma unroll
            for (uint32_t r = 0; r < 4; ++r) m[r] = ((uint64_t)m8.w[2 * r + 1] << 32) | m8.w[2 * r];
            ViscSums<kMixed> vs;
            if constexpr (kVisc) vs.vi = ld16(vel, i);
            auto dnv = [&](const auto& j) { return dn(j, vs); };
            walk_mask<kList ? kBatchDensityList : kBatchDensity, false>(m, gt.start[slot], i, dnv);
            // Crowded buckets: their particles past the kSlots-th, a batch at a time.
            // Mask mode: extraMask[i] bit k says extra particle k (enumerated bucket
            // by bucket in visit order) is kept; all ones when there are more than
            // kExtraBits extras.
            uint32_t k = 0;
            uint64_t bits = 0ull;
            const bool anyCrowded = crowded != 0u;
            for (; crowded; crowded &= crowded - 1u) {
                const uint32_t bi = (uint32_t)__builtin_ctz(crowded);
                const uint32_t s0 = gt.start[slot][bi], e = s0 + gt.cnt[slot][bi];
                for (uint32_t jj = s0 + kSlots; jj < e; jj += kBatchCrowded) {
                    uint32_t j[kBatchCrowded];
#pragma unroll
                    for (uint32_t u = 0; u < kBatchCrowded; ++u) j[u] = jj + u < e ? jj + u : i;
                    const uint32_t hits = dn(j, vs);  // padded entries (j = i) are never kept
                    if (!kList) {
                        if (k < kExtraBits) bits |= (uint64_t)hits << k;
                        k += min(kBatchCrowded, e - jj);
                    }
                }
            }
            flush();
            if (!kList && anyCrowded) out->extraMask[i] = k > kExtraBits ? ~0ull : bits;
            if constexpr (kVisc) out->visc[i] = make_float4(vs.x.s, vs.y.s, vs.z.s, 0.f);
        });
    // The walls: the lane's wall list, or (a group holding a NaN particle) the
    // static buckets of all 27 neighbour cells; none for a poisoned particle.
    uint32_t wl0 = 0u, wl1 = 0u;
    if (kStatic && valid && walls == 0u)
        find_wall_list(out->wallCells, S, cx, cy, cz, wall_octant(xi.x, xi.y, xi.z, cx, cy, cz, h), wl0, wl1);
    if (kStatic && ballot64(valid && ((walls & ~kPoisonBit) != 0u || wl1 > wl0)) != 0ull && lane == 0u)
        out->heavy[blk] = 1u;
    if (!valid) return;
    if (kList) {
        out->nbrCount[i] = (walls & kPoisonBit) ? kListPoison : nc;
    } else {
        uint4* mrow = reinterpret_cast<uint4*>(out->mask + (size_t)i * kCwMaskWords);
        mrow[0] = make_uint4(m8.w[0], m8.w[1], m8.w[2], m8.w[3]);
        mrow[1] = make_uint4(m8.w[4], m8.w[5], m8.w[6], m8.w[7]);
    }
    if (kStatic) {
        walk_walls<kBatchWall>(walls & ~kPoisonBit, cx, cy, cz, S, stable, dWalls);
        const uint32_t* __restrict__ wallList = out->wallList;
        for (uint32_t k = wl0; k < wl1; k += kBatchWall) {
            uint32_t j[kBatchWall];
#pragma unroll
            for (uint32_t u = 0; u < kBatchWall; ++u) j[u] = wallList[k + u < wl1 ? k + u : k];
            dWallBatch(j, wl1 - k);
        }
    }
    // A poisoned particle's density is NaN, as the pile it skipped makes it.
    const float rho = (walls & kPoisonBit) ? __builtin_nanf("")
                                           : S.selfDensity + (sumDyn.s * S.mass + sumSt.s * S.mass) * S.C;
    // Only this lane writes pos[i].w; no lane of this pass reads any .w.
    reinterpret_cast<float*>(pos)[4ull * i + 3] = S.k * (rho - S.rho0);
    if (kDebug) out->dens[i] = rho;
    if (kStatic) {
        // w holds the sum of w2 with its sign bit set when a coincident pair
        // occurred (the sum itself is never negative); then wallZ holds its w1.
        out->wallSum[i] = make_float4(wgx.s, wgy.s, wgz.s, wAny ? -wW2 : wW2);
        if (wAny) out->wallZ[i] = make_float4(wZero, 0.f, 0.f, 0.f);
    }
}

// Force pass of the cell-wave walk: the neighbour sums of one particle
// (SF:557-589) and, in force_finish, the walls, the coincident-pair direction,
// the force and the integration with the new key and histogram (SF:590-656).
// kVisc: the density pass summed the viscosity terms (list mode), so the
// neighbour walk takes only positions with pressure and adds the pressure terms.
template <bool kMixed, bool kVisc>
struct ForceSums {
    float4 xi, vi;   // the particle (pressure in xi.w)
    uint32_t i;      // its index
    float h, h2;
    Acc<kMixed> fpx, fpy, fpz, fvx, fvy, fvz;
    // Pressure weights of coincident pairs (r == 0): their direction is the
    // particle's random direction (SF:566-570), the same for every such pair, so
    // it is evaluated once after the walks.
    float zeroSp = 0.f;
    bool anyZero = false;

    // One dynamic candidate j with its position (pressure in w) and velocity
    // loaded. Candidates out of range (or j == i) add exact zeros: their
    // pressure weight is selected away, since a NaN pressure times a zero weight
    // would not be zero. A kept candidate contributes as in the reference, a zero
    // kernel weight times its terms included.
    __device__ void pair(uint32_t j, const float4 xj, const float4 vj) {
        const float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const bool ok = j != i && within_h2(r2, h2);
        const float r = __builtin_amdgcn_sqrtf(r2);
        const bool zero = r == 0.f;  // r2 == 0 implies r == 0
        const float inv = zero ? 0.f : __builtin_amdgcn_rcpf(r);
        const float q = h - r;
        const bool in = inside_h(r, h);
        const float w1 = in ? -3.0f * q * q : 0.f;
        const float w2 = ok && in ? 6.0f * q : 0.f;
        const float sp = ok ? (xi.w + xj.w) * w1 : 0.f;
        fpx.add((dx * inv) * sp);
        fpy.add((dy * inv) * sp);
        fpz.add((dz * inv) * sp);
        zeroSp += ok && zero ? sp : 0.f;
        anyZero |= ok && zero;
        fvx.add((vj.x - vi.x) * w2);
        fvy.add((vj.y - vi.y) * w2);
        fvz.add((vj.z - vi.z) * w2);
    }
    // The density pass's viscosity sums for this particle (kVisc).
    __device__ void take_visc(const float4* __restrict__ visc) {
        if constexpr (kVisc) {
            const float4 v = ld16(visc, i);
            fvx.s = v.x;
            fvy.s = v.y;
            fvz.s = v.z;
        }
    }
    // The pressure terms of pair alone (kVisc: the viscosity sums come from the
    // density pass).
    __device__ void pair_p(uint32_t j, const float4 xj) {
        const float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const bool ok = j != i && within_h2(r2, h2);
        const float r = __builtin_amdgcn_sqrtf(r2);
        const bool zero = r == 0.f;  // r2 == 0 implies r == 0
        const float inv = zero ? 0.f : __builtin_amdgcn_rcpf(r);
        const float q = h - r;
        const bool in = inside_h(r, h);
        const float w1 = in ? -3.0f * q * q : 0.f;
        const float sp = ok ? (xi.w + xj.w) * w1 : 0.f;
        fpx.add((dx * inv) * sp);
        fpy.add((dy * inv) * sp);
        fpz.add((dz * inv) * sp);
        zeroSp += ok && zero ? sp : 0.f;
        anyZero |= ok && zero;
    }
    // A batch of candidates, their loads issued together.
    template <class J>
    __device__ void batch(const float4* __restrict__ pos, const float4* __restrict__ vel, const J& j) {
        if constexpr (kVisc) {
            constexpr uint32_t B = sizeof(j) / sizeof(j[0]);
            float4 xj[B];
#pragma unroll
            for (uint32_t u = 0; u < B; ++u) xj[u] = ld16(pos, j[u]);
#pragma unroll
            for (uint32_t u = 0; u < B; ++u) pair_p(j[u], xj[u]);
            return;
        }
        constexpr uint32_t B = sizeof(j) / sizeof(j[0]);
        float4 xj[B], vj[B];
#pragma unroll
        for (uint32_t u = 0; u < B; ++u) {
            xj[u] = ld16(pos, j[u]);
            vj[u] = ld16(vel, j[u]);
        }
#pragma unroll
        for (uint32_t u = 0; u < B; ++u) pair(j[u], xj[u], vj[u]);
    }
};

// The rest of the force pass for particle f.i of scene `scene` (valid lanes
// only; the active lanes of the wave are a prefix of it). `poisoned`: the
// density pass skipped the NaN pile for it, so its acceleration is NaN, as the
// pile makes it.
template <bool kStatic, bool kDebug, bool kTri, bool kMixed, bool kVisc>
__device__ inline void force_finish(ForceSums<kMixed, kVisc>& f, const SceneDev& S, uint32_t scene, uint32_t lane,
                                    bool poisoned, const float4* __restrict__ wallSum,
                                    const float4* __restrict__ wallZ, const uint32_t* __restrict__ ids,
                                    const float* __restrict__ tris, float dt, float4* __restrict__ posT,
                                    float4* __restrict__ velT, uint32_t* __restrict__ counts,
                                    float4* __restrict__ accOut, uint32_t* __restrict__ nanNext) {
    const uint32_t i = f.i;
    const float4 xi = f.xi, vi = f.vi;
    const float pi = xi.w;
    const float rho = pi / S.k + S.rho0;  // SF:544
    const float absP = fabsf(pi);
    Acc<kMixed> fsx, fsy, fsz;
    float sumD2s = 0.f;
    if (kStatic) {
        // Walls (SF:590-623): |p_i| times the density pass's geometric sums.
        const float4 ws = ld16(wallSum, i);
        fsx.add(absP * ws.x);
        fsy.add(absP * ws.y);
        fsz.add(absP * ws.z);
        sumD2s = fabsf(ws.w);
        if (__builtin_signbit(ws.w)) {
            f.zeroSp += absP * wallZ[i].x;
            f.anyZero = true;
        }
    }
    if (f.anyZero) {
        const float3 d = random_direction((float)(ids[i] - S.pBegin));
        f.fpx.add(d.x * f.zeroSp);
        f.fpy.add(d.y * f.zeroSp);
        f.fpz.add(d.z * f.zeroSp);
    }
    // SF:626-638
    const float sP = S.mass / (2.0f * rho) * S.C;
    const float sV = S.mu * S.mass * S.C;
    const float sVs = S.mu * S.mass * S.C * 0.0f;
    float Fx = f.fpx.s * sP + f.fvx.s * sV + fsx.s * sP + (-vi.x * sumD2s) * sVs;
    float Fy = f.fpy.s * sP + f.fvy.s * sV + fsy.s * sP + (-vi.y * sumD2s) * sVs;
    float Fz = f.fpz.s * sP + f.fvz.s * sV + fsz.s * sP + (-vi.z * sumD2s) * sVs;
    if (kTri) mesh_force(make_float3(xi.x, xi.y, xi.z), S, tris, Fx, Fy, Fz);
    const float nan = __builtin_nanf("");
    const float ax = poisoned ? nan : Fx / rho + S.gx, ay = poisoned ? nan : Fy / rho + S.gy,
                az = poisoned ? nan : Fz / rho + S.gz;
    if (kDebug) accOut[i] = make_float4(ax, ay, az, 0.f);
    // SF:641-645: symplectic Euler, then the key of the new cell.
    const float h = S.h;
    const float nvx = vi.x + ax * dt, nvy = vi.y + ay * dt, nvz = vi.z + az * dt;
    const float nx = xi.x + nvx * dt, ny = xi.y + nvy * dt, nz = xi.z + nvz * dt;
    const uint32_t ncode = morton3(cell_coord(nx, h), cell_coord(ny, h), cell_coord(nz, h));
    const uint32_t key = fastmod(ncode, S.magicM, S.M);
    posT[i] = make_float4(nx, ny, nz, pi);
    velT[i] = make_float4(nvx, nvy, nvz, __uint_as_float(key));
    const uint32_t gk = S.tBegin + key;
    const Run run = wave_run(gk);
    if (run.is_head) atomicAdd(&counts[gk], run.len);
    // Next step's count of particles with a NaN coordinate in bucket 0. Lane 0
    // is active: the active lanes are a prefix of the wave.
    const uint64_t nb = ballot64(key == 0u && (nx != nx || ny != ny || nz != nz));
    if (nb != 0ull && lane == 0u) atomicAdd(&nanNext[scene], (uint32_t)__popcll(nb));
}

// Force pass over the neighbour lists the density pass recorded (the same
// candidates in the same order as the density pass kept them), so it needs no
// cell groups, tables or masks. A particle whose list overflowed walks its 27
// buckets by Morton runs instead.
template <bool kStatic, bool kDebug, bool kTri, bool kMixed>
__global__ __launch_bounds__(kMaxWG) void k_force_cw(const BlockDesc* __restrict__ blocks, uint32_t numBlocks,
                                                     const SceneDev* __restrict__ scenes,
                                                     const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                     const uint32_t* __restrict__ tab,
                                                     const float4* __restrict__ wallSum,
                                                     const float4* __restrict__ wallZ,
                                                     const float4* __restrict__ visc,
                                                     const uint32_t* __restrict__ nbr,
                                                     const uint32_t* __restrict__ nbrCount,
                                                     uint32_t* __restrict__ nanNext,
                                                     const uint32_t* __restrict__ ids, const float* __restrict__ tris,
                                                     float dt, float4* __restrict__ posT, float4* __restrict__ velT,
                                                     uint32_t* __restrict__ counts, float4* __restrict__ accOut,
                                                     uint32_t numWG, uint32_t run) {
    WaveSlot ws;
    if (!wave_slot(numBlocks, numWG, ws, run)) return;
    const uint32_t blk = ws.tile;
    const BlockDesc bd = blocks[blk];
    const SceneDev& S = scenes[bd.scene];
    const uint32_t lane = __lane_id(), wave = ws.slice;
    const uint32_t i = bd.begin + 64u * wave + lane;
    if (i >= bd.end) return;  // the active lanes of a wave stay a prefix
    ForceSums<kMixed, kViscLists> f;  // the density pass holds the viscosity sums
    f.xi = ld16(pos, i);
    f.vi = ld16(vel, i);
    f.i = i;
    f.h = S.h;
    f.h2 = S.h2;
    const uint32_t n = nbrCount[i];
    if (n <= kListCap) {
        // The list, kBatchForce entries at a time; a wave reads one 256-byte row
        // of indices per slot.
        const uint32_t* list = nbr + (size_t)(blk * kWaves + wave) * (kListCap * 64u) + lane;
        for (uint32_t s = 0; s < n; s += kBatchForce) {
            uint32_t j[kBatchForce];
#pragma unroll
            for (uint32_t u = 0; u < kBatchForce; ++u) j[u] = s + u < n ? list[(s + u) << 6] : i;
            f.batch(pos, vel, j);
        }
    } else if (n != kListPoison) {
        // More neighbours than the list holds: every candidate of the 27 buckets
        // (SF:557-589), each bucket as often as the reference visits it.
        const float h = S.h;
        walk_runs(cell_coord(f.xi.x, h), cell_coord(f.xi.y, h), cell_coord(f.xi.z, h), S.M, S.magicM,
                  tab + S.tBegin, S.pBegin, S.pBegin + S.pCount, [&](uint32_t b, uint32_t e) {
                      for (uint32_t j = b; j < e; ++j) {
                          const uint32_t jj[1] = {j};
                          f.batch(pos, vel, jj);
                      }
                  });
    }
    f.take_visc(visc);
    force_finish<kStatic, kDebug, kTri, kMixed>(f, S, bd.scene, lane, n == kListPoison, wallSum, wallZ, ids, tris, dt,
                                                posT, velT, counts, accOut, nanNext);
}

// Force pass over the density pass's neighbour masks, for scenes too large for
// the lists to stay in the caches: the wave rebuilds its cell groups and their
// bucket tables (no tests), walks the same mask bits as the density pass, and
// takes the crowded buckets' extra particles from the density pass's record.
// Eight waves per SIMD (at most 64 VGPRs) in FP32; the mixed mode's Kahan sums
// need more registers, so that mode settles for six.
template <bool kStatic, bool kDebug, bool kTri, bool kMixed>
__global__ __launch_bounds__(kMaxWG, kMixed ? 6 : 8) void k_force_cw_mask(
    const BlockDesc* __restrict__ blocks, uint32_t numBlocks, const SceneDev* __restrict__ scenes,
    const float4* __restrict__ pos, const float4* __restrict__ vel, const uint32_t* __restrict__ tab,
    const float4* __restrict__ wallSum, const float4* __restrict__ wallZ, const float4* __restrict__ visc,
    const uint32_t* __restrict__ mask,
    const uint64_t* __restrict__ extraMask, const uint32_t* __restrict__ nanCnt, uint32_t* __restrict__ nanNext,
    const uint32_t* __restrict__ ids, const float* __restrict__ tris, float dt, float4* __restrict__ posT,
    float4* __restrict__ velT, uint32_t* __restrict__ counts, float4* __restrict__ accOut, uint32_t numWG,
    uint32_t run) {
    WaveSlot ws;
    if (!wave_slot(numBlocks, numWG, ws, run)) return;
    WaveTables& gt = wave_tables(ws.wave);
    const uint32_t blk = ws.tile;
    const BlockDesc bd = blocks[blk];
    const SceneDev& S = scenes[bd.scene];
    const uint32_t lane = __lane_id(), wave = ws.slice;
    const uint32_t wbegin = bd.begin + 64u * wave;
    if (wbegin >= bd.end) return;
    const uint32_t wcount = min(64u, bd.end - wbegin);
    const uint32_t i = wbegin + lane;
    const bool valid = lane < wcount;
    const uint32_t iv = valid ? i : wbegin;
    ForceSums<kMixed, false> f;
    f.xi = ld16(pos, iv);
    f.vi = ld16(vel, iv);
    f.i = i;
    f.h = S.h;
    f.h2 = S.h2;
    const uint32_t code = morton3(cell_coord(f.xi.x, f.h), cell_coord(f.xi.y, f.h), cell_coord(f.xi.z, f.h));
    uint64_t m[4];
    {
        const uint4* mr = reinterpret_cast<const uint4*>(mask + (size_t)iv * kCwMaskWords);
        const uint4 a = mr[0], b = mr[1];
        m[0] = ((uint64_t)a.y << 32) | a.x;
        m[1] = ((uint64_t)a.w << 32) | a.z;
        m[2] = ((uint64_t)b.y << 32) | b.x;
        m[3] = ((uint64_t)b.w << 32) | b.z;
    }
    auto pairs = [&](const auto& j) { f.batch(pos, vel, j); };
    uint32_t noWalls = 0u;  // walls come from the density pass's sums; kPoisonBit
    const uint64_t nanLanes = ballot64(valid && (f.xi.x != f.xi.x || f.xi.y != f.xi.y || f.xi.z != f.xi.z));
    for_each_group<false, false>(
        lane, valid, code, lane_offsets(lane), S, tab + S.tBegin, nullptr, nanCnt[bd.scene], nanLanes,
        S.pBegin, S.pBegin + S.pCount, gt, noWalls,
        [&](uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, auto&& between) { between(); },
        [&](uint32_t slot, uint32_t crowded) {
            walk_mask<kBatchForceMask, true>(m, gt.start[slot], i, pairs);
            // Crowded buckets' extra particles: the density pass's neighbours
            // among them when there are at most kExtraBits, else every one.
            if (!crowded) return;
            uint64_t xb = extraMask[i];
            const bool all = xb == ~0ull;
            for (; crowded; crowded &= crowded - 1u) {
                const uint32_t bi = (uint32_t)__builtin_ctz(crowded);
                const uint32_t s0 = gt.start[slot][bi] + kSlots, ne = gt.cnt[slot][bi] - kSlots;
                for (uint32_t base = 0; base < ne; base += 32u) {
                    const uint32_t c = min(32u, ne - base);
                    const uint32_t low = c < 32u ? (1u << c) - 1u : ~0u;
                    uint32_t w = all ? low : (uint32_t)xb & low;
                    xb >>= c;
                    for (; w; w &= w - 1u) {  // one at a time: fewer registers
                        const uint32_t j[1] = {s0 + base + (uint32_t)__builtin_ctz(w)};
                        f.batch(pos, vel, j);
                    }
                }
            }
        });
    if (!valid) return;
    f.take_visc(visc);
    force_finish<kStatic, kDebug, kTri, kMixed>(f, S, bd.scene, lane, (noWalls & kPoisonBit) != 0u, wallSum, wallZ,
                                                ids, tris, dt, posT, velT, counts, accOut, nanNext);
}

}  // namespace sph

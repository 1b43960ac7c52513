// Scalar arithmetic shared by the host-side initial sort and the gfx950 kernels:
// cell coordinates, Morton bucket keys, the bucket modulus and the smoothing
// kernels of the reference (SPHFluidSimulation/source/SPH/Kernels/SPHFunctions.cpp,
// "SF" below). Compiled by hipcc for both sides.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sph {

// SF:65-76, SF:99-102. floor(x / h) with an IEEE (correctly rounded) division,
// converted through int32 so a negative cell wraps to 2^32 - k, the OpenCL
// build's as_uint(convert_int_rtn(x / h)).
__host__ __device__ inline uint32_t cell_coord(float x, float h) {
    return (uint32_t)(int32_t)floorf(x / h);
}

// x minus the corner cell * h of a cell, the product rounded to FP32 first and
// never fused with the subtraction, so that every caller forms the same corner
// (the candidate filters rely on targets and candidates sharing it).
__host__ __device__ inline float corner_offset(float x, uint32_t cell, float h) {
#pragma clang fp contract(off)
    const float c = (float)(int32_t)cell * h;
    return x - c;
}

// Hash of a cell's Morton code for the wall-cell bitmap (xor-shift-multiply
// rounds). Cells whose codes alias (1024 cells apart on an axis) share a bit,
// which only makes the bitmap's filter looser.
__host__ __device__ inline uint32_t wall_hash(uint32_t code) {
    uint32_t h = code ^ (code >> 15);
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return h;
}

// SF:79-86 (ExpandBits): the low 10 bits of v, bit k moved to bit 3k.
__host__ __device__ inline uint32_t spread3(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// SF:88-97, SF:103-106: 30-bit Morton code, x on bit 3k, y on 3k+1, z on 3k+2.
__host__ __device__ inline uint32_t morton3(uint32_t cx, uint32_t cy, uint32_t cz) {
    return spread3(cx) | (spread3(cy) << 1) | (spread3(cz) << 2);
}

// Neighbour n = 9a + 3b + c (a, b, c in 0..2) of a cell is the cell offset by
// (a - 1, b - 1, c - 1); its dilated offsets (a, b, c spread as morton3 does).
__host__ __device__ inline uint32_t neighbour_offsets(uint32_t n) {
    return morton3(n / 9u, (n / 3u) % 3u, n % 3u);
}

// Morton code of the neighbour with dilated offsets `off` of the cell whose
// code is c, equal to morton3 of the offset coordinates (each wrapping at 1024
// as morton3 does). Per axis in dilated arithmetic: c's coordinate minus one
// (the borrow runs through the zero gaps), the gaps filled with ones so the
// carry of adding the offset runs through them, then the axis's bits.
__host__ __device__ inline uint32_t neighbour_code(uint32_t c, uint32_t off) {
    constexpr uint32_t mx = 0x09249249u, my = mx << 1, mz = mx << 2;
    const uint32_t bx = (((c & mx) - 1u) & mx) | ~mx;
    const uint32_t by = (((c & my) - 2u) & my) | ~my;
    const uint32_t bz = (((c & mz) - 4u) & mz) | ~mz;
    return ((bx + (off & mx)) & mx) | ((by + (off & my)) & my) | ((bz + (off & mz)) & mz);
}

// Remainder by a runtime divisor d without a hardware divide: Lemire, Kaser and
// Kurz, "Faster Remainder by Direct Computation" (2019), exact for every 32-bit
// numerator and divisor. magic = floor((2^64 - 1) / d) + 1 (0 for d = 1).
__host__ __device__ inline uint64_t fastmod_magic(uint32_t d) { return UINT64_MAX / d + 1u; }

// The high 64 bits of low * d are assembled from two 32x32->64 products, so the
// same code runs on the host (the initial sort and the tests) and the device.
__host__ __device__ inline uint32_t fastmod(uint32_t a, uint64_t magic, uint32_t d) {
    const uint64_t low = magic * a;
    const uint64_t t = (uint64_t)(uint32_t)low * d;
    const uint64_t u = (low >> 32) * d + (t >> 32);
    return (uint32_t)(u >> 32);
}

// SF:115-118: the reference's constant uses 3.1415f, not pi, and multiplies h in
// this order; kept bit-for-bit.
__host__ __device__ inline float kernel_constant(float h) {
    return 15.0f / (3.1415f * h * h * h * h * h * h);
}

// The reference drops a candidate only when distSqr > h^2 (SF:447, SF:488,
// SF:571, SF:605), so a NaN distance is kept, and its kernels return 0 only
// when r >= h (SF:119-139), so a NaN r gives NaN weights: a particle with a NaN
// coordinate poisons every particle that visits its bucket. These two tests
// keep that; for numbers they are r2 <= h2 and r < h.
__host__ __device__ inline bool within_h2(float r2, float h2) { return !(r2 > h2); }
__host__ __device__ inline bool inside_h(float r, float h) { return !(r >= h); }

// SF:119-125 at r = 0, for selfDensity = m * D0(0, h) * C (SimulationEngineCPU.cpp:407).
__host__ __device__ inline float d0_at_zero(float h) { return (h - 0.0f) * (h - 0.0f) * (h - 0.0f); }

// ---------------------------------------------------------------------------
// Neighbour walk by Morton runs.
//
// The reference visits the 27 cells around a particle one by one (SF:397-404)
// and scans each cell's bucket. Per axis the three cells c-1, c, c+1 split into
// an even-aligned pair {c & ~1, (c & ~1) + 1} and a single cell, and the Morton
// codes of an aligned pair differ only in that axis's lowest bit. So the 27
// cells form 14 runs of consecutive codes: pair x pair x pair (8 codes),
// pair x pair x single (4), pair x single x pair (two runs of 2: z adds 4),
// single x pair x pair (four runs of 1: y adds 2, z adds 4), pair x single x
// single (2), single x pair x single (two of 1), single x single x pair (two of
// 1), single x single x single (1). A run of L consecutive codes starting at an
// L-aligned code maps to L consecutive buckets b0 .. b0+L-1 unless it wraps
// past M, and consecutive buckets are one contiguous particle range
// [table[b0], table[b0+L]). Wrapping runs are walked bucket by bucket, so every
// bucket is visited exactly as often as the reference visits it (colliding
// cells included).
struct CellPair {
    uint32_t p, s;  // spread3 of the aligned pair base and of the single cell, pre-shifted
};

__host__ __device__ inline CellPair axis_pair(uint32_t c, uint32_t shift) {
    CellPair r;
    r.p = spread3(c & ~1u) << shift;
    r.s = spread3((c & 1u) ? c + 1u : c - 1u) << shift;
    return r;
}

// Run r (0..13): first Morton code and length in codes.
__host__ __device__ inline void run_of(uint32_t r, const CellPair& X, const CellPair& Y, const CellPair& Z,
                                       uint32_t& code, uint32_t& len) {
    switch (r) {
        case 0: code = X.p | Y.p | Z.p; len = 8; break;
        case 1: code = X.p | Y.p | Z.s; len = 4; break;
        case 2: code = X.p | Y.s | Z.p; len = 2; break;
        case 3: code = (X.p | Y.s | Z.p) + 4u; len = 2; break;
        case 4: code = X.s | Y.p | Z.p; len = 1; break;
        case 5: code = (X.s | Y.p | Z.p) + 2u; len = 1; break;
        case 6: code = (X.s | Y.p | Z.p) + 4u; len = 1; break;
        case 7: code = (X.s | Y.p | Z.p) + 6u; len = 1; break;
        case 8: code = X.p | Y.s | Z.s; len = 2; break;
        case 9: code = X.s | Y.p | Z.s; len = 1; break;
        case 10: code = (X.s | Y.p | Z.s) + 2u; len = 1; break;
        case 11: code = X.s | Y.s | Z.p; len = 1; break;
        case 12: code = (X.s | Y.s | Z.p) + 4u; len = 1; break;
        default: code = X.s | Y.s | Z.s; len = 1; break;
    }
}

// Calls f(begin, end) for every particle range the 27-cell walk visits, in a
// fixed order. `table` is the scene's table segment (M + 1 global indices);
// ranges are clamped to [lo, hi) so a corrupted table cannot read outside the
// particle arrays.
template <class F>
__host__ __device__ inline void walk_runs(uint32_t cx, uint32_t cy, uint32_t cz, uint32_t M, uint64_t magic,
                                          const uint32_t* __restrict__ table, uint32_t lo, uint32_t hi, F&& f) {
    const CellPair X = axis_pair(cx, 0), Y = axis_pair(cy, 1), Z = axis_pair(cz, 2);
    for (uint32_t r = 0; r < 14; ++r) {
        uint32_t code, len;
        run_of(r, X, Y, Z, code, len);
        const uint32_t b0 = fastmod(code, magic, M);
        const bool wraps = b0 + len > M;
        const uint32_t parts = wraps ? len : 1u;
        for (uint32_t part = 0; part < parts; ++part) {
            const uint32_t b = wraps ? fastmod(b0 + part, magic, M) : b0;
            const uint32_t bl = wraps ? 1u : len;
            const uint32_t tb = table[b], te = table[b + bl];
            f(tb > lo ? tb : lo, te < hi ? te : hi);
        }
    }
}

}  // namespace sph

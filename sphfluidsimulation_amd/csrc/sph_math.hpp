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

// Remainder by a runtime divisor d without a hardware divide: Lemire, Kaser and
// Kurz, "Faster Remainder by Direct Computation" (2019), exact for every 32-bit
// numerator and divisor. magic = floor((2^64 - 1) / d) + 1 (0 for d = 1).
__host__ __device__ inline uint64_t fastmod_magic(uint32_t d) { return UINT64_MAX / d + 1u; }

__device__ inline uint32_t fastmod(uint32_t a, uint64_t magic, uint32_t d) {
    const uint64_t low = magic * a;
    return (uint32_t)__umul64hi(low, (uint64_t)d);
}

// SF:115-118: the reference's constant uses 3.1415f, not pi, and multiplies h in
// this order; kept bit-for-bit.
__host__ __device__ inline float kernel_constant(float h) {
    return 15.0f / (3.1415f * h * h * h * h * h * h);
}

// SF:119-125 at r = 0, for selfDensity = m * D0(0, h) * C (SimulationEngineCPU.cpp:407).
__host__ __device__ inline float d0_at_zero(float h) { return (h - 0.0f) * (h - 0.0f) * (h - 0.0f); }

}  // namespace sph

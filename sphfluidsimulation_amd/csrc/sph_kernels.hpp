// gfx950 kernels of the SPH step. One step is four launches:
//   k_density  : 27-cell density sum -> pressure           (SF:358-501)
//   k_force    : 27-cell forces, symplectic Euler, new key, (SF:503-656,
//                wave-aggregated bucket histogram            SF:321-331)
//   k_scan     : single-pass inclusive scan of the counts with decoupled
//                lookback; clears the other table for the next histogram
//                (SimulationEngineCPU.cpp:245-253)
//   k_scatter  : wave-aggregated atomic-decrement scatter that moves the
//                particle state itself into bucket order     (SF:332-356)
// "SF" is SPHFluidSimulation/source/SPH/Kernels/SPHFunctions.cpp.
//
// State lives in HBM as structure-of-arrays in bucket order: pos (x, y, z,
// pressure of the last step), vel (vx, vy, vz, key bits), pres (this step's
// pressure), ids (reference slot of each particle). All scenes of a batch share
// the arrays; a scene owns a contiguous particle range and a contiguous table
// segment of M + 1 entries, and table values are global particle indices, so a
// single global scan is the segmented scan.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_math.hpp"

namespace sph {

constexpr int kBlock = 256;  // particles per workgroup in the particle kernels

struct alignas(16) SceneDev {
    float h, h2, mass, k;
    float rho0, mu, gx, gy;
    float gz, selfDensity, C, h2wide;  // h2wide = h2 * 1.0001f, the candidate filters' radius
    uint32_t pBegin, pCount, tBegin, M;  // particles [pBegin, pBegin + pCount), table [tBegin, tBegin + M]
    uint64_t magicM;
    uint32_t sBegin, sCount;              // static particles [sBegin, sBegin + sCount)
    uint32_t stBegin, Ms;                 // static table [stBegin, stBegin + Ms]
    uint64_t magicMs;
    uint32_t triBegin, triCount;
    uint32_t wmBegin, wmMask;  // wall-cell bitmap: words [wmBegin, wmBegin + wmMask]
    // Boundary-mesh bins (mesh_force): a grid of nbx x nby x nbz bins of edge
    // 1 / binInv from (bx0, by0, bz0); the bin-start table at float offset binStart
    // of the mesh buffer, the triangle lists at float offset binList (uint bits).
    float bx0, by0, bz0, binInv;
    uint32_t nbx, nby, nbz, binStart;
    uint32_t binList;
    // Wall lists (sph_wall_lists.hpp): the scene's part of the cell array starts
    // at word wlBegin, with wlMask + 1 table slots; wlMask == 0 means no table.
    uint32_t wlBegin, wlMask, pad0;
};

struct alignas(16) BlockDesc {
    uint32_t scene, begin, end, pad;
};

typedef _Float16 half2v __attribute__((ext_vector_type(2)));
__device__ inline half2v pack_h2(float lo, float hi) { return half2v{(_Float16)lo, (_Float16)hi}; }
__device__ inline uint32_t h2_bits(half2v v) { return __builtin_bit_cast(uint32_t, v); }
__device__ inline half2v h2_from(uint32_t b) { return __builtin_bit_cast(half2v, b); }

// Mixed precision (SPH_FLAG_FP16_OFFSETS): each particle's position stored as FP16
// offsets from the corner of its own cell, in units of h, with the low bits of
// its cell as a tag, 8 bytes in bucket order beside pos. x: (u_x, u_y); y: u_z in
// the low half and the tag in the high half. The tag holds the low 4 bits of
// each cell coordinate in 5-bit fields whose top bits (kTagGuards) are set, so
// (tag_a - tag_b) & kTagFields is the cell difference modulo 16 per axis with no
// borrow between fields. The offsets use the corner and 1/h of the density
// pass's own FP16 filter.
constexpr uint32_t kTagGuards = 0x4210u;  // bits 4, 9, 14
constexpr uint32_t kTagFields = 0x3DEFu;  // bits 0-3, 5-8, 10-13
__device__ inline uint32_t cell_tag(uint32_t cx, uint32_t cy, uint32_t cz) {
    return (cx & 15u) | ((cy & 15u) << 5) | ((cz & 15u) << 10);
}
__device__ inline uint2 offset_record(float4 p, float h) {
    const uint32_t cx = cell_coord(p.x, h), cy = cell_coord(p.y, h), cz = cell_coord(p.z, h);
    const float invh = 1.0f / h;
    const float ux = corner_offset(p.x, cx, h) * invh;
    const float uy = corner_offset(p.y, cy, h) * invh;
    const float uz = corner_offset(p.z, cz, h) * invh;
    const uint32_t zbits = __builtin_bit_cast(uint16_t, (_Float16)uz);
    return make_uint2(h2_bits(pack_h2(ux, uy)), zbits | ((cell_tag(cx, cy, cz) | kTagGuards) << 16));
}

// ---------------------------------------------------------------------------
// Wave-level helpers (wave64).

// Runs of equal keys in consecutive active lanes. Active lanes form a prefix of
// the wave (the tail of a block's last wave has returned early).
struct Run {
    uint32_t head;  // first lane of this lane's run
    uint32_t len;   // lanes in the run
    uint32_t rank;  // lane - head
    bool is_head;
};

__device__ inline Run wave_run(uint32_t key) {
    const uint32_t lane = __lane_id();
    const uint64_t active = __ballot(1);
    const uint32_t prev = __shfl_up(key, 1);
    const bool head = (lane == 0) || (prev != key);
    const uint64_t heads = __ballot(head);
    const uint64_t upto = (2ull << lane) - 1ull;  // lanes 0..lane
    const uint32_t h = 63u - (uint32_t)__clzll(heads & upto);
    const uint64_t above = heads & ~upto;
    const uint32_t active_end = 64u - (uint32_t)__clzll(active);
    const uint32_t next = above ? (uint32_t)(__ffsll((long long)above) - 1) : active_end;
    Run r;
    r.head = h;
    r.len = next - h;
    r.rank = lane - h;
    r.is_head = head;
    return r;
}

// ---------------------------------------------------------------------------
// Per-particle pieces of SF.

struct CellSpread {
    uint32_t x[3], y[3], z[3];  // spread3 of cell - 1, cell, cell + 1 per axis, pre-shifted for y and z
};

// SF:388-399: neighbour cells cell-1 .. cell+1 per axis with uint wrap; the
// Morton code of a neighbour is the OR of its three spread coordinates.
__device__ inline CellSpread cell_spread(float x, float y, float z, float h) {
    const uint32_t cx = cell_coord(x, h), cy = cell_coord(y, h), cz = cell_coord(z, h);
    CellSpread c;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        c.x[a] = spread3(cx + (uint32_t)a - 1u);
        c.y[a] = spread3(cy + (uint32_t)a - 1u) << 1;
        c.z[a] = spread3(cz + (uint32_t)a - 1u) << 2;
    }
    return c;
}

// SF:140-155 with the C++ modf (sign kept); the reference calls Noise twice
// with the same argument.
__device__ inline float noise(float x) {
    float ip;
    return modff(sinf(x * 112.9898f) * 43758.5453f, &ip);
}
__device__ inline float3 random_direction(float x) {
    const float theta = noise(x) * 2.0f * 3.1415f;
    const float z = noise(x) * 2.0f - 1.0f;
    const float s = sinf(theta), c = cosf(theta);
    const float z2 = sqrtf(1.0f - z * z);
    return make_float3(z2 * c, z2 * s, z);
}

__device__ inline float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }

// SF:156-201
__device__ inline float3 closest_point_on_triangle(float3 p, float3 a, float3 b, float3 c) {
    const float3 ab = sub3(b, a), ac = sub3(c, a), ap = sub3(p, a);
    const float d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    if (d1 <= 0.f && d2 <= 0.f) return a;
    const float3 bp = sub3(p, b);
    const float d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0.f && d4 <= d3) return b;
    const float3 cp = sub3(p, c);
    const float d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 >= 0.f && d5 <= d6) return c;
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
        const float v = d1 / (d1 - d3);
        return make_float3(a.x + ab.x * v, a.y + ab.y * v, a.z + ab.z * v);
    }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {
        const float v = d2 / (d2 - d6);
        return make_float3(a.x + ac.x * v, a.y + ac.y * v, a.z + ac.z * v);
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
        const float v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return make_float3(b.x + (c.x - b.x) * v, b.y + (c.y - b.y) * v, b.z + (c.z - b.z) * v);
    }
    const float denom = 1.f / (va + vb + vc);
    const float v = vb * denom, w = vc * denom;
    return make_float3(a.x + ab.x * v + ac.x * w, a.y + ab.y * v + ac.y * w, a.z + ab.z * v + ac.z * w);
}

// SF:202-223
__device__ inline float3 triangle_force(float3 p, const float* t) {
    const float3 a = make_float3(t[0], t[1], t[2]), b = make_float3(t[3], t[4], t[5]),
                 c = make_float3(t[6], t[7], t[8]);
    const float3 d = sub3(closest_point_on_triangle(p, a, b, c), p);
    const float dist = sqrtf(dot3(d, d));
    if (dist > 2.5f) return make_float3(0.f, 0.f, 0.f);
    const float tt = 1.0f / (dist / 2.5f) - 1.0f;
    const float amp = tt * tt * 50.0f;
    return make_float3(-d.x / dist * amp, -d.y / dist * amp, -d.z / dist * amp);
}

// The boundary-mesh force of SF:636-637, the scene's triangles summed in index
// order onto (fx, fy, fz), over the triangles of the particle's bin only. A bin
// lists, in index order, every triangle whose bounding box widened by the force's
// range (2.5, plus a rounding margin) meets the bin, so each triangle left out is
// farther than the range from the particle and adds an exact zero in the
// reference (SF:211-213); a point outside the grid has no triangle in range. A
// particle with a NaN coordinate takes every triangle, as the reference does.
// `mesh` holds the triangles (9 floats each) followed by the bin tables.
__device__ inline void mesh_force(float3 p, const SceneDev& S, const float* __restrict__ mesh, float& fx, float& fy,
                                  float& fz) {
    const float* tri = mesh + 9u * S.triBegin;
    auto add = [&](uint32_t t) {
        const float3 f = triangle_force(p, tri + 9u * t);
        fx += f.x;
        fy += f.y;
        fz += f.z;
    };
    if (p.x != p.x || p.y != p.y || p.z != p.z) {
        for (uint32_t t = 0; t < S.triCount; ++t) add(t);
        return;
    }
    const float gx = floorf((p.x - S.bx0) * S.binInv), gy = floorf((p.y - S.by0) * S.binInv),
                gz = floorf((p.z - S.bz0) * S.binInv);
    if (!(gx >= 0.f && gy >= 0.f && gz >= 0.f && gx < (float)S.nbx && gy < (float)S.nby && gz < (float)S.nbz)) return;
    const uint32_t b = ((uint32_t)gz * S.nby + (uint32_t)gy) * S.nbx + (uint32_t)gx;
    const uint32_t* start = reinterpret_cast<const uint32_t*>(mesh + S.binStart);
    const uint32_t* list = reinterpret_cast<const uint32_t*>(mesh + S.binList);
    for (uint32_t k = start[b], e = start[b + 1]; k < e; ++k) add(list[k]);
}

constexpr uint32_t kMaskBits = 256;  // candidates per particle recorded by the density pass
constexpr uint32_t kMaskWords = kMaskBits / 32;

// ---------------------------------------------------------------------------
// k_density: SF:358-501 for one particle per lane.
// kMap (particle-map mode): the state stays in memory order and bucket slot k of
// the table holds particle pmap[k] (the reference's particleMap, SF:346-356);
// lane k takes particle pmap[k], so a wave still walks neighbouring cells, and
// each candidate slot j of a bucket is particle pmap[j]. Results go to the
// particle's memory index. Without kMap, slot k is particle k.
template <bool kStatic, bool kDebug, bool kMap>
__global__ __launch_bounds__(kBlock) void k_density(const BlockDesc* __restrict__ blocks,
                                                    const SceneDev* __restrict__ scenes,
                                                    const float4* __restrict__ pos, const uint32_t* __restrict__ tab,
                                                    const float4* __restrict__ spos, const uint32_t* __restrict__ stab,
                                                    const uint32_t* __restrict__ pmap, float* __restrict__ pres,
                                                    float* __restrict__ dens) {
    const BlockDesc bd = blocks[blockIdx.x];
    const SceneDev& S = scenes[bd.scene];
    const uint32_t slot = bd.begin + threadIdx.x;
    if (slot >= bd.end) return;
    // The map is a permutation of the scene's particles; a corrupted entry is
    // clamped so that no access leaves the arrays.
    uint32_t i = kMap ? pmap[slot] : slot;
    if (kMap && (i < S.pBegin || i >= S.pBegin + S.pCount)) i = S.pBegin;
    const float4 xi = pos[i];
    const float h = S.h, h2 = S.h2;
    const uint32_t pEnd = S.pBegin + S.pCount;
    const CellSpread cs = cell_spread(xi.x, xi.y, xi.z, h);
    float sumDyn = 0.f, sumSt = 0.f;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c) {
                const uint32_t code = cs.x[a] | cs.y[b] | cs.z[c];
                const uint32_t bk = S.tBegin + fastmod(code, S.magicM, S.M);
                // Table entries always lie in the scene's range; the clamp keeps
                // a corrupted table from reading outside the particle arrays.
                const uint32_t s = max(tab[bk], S.pBegin), e = min(tab[bk + 1], pEnd);
                for (uint32_t js = s; js < e; ++js) {
                    const uint32_t j = kMap ? pmap[js] : js;
                    if (j == i) continue;
                    const float4 xj = pos[j];
                    const float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
                    const float r2 = dx * dx + dy * dy + dz * dz;
                    if (r2 > h2) continue;
                    const float r = __builtin_amdgcn_sqrtf(r2);
                    const float q = h - r;
                    sumDyn += inside_h(r, h) ? q * q * q : 0.f;
                }
                if (kStatic && S.sCount) {
                    const uint32_t sb = S.stBegin + fastmod(code, S.magicMs, S.Ms);
                    const uint32_t ss = stab[sb], se = stab[sb + 1];
                    for (uint32_t j = ss; j < se; ++j) {
                        const float4 xj = spos[j];
                        const float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
                        const float r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 > h2) continue;
                        const float r = __builtin_amdgcn_sqrtf(r2);
                        const float q = h - r;
                        sumSt += inside_h(r, h) ? q * q * q : 0.f;
                    }
                }
            }
    const float rho = S.selfDensity + (sumDyn * S.mass + sumSt * S.mass) * S.C;
    pres[i] = S.k * (rho - S.rho0);
    if (kDebug) dens[i] = rho;
}

// ---------------------------------------------------------------------------
// k_force: SF:503-656 for one particle per lane, then the new bucket key and
// the wave-aggregated histogram (SF:321-331 fused into the integrator's
// epilogue, as the reference's SF:644-645 already computes the key there).
// kMap: as k_density's.
template <bool kStatic, bool kDebug, bool kTri, bool kMap>
__global__ __launch_bounds__(kBlock) void k_force(const BlockDesc* __restrict__ blocks,
                                                  const SceneDev* __restrict__ scenes,
                                                  const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                  const float* __restrict__ pres, const uint32_t* __restrict__ tab,
                                                  const float4* __restrict__ spos, const uint32_t* __restrict__ stab,
                                                  const uint32_t* __restrict__ pmap, const uint32_t* __restrict__ ids,
                                                  const float* __restrict__ tris, float dt,
                                                  float4* __restrict__ posT, float4* __restrict__ velT,
                                                  uint32_t* __restrict__ counts, float4* __restrict__ accOut) {
    const BlockDesc bd = blocks[blockIdx.x];
    const SceneDev& S = scenes[bd.scene];
    const uint32_t slot = bd.begin + threadIdx.x;
    if (slot >= bd.end) return;
    // A corrupted map entry is clamped, as in k_density; every lane of the wave's
    // prefix stays active for wave_run below.
    uint32_t i = kMap ? pmap[slot] : slot;
    if (kMap && (i < S.pBegin || i >= S.pBegin + S.pCount)) i = S.pBegin;
    const float4 xi = pos[i];
    const float4 vi = vel[i];
    const float pi = pres[i];
    const float h = S.h, h2 = S.h2;
    const uint32_t pEnd = S.pBegin + S.pCount;
    const float rho = pi / S.k + S.rho0;  // SF:544
    const float absP = fabsf(pi);
    const CellSpread cs = cell_spread(xi.x, xi.y, xi.z, h);
    float fpx = 0.f, fpy = 0.f, fpz = 0.f;  // dynamic pressure
    float fvx = 0.f, fvy = 0.f, fvz = 0.f;  // dynamic viscosity
    float fsx = 0.f, fsy = 0.f, fsz = 0.f;  // static pressure
    float sumD2s = 0.f;                      // static viscosity weight (scaled by 0, SF:629)
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c) {
                const uint32_t code = cs.x[a] | cs.y[b] | cs.z[c];
                const uint32_t bk = S.tBegin + fastmod(code, S.magicM, S.M);
                // Table entries always lie in the scene's range; the clamp keeps
                // a corrupted table from reading outside the particle arrays.
                const uint32_t s = max(tab[bk], S.pBegin), e = min(tab[bk + 1], pEnd);
                for (uint32_t js = s; js < e; ++js) {
                    const uint32_t j = kMap ? pmap[js] : js;
                    if (j == i) continue;
                    const float4 xj = pos[j];
                    float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
                    const float r2 = dx * dx + dy * dy + dz * dz;
                    if (r2 > h2) continue;
                    const float r = __builtin_amdgcn_sqrtf(r2);
                    if (r2 == 0.f || r == 0.f) {
                        const float3 d = random_direction((float)(ids[i] - S.pBegin));
                        dx = d.x;
                        dy = d.y;
                        dz = d.z;
                    } else {
                        const float inv = __builtin_amdgcn_rcpf(r);
                        dx *= inv;
                        dy *= inv;
                        dz *= inv;
                    }
                    const float q = h - r;
                    const bool in = inside_h(r, h);
                    const float w1 = in ? -3.0f * q * q : 0.f;
                    const float w2 = in ? 6.0f * q : 0.f;
                    const float pj = pres[j];
                    const float4 vj = vel[j];
                    const float sp = (pi + pj) * w1;
                    fpx += dx * sp;
                    fpy += dy * sp;
                    fpz += dz * sp;
                    fvx += (vj.x - vi.x) * w2;
                    fvy += (vj.y - vi.y) * w2;
                    fvz += (vj.z - vi.z) * w2;
                }
                if (kStatic && S.sCount) {
                    const uint32_t sb = S.stBegin + fastmod(code, S.magicMs, S.Ms);
                    const uint32_t ss = stab[sb], se = stab[sb + 1];
                    for (uint32_t j = ss; j < se; ++j) {
                        const float4 xj = spos[j];
                        float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
                        const float r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 > h2) continue;
                        const float r = __builtin_amdgcn_sqrtf(r2);
                        if (r2 == 0.f || r == 0.f) {
                            const float3 d = random_direction((float)(ids[i] - S.pBegin));
                            dx = d.x;
                            dy = d.y;
                            dz = d.z;
                        } else {
                            const float inv = __builtin_amdgcn_rcpf(r);
                            dx *= inv;
                            dy *= inv;
                            dz *= inv;
                        }
                        const float q = h - r;
                        const bool in = inside_h(r, h);
                        const float sp = absP * (in ? -3.0f * q * q : 0.f);
                        fsx += dx * sp;
                        fsy += dy * sp;
                        fsz += dz * sp;
                        sumD2s += in ? 6.0f * q : 0.f;
                    }
                }
            }
    // SF:626-638
    const float sP = S.mass / (2.0f * rho) * S.C;
    const float sV = S.mu * S.mass * S.C;
    const float sVs = S.mu * S.mass * S.C * 0.0f;
    float Fx = fpx * sP + fvx * sV + fsx * sP + (-vi.x * sumD2s) * sVs;
    float Fy = fpy * sP + fvy * sV + fsy * sP + (-vi.y * sumD2s) * sVs;
    float Fz = fpz * sP + fvz * sV + fsz * sP + (-vi.z * sumD2s) * sVs;
    if (kTri) mesh_force(make_float3(xi.x, xi.y, xi.z), S, tris, Fx, Fy, Fz);
    const float ax = Fx / rho + S.gx, ay = Fy / rho + S.gy, az = Fz / rho + S.gz;
    if (kDebug) accOut[i] = make_float4(ax, ay, az, 0.f);
    // SF:641-645: symplectic Euler, then the key of the new cell.
    const float nvx = vi.x + ax * dt, nvy = vi.y + ay * dt, nvz = vi.z + az * dt;
    const float nx = xi.x + nvx * dt, ny = xi.y + nvy * dt, nz = xi.z + nvz * dt;
    const uint32_t code = morton3(cell_coord(nx, h), cell_coord(ny, h), cell_coord(nz, h));
    const uint32_t key = fastmod(code, S.magicM, S.M);
    posT[i] = make_float4(nx, ny, nz, pi);
    velT[i] = make_float4(nvx, nvy, nvz, __uint_as_float(key));
    const uint32_t gk = S.tBegin + key;
    const Run run = wave_run(gk);
    if (run.is_head) atomicAdd(&counts[gk], run.len);
}

// ---------------------------------------------------------------------------
// Run-walk variants of the two neighbour passes. The density pass also records,
// per particle, which of its first 256 candidates (in walk order) lie within h
// (a slightly widened test, so the set is a superset of the force pass's exact
// test). The force pass walks the same runs of the same table and only visits
// the recorded candidates, plus every candidate past the 256th. The density
// pass stores the pressure in pos.w, so the force pass gets a neighbour's
// position and pressure from one 16-byte load.

// Workgroups are dispatched round-robin over the 8 XCDs (launch index b runs
// on XCD b % 8). The launch is padded to a multiple of 8 and index b takes
// particle tile (b % 8) * chunk + b / 8, so each XCD walks one contiguous
// eighth of the sorted particles and neighbouring tiles share that XCD's L2.
__device__ inline uint32_t xcd_tile(uint32_t b, uint32_t numTiles) {
    const uint32_t chunk = (numTiles + 7u) >> 3;
    return (b & 7u) * chunk + (b >> 3);
}

// The same with the XCDs interleaved in runs of `run` tiles: XCD x takes runs x,
// x + 8, x + 16, ... (the launch padded to a multiple of 8 run), so each XCD's
// share spreads over the whole particle range, which evens out the XCDs' work
// when it varies along the range, while each run of consecutive tiles still
// shares one XCD's L2.
__device__ inline uint32_t xcd_tile_runs(uint32_t b, uint32_t run) {
    const uint32_t x = b & 7u, i = b >> 3;
    return ((i / run) * 8u + x) * run + (i % run);
}
// Runs per XCD of an interleaved launch: eight (measured best against fixed runs
// of 32 to 256 workgroups at 1M and 8M).
constexpr uint32_t kXcdRunsPerXcd = 8;

template <bool kStatic, bool kDebug>
__global__ __launch_bounds__(kBlock) void k_density_runs(const BlockDesc* __restrict__ blocks, uint32_t numBlocks,
                                                         const SceneDev* __restrict__ scenes, float4* pos,
                                                         const uint32_t* __restrict__ tab,
                                                         const float4* __restrict__ spos,
                                                         const uint32_t* __restrict__ stab,
                                                         uint32_t* __restrict__ mask, float* __restrict__ dens) {
    __shared__ uint32_t smask[kMaskWords * kBlock];
    const uint32_t tile = xcd_tile(blockIdx.x, numBlocks);
    if (tile >= numBlocks) return;
    const BlockDesc bd = blocks[tile];
    const SceneDev& S = scenes[bd.scene];
    const uint32_t tid = threadIdx.x;
    const uint32_t i = bd.begin + tid;
    if (i >= bd.end) return;
#pragma unroll
    for (uint32_t w = 0; w < kMaskWords; ++w) smask[w * kBlock + tid] = 0u;
    const float4 xi = pos[i];
    const float h = S.h, h2 = S.h2;
    const float h2wide = S.h2wide;
    const uint32_t cx = cell_coord(xi.x, h), cy = cell_coord(xi.y, h), cz = cell_coord(xi.z, h);
    float sumDyn = 0.f, sumSt = 0.f;
    uint32_t k = 0, bits = 0;
    walk_runs(cx, cy, cz, S.M, S.magicM, tab + S.tBegin, S.pBegin, S.pBegin + S.pCount,
              [&](uint32_t s, uint32_t e) {
                  // Four candidates per iteration, loads issued before use.
                  for (uint32_t j = s; j < e; j += 4u) {
                      const uint32_t last = e - 1u;
                      float4 xs[4];
#pragma unroll
                      for (uint32_t q = 0; q < 4; ++q) xs[q] = pos[min(j + q, last)];
#pragma unroll
                      for (uint32_t q = 0; q < 4; ++q) {
                          if (j + q >= e) break;
                          const float dx = xs[q].x - xi.x, dy = xs[q].y - xi.y, dz = xs[q].z - xi.z;
                          const float r2 = dx * dx + dy * dy + dz * dz;
                          const bool self = j + q == i;
                          if (within_h2(r2, h2wide) && !self) bits |= 1u << (k & 31u);
                          if ((k & 31u) == 31u) {
                              if (k < kMaskBits) smask[(k >> 5) * kBlock + tid] = bits;
                              bits = 0u;
                          }
                          ++k;
                          if (!self && within_h2(r2, h2)) {
                              const float r = __builtin_amdgcn_sqrtf(r2);
                              const float qh = h - r;
                              sumDyn += inside_h(r, h) ? qh * qh * qh : 0.f;
                          }
                      }
                  }
              });
    if ((k & 31u) != 0u && k < kMaskBits) smask[(k >> 5) * kBlock + tid] = bits;
    if (kStatic && S.sCount) {
        walk_runs(cx, cy, cz, S.Ms, S.magicMs, stab + S.stBegin, S.sBegin, S.sBegin + S.sCount,
                  [&](uint32_t s, uint32_t e) {
                      for (uint32_t j = s; j < e; ++j) {
                          const float4 xj = spos[j];
                          const float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
                          const float r2 = dx * dx + dy * dy + dz * dz;
                          if (r2 > h2) continue;
                          const float r = __builtin_amdgcn_sqrtf(r2);
                          const float q = h - r;
                          sumSt += inside_h(r, h) ? q * q * q : 0.f;
                      }
                  });
    }
    const float rho = S.selfDensity + (sumDyn * S.mass + sumSt * S.mass) * S.C;
    // Only this lane writes pos[i].w, and no lane of this pass reads any .w.
    reinterpret_cast<float*>(pos)[4ull * i + 3] = S.k * (rho - S.rho0);
    if (kDebug) dens[i] = rho;
    uint4* mrow = reinterpret_cast<uint4*>(mask + (size_t)i * kMaskWords);
    mrow[0] = make_uint4(smask[0 * kBlock + tid], smask[1 * kBlock + tid], smask[2 * kBlock + tid],
                         smask[3 * kBlock + tid]);
    mrow[1] = make_uint4(smask[4 * kBlock + tid], smask[5 * kBlock + tid], smask[6 * kBlock + tid],
                         smask[7 * kBlock + tid]);
}

template <bool kStatic, bool kDebug, bool kTri>
__global__ __launch_bounds__(kBlock) void k_force_runs(const BlockDesc* __restrict__ blocks, uint32_t numBlocks,
                                                       const SceneDev* __restrict__ scenes,
                                                       const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                       const uint32_t* __restrict__ tab,
                                                       const float4* __restrict__ spos,
                                                       const uint32_t* __restrict__ stab,
                                                       const uint32_t* __restrict__ mask,
                                                       const uint32_t* __restrict__ ids, const float* __restrict__ tris,
                                                       float dt, float4* __restrict__ posT, float4* __restrict__ velT,
                                                       uint32_t* __restrict__ counts, float4* __restrict__ accOut) {
    __shared__ uint32_t smask[(kMaskWords + 1) * kBlock];
    const uint32_t tile = xcd_tile(blockIdx.x, numBlocks);
    if (tile >= numBlocks) return;
    const BlockDesc bd = blocks[tile];
    const SceneDev& S = scenes[bd.scene];
    const uint32_t tid = threadIdx.x;
    const uint32_t i = bd.begin + tid;
    if (i >= bd.end) return;
    {
        const uint4* mrow = reinterpret_cast<const uint4*>(mask + (size_t)i * kMaskWords);
        const uint4 m0 = mrow[0], m1 = mrow[1];
        smask[0 * kBlock + tid] = m0.x;
        smask[1 * kBlock + tid] = m0.y;
        smask[2 * kBlock + tid] = m0.z;
        smask[3 * kBlock + tid] = m0.w;
        smask[4 * kBlock + tid] = m1.x;
        smask[5 * kBlock + tid] = m1.y;
        smask[6 * kBlock + tid] = m1.z;
        smask[7 * kBlock + tid] = m1.w;
        smask[8 * kBlock + tid] = 0u;
    }
    const float4 xi = pos[i];
    const float4 vi = vel[i];
    const float pi = xi.w;  // this step's pressure (density pass)
    const float h = S.h, h2 = S.h2;
    const float rho = pi / S.k + S.rho0;  // SF:544
    const float absP = fabsf(pi);
    const uint32_t cx = cell_coord(xi.x, h), cy = cell_coord(xi.y, h), cz = cell_coord(xi.z, h);
    float fpx = 0.f, fpy = 0.f, fpz = 0.f;
    float fvx = 0.f, fvy = 0.f, fvz = 0.f;
    float fsx = 0.f, fsy = 0.f, fsz = 0.f;
    float sumD2s = 0.f;
    // One dynamic neighbour j (SF:557-589).
    auto pair = [&](uint32_t j, const float4& xj, const float4& vj) {
        float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (j == i || r2 > h2) return;
        const float r = __builtin_amdgcn_sqrtf(r2);
        if (r2 == 0.f || r == 0.f) {
            const float3 d = random_direction((float)(ids[i] - S.pBegin));
            dx = d.x;
            dy = d.y;
            dz = d.z;
        } else {
            const float inv = __builtin_amdgcn_rcpf(r);
            dx *= inv;
            dy *= inv;
            dz *= inv;
        }
        const float q = h - r;
        const bool in = inside_h(r, h);
        const float w1 = in ? -3.0f * q * q : 0.f;
        const float w2 = in ? 6.0f * q : 0.f;
        const float sp = (pi + xj.w) * w1;
        fpx += dx * sp;
        fpy += dy * sp;
        fpz += dz * sp;
        fvx += (vj.x - vi.x) * w2;
        fvy += (vj.y - vi.y) * w2;
        fvz += (vj.z - vi.z) * w2;
    };
    uint32_t k0 = 0;
    walk_runs(cx, cy, cz, S.M, S.magicM, tab + S.tBegin, S.pBegin, S.pBegin + S.pCount,
              [&](uint32_t s, uint32_t e) {
                  const uint32_t n = e > s ? e - s : 0u;
                  for (uint32_t off = 0; off < n; off += 32u) {
                      const uint32_t kk = k0 + off;
                      uint32_t bits;
                      if (kk + 32u <= kMaskBits) {
                          const uint32_t w = kk >> 5;
                          bits = __builtin_amdgcn_alignbit(smask[(w + 1) * kBlock + tid], smask[w * kBlock + tid],
                                                           kk & 31u);
                      } else if (kk >= kMaskBits) {
                          bits = ~0u;  // past the recorded candidates: test them all
                      } else {
                          const uint32_t w = kk >> 5;
                          const uint32_t rec = kMaskBits - kk;  // recorded positions in this chunk (< 32)
                          const uint32_t m = __builtin_amdgcn_alignbit(smask[(w + 1) * kBlock + tid],
                                                                       smask[w * kBlock + tid], kk & 31u);
                          bits = (m & ((1u << rec) - 1u)) | ~((1u << rec) - 1u);
                      }
                      if (n - off < 32u) bits &= (1u << (n - off)) - 1u;
                      // Two neighbours per iteration, all four loads issued before use.
                      while (bits) {
                          const uint32_t b0 = (uint32_t)__builtin_ctz(bits);
                          bits &= bits - 1u;
                          const bool two = bits != 0u;
                          const uint32_t b1 = two ? (uint32_t)__builtin_ctz(bits) : b0;
                          if (two) bits &= bits - 1u;
                          const uint32_t j0 = s + off + b0, j1 = s + off + b1;
                          const float4 x0 = pos[j0], v0 = vel[j0];
                          const float4 x1 = pos[j1], v1 = vel[j1];
                          pair(j0, x0, v0);
                          if (two) pair(j1, x1, v1);
                      }
                  }
                  k0 += n;
              });
    if (kStatic && S.sCount) {
        walk_runs(cx, cy, cz, S.Ms, S.magicMs, stab + S.stBegin, S.sBegin, S.sBegin + S.sCount,
                  [&](uint32_t s, uint32_t e) {
                      for (uint32_t j = s; j < e; ++j) {
                          const float4 xj = spos[j];
                          float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
                          const float r2 = dx * dx + dy * dy + dz * dz;
                          if (r2 > h2) continue;
                          const float r = __builtin_amdgcn_sqrtf(r2);
                          if (r2 == 0.f || r == 0.f) {
                              const float3 d = random_direction((float)(ids[i] - S.pBegin));
                              dx = d.x;
                              dy = d.y;
                              dz = d.z;
                          } else {
                              const float inv = __builtin_amdgcn_rcpf(r);
                              dx *= inv;
                              dy *= inv;
                              dz *= inv;
                          }
                          const float q = h - r;
                          const bool in = inside_h(r, h);
                          const float sp = absP * (in ? -3.0f * q * q : 0.f);
                          fsx += dx * sp;
                          fsy += dy * sp;
                          fsz += dz * sp;
                          sumD2s += in ? 6.0f * q : 0.f;
                      }
                  });
    }
    // SF:626-638
    const float sP = S.mass / (2.0f * rho) * S.C;
    const float sV = S.mu * S.mass * S.C;
    const float sVs = S.mu * S.mass * S.C * 0.0f;
    float Fx = fpx * sP + fvx * sV + fsx * sP + (-vi.x * sumD2s) * sVs;
    float Fy = fpy * sP + fvy * sV + fsy * sP + (-vi.y * sumD2s) * sVs;
    float Fz = fpz * sP + fvz * sV + fsz * sP + (-vi.z * sumD2s) * sVs;
    if (kTri) mesh_force(make_float3(xi.x, xi.y, xi.z), S, tris, Fx, Fy, Fz);
    const float ax = Fx / rho + S.gx, ay = Fy / rho + S.gy, az = Fz / rho + S.gz;
    if (kDebug) accOut[i] = make_float4(ax, ay, az, 0.f);
    const float nvx = vi.x + ax * dt, nvy = vi.y + ay * dt, nvz = vi.z + az * dt;
    const float nx = xi.x + nvx * dt, ny = xi.y + nvy * dt, nz = xi.z + nvz * dt;
    const uint32_t code = morton3(cell_coord(nx, h), cell_coord(ny, h), cell_coord(nz, h));
    const uint32_t key = fastmod(code, S.magicM, S.M);
    posT[i] = make_float4(nx, ny, nz, pi);
    velT[i] = make_float4(nvx, nvy, nvz, __uint_as_float(key));
    const uint32_t gk = S.tBegin + key;
    const Run run = wave_run(gk);
    if (run.is_head) atomicAdd(&counts[gk], run.len);
}

// ---------------------------------------------------------------------------
// k_scan: inclusive scan of n table entries in place, single pass with
// decoupled lookback (Merrill and Garland, "Single-pass Parallel Prefix Scan
// with Decoupled Look-back", 2016). Tiles take a ticket in launch order, so a
// tile only waits on tiles that are already running. Status words carry a
// 2-bit flag (1 = aggregate, 2 = inclusive prefix) over a 32-bit value in one
// 64-bit atomic. The kernel also zeroes `zero_out` (the other table, which
// collects the next step's histogram), the other status array and the n_clear
// words of `clear` (the per-scene counts the next force pass accumulates), so
// no memset is needed between steps and the launch can live in a hipGraph.
constexpr int kScanThreads = 256;
#ifndef SPH_SCAN_ROWS
#define SPH_SCAN_ROWS 16
#endif
constexpr int kScanRows = SPH_SCAN_ROWS;
constexpr uint32_t kScanTile = kScanThreads * 4u * kScanRows;  // 16384 entries per tile
constexpr unsigned long long kStatAgg = 1ull << 32;
constexpr unsigned long long kStatIncl = 2ull << 32;

__device__ inline uint32_t wave_inclusive_scan(uint32_t v) {
    const uint32_t lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d);
        if (lane >= (uint32_t)d) v += t;
    }
    return v;
}

// Inclusive scan across the wave: DPP row shifts within each row of 16 lanes,
// then the rows' totals (lanes 15, 31, 47) carried into the later rows.
__device__ inline uint32_t wave_inclusive_scan_dpp(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
    const uint32_t lane = __lane_id();
    const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)v, 15);
    const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)v, 31);
    const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)v, 47);
    return v + (lane >= 16u ? r0 : 0u) + (lane >= 32u ? r1 : 0u) + (lane >= 48u ? r2 : 0u);
}

__device__ inline uint32_t block_exclusive_scan(uint32_t x, uint32_t& total, uint32_t* s_w) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t inc = wave_inclusive_scan(x);
    if (lane == 63u) s_w[wave] = inc;
    __syncthreads();
    const uint32_t w0 = s_w[0], w1 = s_w[1], w2 = s_w[2], w3 = s_w[3];
    __syncthreads();
    total = w0 + w1 + w2 + w3;
    const uint32_t off = (wave > 0 ? w0 : 0u) + (wave > 1 ? w1 : 0u) + (wave > 2 ? w2 : 0u);
    return off + inc - x;
}

__global__ __launch_bounds__(kScanThreads) void k_scan(uint32_t* __restrict__ data, uint32_t* __restrict__ zero_out,
                                                       uint32_t n, unsigned long long* __restrict__ status,
                                                       unsigned long long* __restrict__ status_clear,
                                                       uint32_t* __restrict__ ticket, uint32_t num_tiles,
                                                       uint32_t* __restrict__ clear, uint32_t n_clear) {
    __shared__ uint32_t s_tile, s_prefix;
    __shared__ uint32_t s_tot[kScanRows][4];  // wave totals of each row
    const uint32_t tid = threadIdx.x;
    if (tid == 0) {
        const uint32_t t = atomicAdd(ticket, 1u);
        if (t == num_tiles - 1u) atomicExch(ticket, 0u);  // every tile holds its ticket by now
        s_tile = t;
    }
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * kScanTile;
    uint4 v[kScanRows];
    uint32_t excl[kScanRows];  // entries of the tile before this thread's four in row r
    const uint32_t lane = tid & 63u, wave = tid >> 6;
#pragma unroll
    for (int r = 0; r < kScanRows; ++r) {
        const uint32_t idx = base + (uint32_t)r * (kScanThreads * 4u) + tid * 4u;
        if (idx + 3u < n) {
            v[r] = *reinterpret_cast<const uint4*>(data + idx);
        } else {
            v[r].x = idx < n ? data[idx] : 0u;
            v[r].y = idx + 1u < n ? data[idx + 1] : 0u;
            v[r].z = idx + 2u < n ? data[idx + 2] : 0u;
            v[r].w = 0u;
        }
    }
    // Every row's wave scans at once, then one barrier for the wave totals.
#pragma unroll
    for (int r = 0; r < kScanRows; ++r) {
        const uint32_t sum = v[r].x + v[r].y + v[r].z + v[r].w;
        const uint32_t inc = wave_inclusive_scan_dpp(sum);
        if (lane == 63u) s_tot[r][wave] = inc;
        excl[r] = inc - sum;
    }
    __syncthreads();
    uint32_t agg = 0;
#pragma unroll
    for (int r = 0; r < kScanRows; ++r) {
        const uint32_t t0 = s_tot[r][0], t1 = s_tot[r][1], t2 = s_tot[r][2], t3 = s_tot[r][3];
        excl[r] += agg + (wave > 0u ? t0 : 0u) + (wave > 1u ? t1 : 0u) + (wave > 2u ? t2 : 0u);
        agg += t0 + t1 + t2 + t3;
    }
    if (tid < 64u) {
        // Look-back by the first wave, 64 predecessors at a time: lane l reads
        // the status of tile t - 1 - l; the lanes up to the nearest inclusive
        // prefix must all be published, and their values sum to this tile's
        // exclusive prefix (tiles before the first count as an inclusive 0).
        uint32_t prefix = 0;
        if (tile == 0) {
            if (tid == 0) __hip_atomic_store(&status[0], kStatIncl | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (tid == 0) __hip_atomic_store(&status[tile], kStatAgg | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int32_t j0 = (int32_t)tile - 1;
            for (;;) {
                const int32_t j = j0 - (int32_t)tid;
                const unsigned long long st =
                    j >= 0 ? __hip_atomic_load(&status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : kStatIncl;
                const uint32_t flag = (uint32_t)(st >> 32);
                const uint64_t incl = __ballot(flag == 2u);
                const uint32_t first = incl ? (uint32_t)__builtin_ctzll(incl) : 63u;  // last lane that counts
                const uint64_t upto = first == 63u ? ~0ull : (2ull << first) - 1ull;
                if (__ballot(flag == 0u) & upto) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                uint32_t val = tid <= first ? (uint32_t)st : 0u;
#pragma unroll
                for (int d = 32; d >= 1; d >>= 1) val += (uint32_t)__shfl_xor((int)val, d);
                prefix += val;
                if (incl) break;
                j0 -= 64;
            }
            if (tid == 0)
                __hip_atomic_store(&status[tile], kStatIncl | (uint32_t)(prefix + agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        if (tid == 0) {
            s_prefix = prefix;
            status_clear[tile] = 0ull;
        }
    }
    __syncthreads();
    const uint32_t prefix = s_prefix;
#pragma unroll
    for (int r = 0; r < kScanRows; ++r) {
        const uint32_t idx = base + (uint32_t)r * (kScanThreads * 4u) + tid * 4u;
        uint4 o;
        o.x = prefix + excl[r] + v[r].x;
        o.y = o.x + v[r].y;
        o.z = o.y + v[r].z;
        o.w = o.z + v[r].w;
        if (idx + 3u < n) {
            *reinterpret_cast<uint4*>(data + idx) = o;
            *reinterpret_cast<uint4*>(zero_out + idx) = make_uint4(0u, 0u, 0u, 0u);
        } else {
            if (idx < n) {
                data[idx] = o.x;
                zero_out[idx] = 0u;
            }
            if (idx + 1u < n) {
                data[idx + 1] = o.y;
                zero_out[idx + 1] = 0u;
            }
            if (idx + 2u < n) {
                data[idx + 2] = o.z;
                zero_out[idx + 2] = 0u;
            }
        }
    }
    if (tile == 0u)
        for (uint32_t t = tid; t < n_clear; t += kScanThreads) clear[t] = 0u;
}

// ---------------------------------------------------------------------------
// k_scatter: SF:346-356 / SF:332-345. After the inclusive scan cursor[b] is the
// end of bucket b; each run of equal keys in a wave takes len slots with one
// atomic and keeps its wave order, so afterwards cursor[b] is the bucket start
// (the reference's final hashMap). The particle state itself moves, so the next
// step's neighbour reads are contiguous per bucket.
// pmapOut (particle-map mode, a reordering step: SF:332-345): the state moves
// into bucket order and the map becomes the identity.
// off16 (mixed precision): also the particle's offset record at its new slot.
__device__ inline void scatter_block(uint32_t blk, const BlockDesc* __restrict__ blocks,
                                     const SceneDev* __restrict__ scenes, const float4* __restrict__ posT,
                                     const float4* __restrict__ velT, const uint32_t* __restrict__ idIn,
                                     uint32_t* __restrict__ cursor, float4* __restrict__ pos,
                                     float4* __restrict__ vel, uint32_t* __restrict__ idOut,
                                     uint32_t* __restrict__ pmapOut, uint2* __restrict__ off16) {
    const BlockDesc bd = blocks[blk];
    const SceneDev& S = scenes[bd.scene];
    const uint32_t tBegin = S.tBegin;
    const uint32_t i = bd.begin + threadIdx.x;
    if (i >= bd.end) return;
    const float4 p = posT[i];
    const float4 v = velT[i];
    const uint32_t id = idIn[i];
    const uint32_t gk = tBegin + __float_as_uint(v.w);
    const Run run = wave_run(gk);
    uint32_t old = 0;
    if (run.is_head) old = atomicSub(&cursor[gk], run.len);
    old = __shfl(old, (int)run.head);
    const uint32_t dst = old - run.len + run.rank;
    // A correct cursor table always yields dst in the scene's range; the guard
    // keeps a corrupted one from writing outside the arrays (sph_validate then
    // reports the broken buckets).
    if (dst < S.pBegin || dst >= S.pBegin + S.pCount) return;
    pos[dst] = p;
    vel[dst] = v;
    idOut[dst] = id;
    if (pmapOut) pmapOut[dst] = dst;
    if (off16) off16[dst] = offset_record(p, S.h);
}

__global__ __launch_bounds__(kBlock) void k_scatter(const BlockDesc* __restrict__ blocks,
                                                    const SceneDev* __restrict__ scenes,
                                                    const float4* __restrict__ posT, const float4* __restrict__ velT,
                                                    const uint32_t* __restrict__ idIn, uint32_t* __restrict__ cursor,
                                                    float4* __restrict__ pos, float4* __restrict__ vel,
                                                    uint32_t* __restrict__ idOut, uint32_t* __restrict__ pmapOut,
                                                    uint2* __restrict__ off16) {
    scatter_block(blockIdx.x, blocks, scenes, posT, velT, idIn, cursor, pos, vel, idOut, pmapOut, off16);
}

// The offset records of the current state (mixed precision), after the state was
// set other than by the scatter: at initialization and after sph_write.
__global__ __launch_bounds__(kBlock) void k_offsets(const BlockDesc* __restrict__ blocks,
                                                    const SceneDev* __restrict__ scenes,
                                                    const float4* __restrict__ pos, uint2* __restrict__ off16) {
    const BlockDesc bd = blocks[blockIdx.x];
    const uint32_t i = bd.begin + threadIdx.x;
    if (i >= bd.end) return;
    off16[i] = offset_record(pos[i], scenes[bd.scene].h);
}

// Particle-map mode, a step that does not reorder (SF:346-356,
// FillDynamicParticleMapAndFinishHashMap): the state keeps its memory order (this
// step's output becomes the next step's input at the same index) and slot
// --cursor[key] of the particle's bucket takes its memory index, the runs of
// equal keys in a wave taking their slots with one atomicSub as in scatter_block.
__global__ __launch_bounds__(kBlock) void k_fill_map(const BlockDesc* __restrict__ blocks,
                                                     const SceneDev* __restrict__ scenes,
                                                     const float4* __restrict__ posT, const float4* __restrict__ velT,
                                                     const uint32_t* __restrict__ idIn, uint32_t* __restrict__ cursor,
                                                     float4* __restrict__ pos, float4* __restrict__ vel,
                                                     uint32_t* __restrict__ idOut, uint32_t* __restrict__ pmap) {
    const BlockDesc bd = blocks[blockIdx.x];
    const SceneDev& S = scenes[bd.scene];
    const uint32_t i = bd.begin + threadIdx.x;
    if (i >= bd.end) return;
    const float4 p = posT[i];
    const float4 v = velT[i];
    pos[i] = p;
    vel[i] = v;
    idOut[i] = idIn[i];
    const uint32_t gk = S.tBegin + __float_as_uint(v.w);
    const Run run = wave_run(gk);
    uint32_t old = 0;
    if (run.is_head) old = atomicSub(&cursor[gk], run.len);
    old = __shfl(old, (int)run.head);
    const uint32_t dst = old - run.len + run.rank;
    if (dst < S.pBegin || dst >= S.pBegin + S.pCount) return;  // as in scatter_block
    pmap[dst] = i;
}

// ---------------------------------------------------------------------------
// Diagnostics and readback (not on the step path).

// FP64 partial sums per block: {sum vx, vy, vz, sum |v|^2, sum x, y, z, count}.
__global__ __launch_bounds__(kBlock) void k_conserved_partial(const BlockDesc* __restrict__ blocks,
                                                              const float4* __restrict__ pos,
                                                              const float4* __restrict__ vel,
                                                              double* __restrict__ partial) {
    __shared__ double s[4][8];
    const BlockDesc bd = blocks[blockIdx.x];
    const uint32_t i = bd.begin + threadIdx.x;
    double q[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (i < bd.end) {
        const float4 p = pos[i], v = vel[i];
        q[0] = v.x;
        q[1] = v.y;
        q[2] = v.z;
        q[3] = (double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z;
        q[4] = p.x;
        q[5] = p.y;
        q[6] = p.z;
        q[7] = 1.0;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        double x = q[k];
        for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d);
        q[k] = x;
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < 8; ++k) s[wave][k] = q[k];
    __syncthreads();
    if (threadIdx.x < 8) {
        const int k = threadIdx.x;
        partial[8ull * blockIdx.x + k] = ((s[0][k] + s[1][k]) + s[2][k]) + s[3][k];
    }
}

// One workgroup per scene sums its blocks' partials in a fixed order.
__global__ __launch_bounds__(64) void k_conserved_final(const uint32_t* __restrict__ sceneBlocks,
                                                        const SceneDev* __restrict__ scenes,
                                                        const double* __restrict__ partial, double* __restrict__ out) {
    const uint32_t s = blockIdx.x;
    const uint32_t b0 = sceneBlocks[s], b1 = sceneBlocks[s + 1];
    const uint32_t lane = threadIdx.x;
    double q[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t b = b0 + lane; b < b1; b += 64u)
        for (int k = 0; k < 8; ++k) q[k] += partial[8ull * b + k];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        double x = q[k];
        for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d);
        q[k] = x;
    }
    if (lane == 0) {
        const double m = scenes[s].mass;
        const double mass = q[7] * m;
        out[8 * s + 0] = mass;
        out[8 * s + 1] = m * q[0];
        out[8 * s + 2] = m * q[1];
        out[8 * s + 3] = m * q[2];
        out[8 * s + 4] = 0.5 * m * q[3];
        out[8 * s + 5] = mass > 0 ? m * q[4] / mass : 0.0;
        out[8 * s + 6] = mass > 0 ? m * q[5] / mass : 0.0;
        out[8 * s + 7] = mass > 0 ? m * q[6] / mass : 0.0;
    }
}

// SimulationEngine.h:115-177 and 229-282 for every dynamic particle: finite
// state, stored key equal to the key of its position, and the particle inside
// its bucket's range.
// In particle-map mode (pmap non-null) bucket slot i holds particle pmap[i], whose
// stored key must be that bucket's (SimulationEngine.h:266-274).
__global__ __launch_bounds__(kBlock) void k_validate_particles(const BlockDesc* __restrict__ blocks,
                                                               const SceneDev* __restrict__ scenes,
                                                               const float4* __restrict__ pos,
                                                               const float4* __restrict__ vel,
                                                               const uint32_t* __restrict__ tab,
                                                               const uint32_t* __restrict__ pmap,
                                                               uint32_t* __restrict__ bits) {
    const BlockDesc bd = blocks[blockIdx.x];
    const SceneDev& S = scenes[bd.scene];
    const uint32_t i = bd.begin + threadIdx.x;
    if (i >= bd.end) return;
    const float4 p = pos[i], v = vel[i];
    uint32_t e = 0;
    if (!isfinite(p.x) || !isfinite(p.y) || !isfinite(p.z) || !isfinite(p.w) || !isfinite(v.x) ||
        !isfinite(v.y) || !isfinite(v.z))
        e |= 0x1u;
    const uint32_t key = __float_as_uint(v.w);
    const uint32_t want = fastmod(morton3(cell_coord(p.x, S.h), cell_coord(p.y, S.h), cell_coord(p.z, S.h)),
                                  S.magicM, S.M);
    if (key != want) e |= 0x2u;
    uint32_t sk = key;  // key of the particle listed in slot i
    if (pmap) {
        const uint32_t j = pmap[i];
        sk = (j >= S.pBegin && j < S.pBegin + S.pCount) ? __float_as_uint(vel[j].w) : 0xFFFFFFFFu;
    }
    if (sk < S.M) {
        if (i < tab[S.tBegin + sk] || i >= tab[S.tBegin + sk + 1]) e |= 0x20u;
    } else {
        e |= 0x20u;
    }
    if (e) atomicOr(bits, e);
}

// Table invariants of SimulationEngine.h:229-282 over one table: per scene the
// first entry is the scene's first particle, the sentinel is one past its last,
// and entries never decrease.
__global__ __launch_bounds__(kBlock) void k_validate_table(const uint32_t* __restrict__ tab, uint32_t n,
                                                           const uint32_t* __restrict__ firstIdx,
                                                           const uint32_t* __restrict__ firstVal,
                                                           const uint32_t* __restrict__ lastIdx,
                                                           const uint32_t* __restrict__ lastVal, uint32_t nScenes,
                                                           uint32_t shift, uint32_t* __restrict__ bits) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t e = 0;
    if (x + 1u < n && tab[x] > tab[x + 1]) e |= 0x10u;
    if (x < nScenes) {
        if (tab[firstIdx[x]] != firstVal[x]) e |= 0x4u;
        if (tab[lastIdx[x]] != lastVal[x]) e |= 0x8u;
    }
    if (e) atomicOr(bits, e << shift);
}

// Static particles: key of the position must place the particle in its bucket.
__global__ __launch_bounds__(kBlock) void k_validate_static(const SceneDev* __restrict__ scenes, uint32_t nScenes,
                                                            const uint32_t* __restrict__ staticScene,
                                                            const float4* __restrict__ spos,
                                                            const uint32_t* __restrict__ stab, uint32_t ns,
                                                            uint32_t* __restrict__ bits) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    const SceneDev& S = scenes[staticScene[i]];
    const float4 p = spos[i];
    const uint32_t key =
        fastmod(morton3(cell_coord(p.x, S.h), cell_coord(p.y, S.h), cell_coord(p.z, S.h)), S.magicMs, S.Ms);
    if (i < stab[S.stBegin + key] || i >= stab[S.stBegin + key + 1] || !isfinite(p.x) || !isfinite(p.y) ||
        !isfinite(p.z))
        atomicOr(bits, 0x40u);
}

// Current state back to original order.
// outIds[k]: the original index of the particle in bucket slot k (slot k's
// particle is pmap[k] in particle-map mode, else k).
__global__ __launch_bounds__(kBlock) void k_unsort(const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                   const uint32_t* __restrict__ ids,
                                                   const uint32_t* __restrict__ origOfSlot, uint32_t n,
                                                   float* __restrict__ outPos, float* __restrict__ outVel,
                                                   float* __restrict__ outP, uint32_t* __restrict__ outKey,
                                                   uint32_t* __restrict__ outIds, const uint32_t* __restrict__ pmap) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t o = origOfSlot[ids[k]];
    const float4 p = pos[k], v = vel[k];
    outPos[3ull * o + 0] = p.x;
    outPos[3ull * o + 1] = p.y;
    outPos[3ull * o + 2] = p.z;
    outVel[3ull * o + 0] = v.x;
    outVel[3ull * o + 1] = v.y;
    outVel[3ull * o + 2] = v.z;
    outP[o] = p.w;
    if (outKey) outKey[o] = __float_as_uint(v.w);
    if (outIds) outIds[k] = pmap ? origOfSlot[ids[pmap[k]]] : o;
}

// The inverse of k_unsort for a write: slot k takes the position and velocity of
// its original index from inPos / inVel (either may be null: kept), its pressure
// (pos.w) and stored key (vel.w) unchanged.
__global__ __launch_bounds__(kBlock) void k_write(float4* __restrict__ pos, float4* __restrict__ vel,
                                                  const uint32_t* __restrict__ ids,
                                                  const uint32_t* __restrict__ origOfSlot, uint32_t n,
                                                  const float* __restrict__ inPos, const float* __restrict__ inVel) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t o = origOfSlot[ids[k]];
    if (inPos) {
        float4 p = pos[k];
        p.x = inPos[3ull * o + 0];
        p.y = inPos[3ull * o + 1];
        p.z = inPos[3ull * o + 2];
        pos[k] = p;
    }
    if (inVel) {
        float4 v = vel[k];
        v.x = inVel[3ull * o + 0];
        v.y = inVel[3ull * o + 1];
        v.z = inVel[3ull * o + 2];
        vel[k] = v;
    }
}

// Per-particle debug arrays of the last step (indexed in that step's order).
__global__ __launch_bounds__(kBlock) void k_unsort_debug(const float* __restrict__ dens,
                                                         const float4* __restrict__ acc,
                                                         const uint32_t* __restrict__ ids,
                                                         const uint32_t* __restrict__ origOfSlot, uint32_t n,
                                                         float* __restrict__ outD, float* __restrict__ outA) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t o = origOfSlot[ids[k]];
    outD[o] = dens[k];
    const float4 a = acc[k];
    outA[3ull * o + 0] = a.x;
    outA[3ull * o + 1] = a.y;
    outA[3ull * o + 2] = a.z;
}

// Plain streaming copy of n16 16-byte words (counter calibration).
__global__ __launch_bounds__(256) void k_copy_probe(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                    uint64_t n16) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n16; k += stride) dst[k] = src[k];
}

}  // namespace sph

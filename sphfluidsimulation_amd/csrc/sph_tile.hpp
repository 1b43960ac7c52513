// LDS-staged neighbour passes ("tile walk", selected with SPH_FLAG_TILE_WALK).
//
// A workgroup owns 256 consecutive particles in bucket order. Before any
// neighbour is read it builds, in LDS, the set of distinct buckets its
// particles' 27 neighbour cells map to (an open-addressing map keyed by bucket),
// then copies those buckets' particles from HBM into LDS once. Each particle
// then walks its 27 neighbour cells in the reference's order (SF:397-404),
// looks each cell's bucket up in the map and reads that bucket's candidates
// from LDS. A bucket that two of the 27 cells collide on is walked twice, as
// the reference does (SF:397-404 visits each cell's bucket; with M below the
// Morton range two cells can share one); a bucket that does not fit in LDS, or
// did not fit in the map, is read from HBM instead, so every M works.
//
// Particles in the same cell share their 27 buckets, so only the first
// particle of each cell inserts them (a 256-particle tile holds 32-64 cells
// and needs about 144 distinct buckets on a fluid lattice at spacing h/2).
//
// The density pass first tests each bucket's candidates against a widened
// radius only, records the in-range ones in a per-particle byte mask, and then
// evaluates the kernel for the flagged candidates alone, so the square root and
// cube are paid per neighbour rather than per candidate. The force pass reads
// the mask and evaluates only the recorded neighbours. SF =
// SPHFluidSimulation/source/SPH/Kernels/SPHFunctions.cpp.
#pragma once

#include "sph_kernels.hpp"

namespace sph {

constexpr uint32_t kMapSlots = 512;      // bucket map slots per workgroup
constexpr uint32_t kMapProbe = 32;       // linear-probe limit; past it a bucket stays unmapped
constexpr uint32_t kCapDensity = 1280;   // staged candidates of the density pass (16 B each)
constexpr uint32_t kCapForce = 1024;     // staged candidates of the force pass (32 B each)
constexpr uint32_t kNotStaged = 0xFFFFFFFFu;
static_assert(kMapSlots == 2 * kBlock, "staging assigns two map slots per thread");

__device__ inline uint32_t map_hash(uint32_t b) { return (b * 2654435761u) >> 23; }  // 9 bits

// cs.x[a] | cs.y[b] | cs.z[c] with selects instead of indexed loads, so the
// rolled 27-cell loops keep the spread coordinates in registers.
__device__ inline uint32_t pick3(const uint32_t (&v)[3], int k) { return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]); }
__device__ inline uint32_t cell_code(const CellSpread& cs, int a, int b, int c) {
    return pick3(cs.x, a) | pick3(cs.y, b) | pick3(cs.z, c);
}

// Bucket map in LDS: key = bucket + 1 (0 = empty), then the bucket's global
// particle range and its LDS offset (kNotStaged if it lives in HBM only).
struct BucketMap {
    uint32_t key[kMapSlots];
    uint32_t beg[kMapSlots];
    uint32_t cnt[kMapSlots];
    uint32_t off[kMapSlots];
};

__device__ inline void map_insert(BucketMap& m, uint32_t b) {
    uint32_t s = map_hash(b);
    for (uint32_t p = 0; p < kMapProbe; ++p) {
        const uint32_t prev = atomicCAS(&m.key[s], 0u, b + 1u);
        if (prev == 0u || prev == b + 1u) return;
        s = (s + 1u) & (kMapSlots - 1u);
    }
}

__device__ inline int map_find(const BucketMap& m, uint32_t b) {
    uint32_t s = map_hash(b);
    for (uint32_t p = 0; p < kMapProbe; ++p) {
        const uint32_t k = m.key[s];
        if (k == b + 1u) return (int)s;
        if (k == 0u) return -1;
        s = (s + 1u) & (kMapSlots - 1u);
    }
    return -1;
}

// One neighbour bucket of the walk: global range [beg, beg + cnt) and the LDS
// offset of its staged copy.
struct BucketView {
    uint32_t beg, cnt, off;
};

__device__ inline BucketView bucket_view(const BucketMap& m, uint32_t bucket, const uint32_t* __restrict__ table,
                                         uint32_t lo, uint32_t hi) {
    BucketView v;
    const int slot = map_find(m, bucket);
    if (slot >= 0) {
        v.beg = m.beg[slot];
        v.cnt = m.cnt[slot];
        v.off = m.off[slot];
    } else {
        const uint32_t tb = table[bucket], te = table[bucket + 1];
        const uint32_t s = tb > lo ? tb : lo, e = te < hi ? te : hi;
        v.beg = s;
        v.cnt = e > s ? e - s : 0u;
        v.off = kNotStaged;
    }
    return v;
}

// Builds the bucket map of a tile and stages the mapped buckets' particles.
// `valid` lanes hold particle i in cell code `code`; every lane must call this
// (it contains barriers).
template <bool kVel>
__device__ inline void stage_tile(BucketMap& m, uint32_t* s_code, uint32_t* s_w, bool valid, const CellSpread& cs,
                                  uint32_t code, const SceneDev& S, const uint32_t* __restrict__ table,
                                  const float4* __restrict__ pos, const float4* __restrict__ vel, float4* candP,
                                  float4* candV, uint32_t cap) {
    const uint32_t tid = threadIdx.x;
    m.key[tid] = 0u;
    m.key[tid + kBlock] = 0u;
    s_code[tid] = valid ? code : 0xFFFFFFFFu;
    __syncthreads();
    const uint32_t lo = S.pBegin, hi = S.pBegin + S.pCount;
    // The first particle of each cell inserts the cell's 27 buckets.
    if (valid && (tid == 0 || s_code[tid - 1] != code)) {
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c) map_insert(m, fastmod(cell_code(cs, a, b, c), S.magicM, S.M));
    }
    __syncthreads();
    // Two slots per thread: range, count, then an exclusive scan for LDS offsets.
    uint32_t cnt2[2], beg2[2], sum = 0;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const uint32_t slot = 2u * tid + (uint32_t)q;
        const uint32_t k = m.key[slot];
        cnt2[q] = 0u;
        beg2[q] = 0u;
        if (k) {
            const uint32_t bk = k - 1u;
            const uint32_t tb = table[bk], te = table[bk + 1];
            const uint32_t s = tb > lo ? tb : lo, e = te < hi ? te : hi;
            beg2[q] = s;
            cnt2[q] = e > s ? e - s : 0u;
        }
        sum += cnt2[q];
    }
    uint32_t total;
    uint32_t o = block_exclusive_scan(sum, total, s_w);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const uint32_t slot = 2u * tid + (uint32_t)q;
        const bool fits = o + cnt2[q] <= cap;
        m.beg[slot] = beg2[q];
        m.cnt[slot] = cnt2[q];
        m.off[slot] = fits ? o : kNotStaged;
        if (fits)
            for (uint32_t t = 0; t < cnt2[q]; t += 4u) {
                float4 p4[4], v4[4];
                const uint32_t last = cnt2[q] - 1u;
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u) {
                    p4[u] = pos[beg2[q] + min(t + u, last)];
                    if (kVel) v4[u] = vel[beg2[q] + min(t + u, last)];
                }
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u)
                    if (t + u < cnt2[q]) {
                        candP[o + t + u] = p4[u];
                        if (kVel) candV[o + t + u] = v4[u];
                    }
            }
        o += cnt2[q];
    }
    __syncthreads();
}

// Candidate mask of the tile walk: byte v (v = 0..26, the neighbour cells in
// the reference's x-y-z loop order) flags which of the first 8 candidates of
// that cell's bucket lie within h (widened test). Candidates past the 8th of a
// bucket are not recorded; the force pass tests them itself. 27 bytes padded
// to 8 words per particle.
constexpr uint32_t kTileMaskWords = 8;

template <bool kStatic, bool kDebug>
__global__ __launch_bounds__(kBlock) void k_density_tile(const BlockDesc* __restrict__ blocks, uint32_t numBlocks,
                                                         const SceneDev* __restrict__ scenes, float4* pos,
                                                         const uint32_t* __restrict__ tab,
                                                         const float4* __restrict__ spos,
                                                         const uint32_t* __restrict__ stab,
                                                         uint32_t* __restrict__ mask, float* __restrict__ dens) {
    __shared__ BucketMap m;
    __shared__ float4 candP[kCapDensity];
    __shared__ uint32_t smask[kTileMaskWords * kBlock];
    __shared__ uint32_t s_code[kBlock];
    __shared__ uint32_t s_w[4];
    const uint32_t tile = xcd_tile(blockIdx.x, numBlocks);
    if (tile >= numBlocks) return;  // whole workgroup leaves together
    const BlockDesc bd = blocks[tile];
    const SceneDev& S = scenes[bd.scene];
    const uint32_t tid = threadIdx.x;
    const uint32_t i = bd.begin + tid;
    const bool valid = i < bd.end;
    const float4 xi = valid ? pos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float h = S.h, h2 = S.h2;
    const float h2wide = S.h2wide;
    const uint32_t cx = cell_coord(xi.x, h), cy = cell_coord(xi.y, h), cz = cell_coord(xi.z, h);
    const CellSpread cs = cell_spread(xi.x, xi.y, xi.z, h);
    const uint32_t code = morton3(cx, cy, cz);
    const uint32_t* table = tab + S.tBegin;
    const uint32_t lo = S.pBegin, hi = S.pBegin + S.pCount;
    stage_tile<false>(m, s_code, s_w, valid, cs, code, S, table, pos, nullptr, candP, nullptr, kCapDensity);
    if (!valid) return;

    float sumDyn = 0.f, sumSt = 0.f;
    // The candidates of one bucket visit, read through `src` (the LDS copy or
    // HBM): widened test of the first 8 (four loads in flight), then the exact
    // contribution (SF:426-453) of the flagged ones, then any past the 8th.
    // Returns the visit's mask byte.
    auto visit_bucket = [&](const float4* src, uint32_t beg, uint32_t cnt) -> uint32_t {
        const uint32_t n8 = cnt < 8u ? cnt : 8u;
        const uint32_t selfT = i - beg;  // >= cnt unless i is in this bucket
        uint32_t bits8 = 0;
        for (uint32_t t = 0; t < n8; t += 4u) {
            float4 xs[4];
#pragma unroll
            for (uint32_t q = 0; q < 4; ++q) xs[q] = src[min(t + q, n8 - 1u)];
#pragma unroll
            for (uint32_t q = 0; q < 4; ++q) {
                const float dx = xs[q].x - xi.x, dy = xs[q].y - xi.y, dz = xs[q].z - xi.z;
                const float r2 = dx * dx + dy * dy + dz * dz;
                const bool in = (t + q < n8) && (t + q != selfT) && within_h2(r2, h2wide);
                bits8 |= in ? (1u << (t + q)) : 0u;
            }
        }
        auto d0 = [&](uint32_t t) {
            const float4 xj = src[t];
            const float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (within_h2(r2, h2) && t != selfT) {
                const float r = __builtin_amdgcn_sqrtf(r2);
                const float q = h - r;
                sumDyn += inside_h(r, h) ? q * q * q : 0.f;
            }
        };
        for (uint32_t mb = bits8; mb; mb &= mb - 1u) d0((uint32_t)__builtin_ctz(mb));
        for (uint32_t t = 8; t < cnt; ++t) d0(t);
        return bits8;
    };
    uint32_t word = 0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c) {
                const uint32_t visit = (uint32_t)(a * 9 + b * 3 + c);
                const uint32_t bucket = fastmod(cell_code(cs, a, b, c), S.magicM, S.M);
                const BucketView v = bucket_view(m, bucket, table, lo, hi);
                const uint32_t bits8 = v.off != kNotStaged ? visit_bucket(candP + v.off, v.beg, v.cnt)
                                                           : visit_bucket(pos + v.beg, v.beg, v.cnt);
                word |= bits8 << (8u * (visit & 3u));
                if ((visit & 3u) == 3u || visit == 26u) {
                    smask[(visit >> 2) * kBlock + tid] = word;
                    word = 0u;
                }
            }
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
    // Only this lane writes pos[i].w; no lane of this pass reads any .w.
    reinterpret_cast<float*>(pos)[4ull * i + 3] = S.k * (rho - S.rho0);
    if (kDebug) dens[i] = rho;
    uint4* mrow = reinterpret_cast<uint4*>(mask + (size_t)i * kTileMaskWords);
    mrow[0] = make_uint4(smask[0 * kBlock + tid], smask[1 * kBlock + tid], smask[2 * kBlock + tid],
                         smask[3 * kBlock + tid]);
    mrow[1] = make_uint4(smask[4 * kBlock + tid], smask[5 * kBlock + tid], smask[6 * kBlock + tid], 0u);
}

template <bool kStatic, bool kDebug, bool kTri>
__global__ __launch_bounds__(kBlock) void k_force_tile(const BlockDesc* __restrict__ blocks, uint32_t numBlocks,
                                                       const SceneDev* __restrict__ scenes,
                                                       const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                       const uint32_t* __restrict__ tab,
                                                       const float4* __restrict__ spos,
                                                       const uint32_t* __restrict__ stab,
                                                       const uint32_t* __restrict__ mask,
                                                       const uint32_t* __restrict__ ids, const float* __restrict__ tris,
                                                       float dt, float4* __restrict__ posT, float4* __restrict__ velT,
                                                       uint32_t* __restrict__ counts, float4* __restrict__ accOut) {
    __shared__ BucketMap m;
    __shared__ float4 candP[kCapForce];
    __shared__ float4 candV[kCapForce];
    __shared__ uint32_t smask[kTileMaskWords * kBlock];
    __shared__ uint32_t s_code[kBlock];
    __shared__ uint32_t s_w[4];
    const uint32_t tile = xcd_tile(blockIdx.x, numBlocks);
    if (tile >= numBlocks) return;
    const BlockDesc bd = blocks[tile];
    const SceneDev& S = scenes[bd.scene];
    const uint32_t tid = threadIdx.x;
    const uint32_t i = bd.begin + tid;
    const bool valid = i < bd.end;
    if (valid) {
        const uint4* mr = reinterpret_cast<const uint4*>(mask + (size_t)i * kTileMaskWords);
        const uint4 m0 = mr[0], m1 = mr[1];
        smask[0 * kBlock + tid] = m0.x;
        smask[1 * kBlock + tid] = m0.y;
        smask[2 * kBlock + tid] = m0.z;
        smask[3 * kBlock + tid] = m0.w;
        smask[4 * kBlock + tid] = m1.x;
        smask[5 * kBlock + tid] = m1.y;
        smask[6 * kBlock + tid] = m1.z;
    }
    const float4 xi = valid ? pos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 vi = valid ? vel[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float pi = xi.w;  // this step's pressure (density pass)
    const float h = S.h, h2 = S.h2;
    const uint32_t cx = cell_coord(xi.x, h), cy = cell_coord(xi.y, h), cz = cell_coord(xi.z, h);
    const CellSpread cs = cell_spread(xi.x, xi.y, xi.z, h);
    const uint32_t code = morton3(cx, cy, cz);
    const uint32_t* table = tab + S.tBegin;
    const uint32_t lo = S.pBegin, hi = S.pBegin + S.pCount;
    stage_tile<true>(m, s_code, s_w, valid, cs, code, S, table, pos, vel, candP, candV, kCapForce);
    if (!valid) return;

    const float rho = pi / S.k + S.rho0;  // SF:544
    const float absP = fabsf(pi);
    float fpx = 0.f, fpy = 0.f, fpz = 0.f;
    float fvx = 0.f, fvy = 0.f, fvz = 0.f;
    float fsx = 0.f, fsy = 0.f, fsz = 0.f;
    float sumD2s = 0.f;
    // One dynamic candidate (position xj with this step's pressure in w,
    // velocity from vsrc, global index j) of the walk (SF:557-589).
    auto pair = [&](uint32_t j, const float4& xj, const float4* vsrc) {
        float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (j == i || r2 > h2) return;
        const float4 vj = *vsrc;
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
    // The recorded neighbours of one bucket visit, read through psrc/vsrc (LDS
    // copies or HBM), then any candidate past the 8th.
    auto visit_bucket = [&](const float4* psrc, const float4* vsrc, uint32_t beg, uint32_t cnt, uint32_t bits8) {
        for (uint32_t mb = bits8; mb; mb &= mb - 1u) {
            const uint32_t t = (uint32_t)__builtin_ctz(mb);
            pair(beg + t, psrc[t], vsrc + t);
        }
        for (uint32_t t = 8; t < cnt; ++t) pair(beg + t, psrc[t], vsrc + t);
    };
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c) {
                const uint32_t visit = (uint32_t)(a * 9 + b * 3 + c);
                const uint32_t bucket = fastmod(cell_code(cs, a, b, c), S.magicM, S.M);
                const BucketView v = bucket_view(m, bucket, table, lo, hi);
                const uint32_t bits8 = (smask[(visit >> 2) * kBlock + tid] >> (8u * (visit & 3u))) & 0xFFu;
                if (v.off != kNotStaged)
                    visit_bucket(candP + v.off, candV + v.off, v.beg, v.cnt, bits8);
                else
                    visit_bucket(pos + v.beg, vel + v.beg, v.beg, v.cnt, bits8);
            }
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
    // SF:641-645: symplectic Euler, then the key of the new cell.
    const float nvx = vi.x + ax * dt, nvy = vi.y + ay * dt, nvz = vi.z + az * dt;
    const float nx = xi.x + nvx * dt, ny = xi.y + nvy * dt, nz = xi.z + nvz * dt;
    const uint32_t ncode = morton3(cell_coord(nx, h), cell_coord(ny, h), cell_coord(nz, h));
    const uint32_t key = fastmod(ncode, S.magicM, S.M);
    posT[i] = make_float4(nx, ny, nz, pi);
    velT[i] = make_float4(nvx, nvy, nvz, __uint_as_float(key));
    const uint32_t gk = S.tBegin + key;
    const Run run = wave_run(gk);
    if (run.is_head) atomicAdd(&counts[gk], run.len);
}

}  // namespace sph

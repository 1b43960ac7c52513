// Host side of the MI355X SPH step engine and its C ABI (include/sph_hip.h).
//
// The engine mirrors SPH::SimulationEngine (SPHFluidSimulation/include/SPH/Core/
// SimulationEngine.h:35-48): Initialize sorts the scene once on the host, in the
// order a one-thread reference run produces, and uploads it; Update enqueues the
// four step kernels of sph_kernels.hpp per step on the engine's own stream
// (optionally replayed from captured hipGraphs); readers copy the state back in
// original order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sph_hip.h"
#include "sph_kernels.hpp"
#include "sph_tile.hpp"
#include "sph_cellwave.hpp"
#include "sph_wall_lists.hpp"

using namespace sph;

// Staging of sph_read_noalloc: positions (3 N), velocities (3 N) and pressures
// (N) in original order, N <= SPH_NOALLOC_MAX_PARTICLES. Each GPU holds its own
// copy from the module's load until the process ends. Host code never uses this
// name as an address: sph_create resolves each device's copy with
// hipGetSymbolAddress (noalloc_scratch) and the engine passes that pointer.
static __device__ __align__(256) float scratch[7ull * SPH_NOALLOC_MAX_PARTICLES];

namespace {

struct StageEvent {
    int stage;
    hipEvent_t start, stop;
};

}  // namespace

struct sph_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    float* scratch = nullptr;  // this device's copy of the file-scope `scratch` array
    std::string err;
    bool initialized = false;
    uint32_t flags = 0;

    // Sizes.
    uint32_t nScenes = 0;
    uint64_t N = 0, Ns = 0, L = 0, Ls = 0, nTris = 0;
    uint32_t numBlocks = 0, numTiles = 0;
    bool hasStatic = false, hasTris = false;
    std::vector<SceneDev> hScenes;
    std::vector<BlockDesc> hBlocks;
    std::vector<uint32_t> hSceneBlocks;  // nScenes + 1 block offsets
    std::vector<uint32_t> hStaticIds;    // original global wall index per sorted wall slot
    std::vector<float> hTris;

    // Device buffers.
    SceneDev* dScenes = nullptr;
    BlockDesc* dBlocks = nullptr;
    uint32_t* dSceneBlocks = nullptr;
    float4 *pos = nullptr, *vel = nullptr, *posT = nullptr, *velT = nullptr;
    float* pres = nullptr;
    uint32_t* ids[2] = {nullptr, nullptr};
    uint32_t* origOfSlot = nullptr;
    uint32_t* tab[2] = {nullptr, nullptr};
    unsigned long long* status[2] = {nullptr, nullptr};
    uint32_t* ticket = nullptr;
    float4* spos = nullptr;
    uint32_t* stab = nullptr;
    uint32_t* wallMap = nullptr;  // per scene: bit wall_hash(Morton code of c) set for every cell c holding walls
    uint32_t* wallCells = nullptr;  // wall lists (sph_wall_lists.hpp): per-scene cell tables
    uint32_t* wallList = nullptr;   // and the octants' lists of static particle indices
    uint32_t* tileOrder = nullptr;  // cell-wave launch order (tile_order_chunk), numBlocks entries
    uint32_t* heavy = nullptr;      // per block: particles near walls in the last density pass
    float4* wallSum = nullptr;      // cell-wave: the walls' geometric sums of each particle, for the force pass
    float4* wallZ = nullptr;        // and the coincident-pair weight where one occurred
    float4* visc = nullptr;         // cell-wave list mode: the density pass's viscosity sums
    uint64_t wallWords = 0;
    uint32_t* staticScene = nullptr;
    float* tris = nullptr;
    float* dens = nullptr;
    float4* acc = nullptr;
    uint32_t* mask = nullptr;       // candidate mask from the density pass (run, tile and cell-wave walks)
    bool cwList = false;            // cell-wave: neighbour lists (else masks) between the passes
    uint64_t* extraMask = nullptr;  // cell-wave, mask mode: crowded buckets' neighbours
    uint2* off16 = nullptr;         // cell-wave, mixed precision: FP16 offset records (offset_record)
    uint32_t* nbr = nullptr;        // cell-wave, list mode: neighbour lists (kListCap entries per particle)
    uint32_t* nbrCount = nullptr;   // cell-wave: list lengths
    DensityOut* densOut = nullptr;  // cell-wave density pass: its output pointers (device copy)
    uint32_t* nanCnt = nullptr;     // 2 x nScenes: NaN particles in bucket 0, per parity
    // Launch shape of the cell-wave passes: waves per workgroup of the density and
    // force passes; shapeSet once the caller chose it (sph_set_launch_shape).
    uint32_t densWaves = 4, forceWaves = 4;
    bool shapeSet = false;
    double* consPartial = nullptr;
    double* consOut = nullptr;
    uint32_t* valBits = nullptr;
    uint32_t* valIdx = nullptr;  // 4 * nScenes: firstIdx, firstVal, lastIdx, lastVal (dynamic), then static
    // Particle-map mode (SPH_FLAG_PARTICLE_MAP): bucket slot k holds particle pmap[k]
    // while the state keeps its memory order; a reordering step sorts the state and
    // resets the map (SimulationEngineCPU.cpp:280-316, 420-459).
    uint32_t* pmap = nullptr;
    float reorderInterval = FLT_MAX;  // reorderTimeInterval (SimulationEngineCPU.cpp:355)
    float reorderElapsed = 0.f;
    bool reorderNow = false;          // the steps being enqueued reorder
    int cur = 0;                 // ids[cur] / tab[cur] hold the current sorted state
    bool debugValid = false;
    double simTime = 0.0;

    // Asynchronous readback ring: kReadbackSlots pinned host buffers, each
    // [pos 3N | vel 3N | pressure N | key N] in original particle order, filled
    // from one device staging buffer by stream-ordered copies.
    static constexpr uint32_t kReadbackSlots = 3;
    float* rbDev = nullptr;  // staging: the host layout, then N slot ids
    uint64_t rbN = 0;
    void* rbHost[kReadbackSlots] = {nullptr, nullptr, nullptr};
    hipEvent_t rbEvent[kReadbackSlots] = {nullptr, nullptr, nullptr};
    bool rbFilled[kReadbackSlots] = {false, false, false};
    double rbTime[kReadbackSlots] = {0, 0, 0};
    uint32_t rbNext = 0;

    // hipGraph replay: one graph per starting parity and (particle-map mode)
    // reordering or not, index 2 * reorderNow + parity, keyed by dt.
    hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
    float graphDt[4] = {0.f, 0.f, 0.f, 0.f};

    // Profiling.
    uint32_t profiling = 0;  // stage mask, SPH_PROFILE_ROTATE
    uint32_t profStep = 0;   // with SPH_PROFILE_ROTATE: steps profiled so far
    std::vector<StageEvent> pending;
    std::vector<hipEvent_t> eventPool;
    double stageMs[SPH_STAGE_COUNT] = {0, 0, 0, 0};
    uint64_t stageLaunches[SPH_STAGE_COUNT] = {0, 0, 0, 0};
};

#define SPH_HIP(h, expr)                                                                      \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            (h)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                     \
            return SPH_E_HIP;                                                                 \
        }                                                                                     \
    } while (0)

namespace {

template <class T>
int dalloc(sph_engine* h, T** p, uint64_t count) {
    *p = nullptr;
    if (count == 0) return SPH_OK;
    SPH_HIP(h, hipMalloc((void**)p, sizeof(T) * count));
    return SPH_OK;
}

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

void release(sph_engine* h) {
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (auto& g : h->graph) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    for (auto& pe : h->pending) {
        h->eventPool.push_back(pe.start);
        h->eventPool.push_back(pe.stop);
    }
    h->pending.clear();
    dfree(h->dScenes);
    dfree(h->dBlocks);
    dfree(h->dSceneBlocks);
    dfree(h->pos);
    dfree(h->vel);
    dfree(h->posT);
    dfree(h->velT);
    dfree(h->pres);
    dfree(h->ids[0]);
    dfree(h->ids[1]);
    dfree(h->origOfSlot);
    dfree(h->tab[0]);
    dfree(h->tab[1]);
    dfree(h->status[0]);
    dfree(h->status[1]);
    dfree(h->ticket);
    dfree(h->spos);
    dfree(h->stab);
    dfree(h->wallMap);
    dfree(h->wallCells);
    dfree(h->wallList);
    dfree(h->tileOrder);
    dfree(h->heavy);
    dfree(h->wallSum);
    dfree(h->wallZ);
    dfree(h->visc);
    h->wallWords = 0;
    dfree(h->staticScene);
    dfree(h->tris);
    dfree(h->dens);
    dfree(h->acc);
    dfree(h->mask);
    dfree(h->nbr);
    dfree(h->extraMask);
    dfree(h->off16);
    dfree(h->nbrCount);
    dfree(h->densOut);
    dfree(h->nanCnt);
    dfree(h->consPartial);
    dfree(h->consOut);
    dfree(h->valBits);
    dfree(h->valIdx);
    dfree(h->pmap);
    h->reorderInterval = FLT_MAX;  // SimulationEngineCPU.cpp:389-390
    h->reorderElapsed = 0.f;
    h->reorderNow = false;
    dfree(h->rbDev);
    for (uint32_t k = 0; k < sph_engine::kReadbackSlots; ++k) {
        if (h->rbHost[k]) (void)hipHostFree(h->rbHost[k]);
        h->rbHost[k] = nullptr;
        if (h->rbEvent[k]) (void)hipEventDestroy(h->rbEvent[k]);
        h->rbEvent[k] = nullptr;
        h->rbFilled[k] = false;
    }
    h->rbN = 0;
    h->rbNext = 0;
    h->hScenes.clear();
    h->hBlocks.clear();
    h->hSceneBlocks.clear();
    h->hStaticIds.clear();
    h->hTris.clear();
    h->N = h->Ns = h->L = h->Ls = h->nTris = 0;
    h->nScenes = 0;
    h->numBlocks = h->numTiles = 0;
    h->hasStatic = h->hasTris = false;
    h->cur = 0;
    h->simTime = 0.0;
    h->debugValid = false;
    h->initialized = false;
}

// Reference counting sort of one scene's keys, single-threaded, as
// SimulationEngineCPU.cpp:55-131 with one thread (and SimulationEngine.h:79-100
// for walls): histogram, inclusive scan, then a pre-decrement scatter in input
// order, so equal keys end in descending input order and table[b] is the start
// of bucket b. Returns slot[i] for input i.
void counting_sort(const std::vector<uint32_t>& keys, uint64_t M, std::vector<uint32_t>& table,
                   std::vector<uint32_t>& slotOf) {
    const uint64_t n = keys.size();
    table.assign(M + 1, 0u);
    for (uint64_t i = 0; i < n; ++i) ++table[keys[i]];
    uint32_t sum = 0;
    for (uint64_t b = 0; b <= M; ++b) {
        sum += table[b];
        table[b] = sum;
    }
    slotOf.resize(n);
    for (uint64_t i = 0; i < n; ++i) slotOf[i] = --table[keys[i]];
}

// Bins of one scene's triangles for mesh_force. The grid covers the triangles'
// bounding boxes widened by the force's range (SF:205, 2.5) plus a margin for the
// rounding of the device's distance and bin index; bins have an edge of half the
// range (larger when that would make more than 2^21 bins), and each lists, in
// index order, the triangles whose widened box meets it. Appends the bin-start
// table and the lists to `tables` and records the grid in S (offsets relative to
// `tableBase`, the float offset the tables will have in the mesh buffer).
void build_mesh_bins(const float* tri9, uint64_t count, SceneDev& S, std::vector<uint32_t>& tables,
                     uint64_t tableBase) {
    S.nbx = S.nby = S.nbz = 0;
    if (count == 0) return;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, maxAbs = 0.0;
    for (uint64_t k = 0; k < 9 * count; ++k) {
        const double v = tri9[k];
        lo[k % 3] = std::min(lo[k % 3], v);
        hi[k % 3] = std::max(hi[k % 3], v);
        maxAbs = std::max(maxAbs, std::fabs(v));
    }
    const double widen = 2.5 * 1.001 + 1e-4 * (1.0 + maxAbs);
    double edge = 1.25;
    uint64_t n[3];
    for (;;) {
        uint64_t total = 1;
        for (int a = 0; a < 3; ++a) {
            n[a] = std::max<uint64_t>(1, (uint64_t)std::ceil((hi[a] - lo[a] + 2.0 * widen) / edge));
            total *= n[a];
        }
        if (total <= (1ull << 21)) break;
        edge *= 1.25;
    }
    const double o[3] = {lo[0] - widen, lo[1] - widen, lo[2] - widen};
    S.bx0 = (float)o[0];
    S.by0 = (float)o[1];
    S.bz0 = (float)o[2];
    S.binInv = (float)(1.0 / edge);
    S.nbx = (uint32_t)n[0];
    S.nby = (uint32_t)n[1];
    S.nbz = (uint32_t)n[2];
    const uint64_t nb = n[0] * n[1] * n[2];
    // The bins each triangle's widened box meets, in the float32 grid the device
    // uses (origin and edge as stored), so the two agree on every boundary.
    const double fo[3] = {S.bx0, S.by0, S.bz0}, finv = S.binInv;
    auto range = [&](uint64_t t, int a, uint64_t& b0, uint64_t& b1) {
        double mn = 1e300, mx = -1e300;
        for (int v = 0; v < 3; ++v) {
            mn = std::min(mn, (double)tri9[9 * t + 3 * v + a]);
            mx = std::max(mx, (double)tri9[9 * t + 3 * v + a]);
        }
        const double g0 = std::floor((mn - widen - fo[a]) * finv), g1 = std::floor((mx + widen - fo[a]) * finv);
        b0 = (uint64_t)std::max(0.0, g0);
        b1 = (uint64_t)std::min((double)(n[a] - 1), std::max(0.0, g1));
    };
    std::vector<uint32_t> cnt(nb + 1, 0u);
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<uint32_t> fill;
        if (pass == 1) {
            uint32_t sum = 0;
            for (uint64_t b = 0; b <= nb; ++b) {
                const uint32_t c = cnt[b];
                cnt[b] = sum;
                sum += c;
            }
            fill.assign(cnt.begin(), cnt.end() - 1);
            S.binStart = (uint32_t)(tableBase + tables.size());
            tables.insert(tables.end(), cnt.begin(), cnt.end());
            S.binList = (uint32_t)(tableBase + tables.size());
            tables.resize(tables.size() + cnt[nb]);
        }
        for (uint64_t t = 0; t < count; ++t) {
            uint64_t x0, x1, y0, y1, z0, z1;
            range(t, 0, x0, x1);
            range(t, 1, y0, y1);
            range(t, 2, z0, z1);
            for (uint64_t z = z0; z <= z1; ++z)
                for (uint64_t y = y0; y <= y1; ++y)
                    for (uint64_t x = x0; x <= x1; ++x) {
                        const uint64_t b = (z * n[1] + y) * n[0] + x;
                        if (pass == 0)
                            ++cnt[b];
                        else
                            tables[S.binList - tableBase + fill[b]++] = (uint32_t)t;
                    }
        }
    }
}

uint32_t key_of(const float* p, float h, uint64_t M) {
    const uint32_t code = morton3(cell_coord(p[0], h), cell_coord(p[1], h), cell_coord(p[2], h));
    return (uint32_t)(code % M);
}

hipEvent_t take_event(sph_engine* h) {
    if (!h->eventPool.empty()) {
        hipEvent_t e = h->eventPool.back();
        h->eventPool.pop_back();
        return e;
    }
    // Timing only: no system-scope fence when the event is recorded (that fence writes
    // back and invalidates the caches, which costs the stream time and slows the next
    // launch), so the events change what they time as little as possible.
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    return e;
}

// The engine's device staging buffer for readbacks (9 N floats: positions,
// velocities, pressures and keys in original order, then N slot ids), allocated
// on the first readback after initialize and kept: the synchronous readers and
// the asynchronous ring share it, all in stream order on the engine's stream.
int ensure_staging(sph_engine* h) {
    if (h->rbDev || h->N == 0) return SPH_OK;
    SPH_HIP(h, hipMalloc((void**)&h->rbDev, sizeof(float) * 9 * h->N));
    return SPH_OK;
}

// Folds finished stage events into the totals.
int collect_events(sph_engine* h) {
    for (auto& pe : h->pending) {
        SPH_HIP(h, hipEventSynchronize(pe.stop));
        float ms = 0.f;
        SPH_HIP(h, hipEventElapsedTime(&ms, pe.start, pe.stop));
        h->stageMs[pe.stage] += ms;
        h->stageLaunches[pe.stage] += 1;
        h->eventPool.push_back(pe.start);
        h->eventPool.push_back(pe.stop);
    }
    h->pending.clear();
    return SPH_OK;
}

// Neighbour-pass variant chosen by the initialize flags.
enum class Walk { CellWave, Runs, Tiles, Cells };

Walk walk_of(uint32_t flags) {
    if (flags & (SPH_FLAG_CELL_WALK | SPH_FLAG_PARTICLE_MAP)) return Walk::Cells;
    if (flags & SPH_FLAG_RUN_WALK) return Walk::Runs;
    if (flags & SPH_FLAG_TILE_WALK) return Walk::Tiles;
    return Walk::CellWave;
}

// Enqueues one step starting from parity h->cur (does not flip it).
// Mixed precision: the offset records of the current state (the scatter writes
// them every step; this covers a state set otherwise). On the engine's stream.
int refresh_offsets(sph_engine* h) {
    if (!h->off16 || h->N == 0) return SPH_OK;
    hipLaunchKernelGGL(k_offsets, dim3(h->numBlocks), dim3(kBlock), 0, h->stream, h->dBlocks, h->dScenes, h->pos,
                       h->off16);
    SPH_HIP(h, hipGetLastError());
    return SPH_OK;
}

int enqueue_step(sph_engine* h, float dt, bool capture) {
    const int c = h->cur;
    const bool dbg = (h->flags & SPH_FLAG_RECORD_DEBUG) != 0;
    uint32_t prof = capture ? 0u : h->profiling & SPH_PROFILE_ALL;
    if (prof && (h->profiling & SPH_PROFILE_ROTATE)) {
        // One masked stage per step, in turn.
        uint32_t k = h->profStep++ % (uint32_t)__builtin_popcount(prof);
        for (; k; --k) prof &= prof - 1u;
        prof &= ~prof + 1u;
    }
    const dim3 grid(h->numBlocks), block(kBlock);
    const dim3 gridX(((h->numBlocks + 7u) / 8u) * 8u);  // padded for the XCD-aware tile order
    // Cell-wave passes: workgroups of W waves, each wave a 64-particle slice of a block.
    auto wgs = [&](uint32_t W) { return (h->numBlocks * kWaves + W - 1u) / W; };
    const uint32_t densWG = wgs(h->densWaves), forceWG = wgs(h->forceWaves);
    const dim3 gridD(((densWG + 7u) / 8u) * 8u), blockD(64u * h->densWaves);
    // The force pass of a single scene takes its tiles with the XCDs interleaved,
    // kXcdRunsPerXcd runs on each (xcd_tile_runs): its work per tile varies along
    // the particle range, and contiguous eighths left the XCDs unevenly loaded. A
    // batch keeps each XCD on contiguous scenes, whose data then stays in that
    // XCD's L2.
    const uint32_t forceRun =
        h->nScenes == 1 ? std::max(1u, (forceWG + 8u * kXcdRunsPerXcd - 1u) / (8u * kXcdRunsPerXcd)) : 0u;
    const uint32_t padF = forceRun ? 8u * forceRun : 8u;
    const dim3 gridF(((forceWG + padF - 1u) / padF) * padF), blockF(64u * h->forceWaves);
    const size_t ldsD = sizeof(WaveTables) * h->densWaves, ldsF = sizeof(WaveTables) * h->forceWaves;
    StageEvent ev{};
    bool timing = false;
    auto begin = [&](int stage) -> int {
        timing = (prof >> stage) & 1u;
        if (!timing) return SPH_OK;
        ev.stage = stage;
        ev.start = take_event(h);
        ev.stop = take_event(h);
        if (!ev.start || !ev.stop) {
            h->err = "hipEventCreate failed";
            return SPH_E_HIP;
        }
        SPH_HIP(h, hipEventRecord(ev.start, h->stream));
        return SPH_OK;
    };
    auto end = [&]() -> int {
        if (!timing) return SPH_OK;
        SPH_HIP(h, hipEventRecord(ev.stop, h->stream));
        h->pending.push_back(ev);
        return SPH_OK;
    };
    int rc;
    // Density / pressure.
    if ((rc = begin(SPH_STAGE_DENSITY))) return rc;
    const Walk walk = walk_of(h->flags);
    const bool mixed = (h->flags & (SPH_FLAG_FP16_OFFSETS | SPH_FLAG_FP16_STORED)) != 0;  // cell-wave walk only
    const bool stored = (h->flags & SPH_FLAG_FP16_STORED) != 0;
    // Per-scene counts of NaN particles in bucket 0: read this step, accumulated for the next.
    const uint32_t* nanNow = h->nanCnt + (size_t)c * h->nScenes;
    uint32_t* nanNext = h->nanCnt + (size_t)(c ^ 1) * h->nScenes;
    if (walk != Walk::Cells) {
#define SPH_DENS_CW(S, D, X, L, T)                                                                             \
        hipLaunchKernelGGL((k_density_cw<S, D, X, L, T>), gridD, blockD, ldsD, h->stream, h->dBlocks,         \
                           h->numBlocks, h->dScenes, h->pos, h->vel, h->tab[c], h->spos, h->stab, h->wallMap,  \
                           h->tileOrder, h->nbr, nanNow, h->densOut, h->off16, densWG)
#define SPH_DENS(S, D)                                                                                         \
    if (walk == Walk::CellWave && stored && h->cwList)                                                         \
        SPH_DENS_CW(S, D, true, true, true);                                                                   \
    else if (walk == Walk::CellWave && stored)                                                                 \
        SPH_DENS_CW(S, D, true, false, true);                                                                  \
    else if (walk == Walk::CellWave && mixed && h->cwList)                                                     \
        SPH_DENS_CW(S, D, true, true, false);                                                                  \
    else if (walk == Walk::CellWave && mixed)                                                                  \
        SPH_DENS_CW(S, D, true, false, false);                                                                 \
    else if (walk == Walk::CellWave && h->cwList)                                                              \
        SPH_DENS_CW(S, D, false, true, false);                                                                 \
    else if (walk == Walk::CellWave)                                                                           \
        SPH_DENS_CW(S, D, false, false, false);                                                                \
    else if (walk == Walk::Tiles)                                                                              \
        hipLaunchKernelGGL((k_density_tile<S, D>), gridX, block, 0, h->stream, h->dBlocks, h->numBlocks,       \
                           h->dScenes, h->pos, h->tab[c], h->spos, h->stab, h->mask, h->dens);                 \
    else                                                                                                       \
        hipLaunchKernelGGL((k_density_runs<S, D>), gridX, block, 0, h->stream, h->dBlocks, h->numBlocks,       \
                           h->dScenes, h->pos, h->tab[c], h->spos, h->stab, h->mask, h->dens)
        if (h->hasStatic) {
            if (dbg) {
                SPH_DENS(true, true);
            } else {
                SPH_DENS(true, false);
            }
        } else {
            if (dbg) {
                SPH_DENS(false, true);
            } else {
                SPH_DENS(false, false);
            }
        }
#undef SPH_DENS
#undef SPH_DENS_CW
    } else {
#define SPH_DENS_CELL(S, D, P)                                                                                   \
    hipLaunchKernelGGL((k_density<S, D, P>), grid, block, 0, h->stream, h->dBlocks, h->dScenes, h->pos, h->tab[c], \
                       h->spos, h->stab, h->pmap, h->pres, h->dens)
        const bool map = h->pmap != nullptr;
        if (h->hasStatic) {
            if (dbg) {
                if (map) SPH_DENS_CELL(true, true, true); else SPH_DENS_CELL(true, true, false);
            } else {
                if (map) SPH_DENS_CELL(true, false, true); else SPH_DENS_CELL(true, false, false);
            }
        } else {
            if (dbg) {
                if (map) SPH_DENS_CELL(false, true, true); else SPH_DENS_CELL(false, true, false);
            } else {
                if (map) SPH_DENS_CELL(false, false, true); else SPH_DENS_CELL(false, false, false);
            }
        }
#undef SPH_DENS_CELL
    }
    SPH_HIP(h, hipGetLastError());
    if ((rc = end())) return rc;

    // Forces, integration, new keys, histogram into the other table.
    if ((rc = begin(SPH_STAGE_FORCE))) return rc;
#define SPH_FORCE_LIST(S, D, T, X)                                                                                \
        hipLaunchKernelGGL((k_force_cw<S, D, T, X>), gridF, blockF, 0, h->stream, h->dBlocks, h->numBlocks,      \
                           h->dScenes, h->pos, h->vel, h->tab[c], h->wallSum, h->wallZ, h->visc, h->nbr,          \
                           h->nbrCount, nanNext, h->ids[c], h->tris, dt, h->posT, h->velT, h->tab[c ^ 1], h->acc, \
                           forceWG, forceRun)
#define SPH_FORCE_MASK(S, D, T, X)                                                                                \
        hipLaunchKernelGGL((k_force_cw_mask<S, D, T, X>), gridF, blockF, ldsF, h->stream, h->dBlocks,            \
                           h->numBlocks, h->dScenes, h->pos, h->vel, h->tab[c], h->wallSum, h->wallZ, h->visc,    \
                           h->mask, h->extraMask, nanNow, nanNext, h->ids[c], h->tris, dt, h->posT, h->velT,      \
                           h->tab[c ^ 1], h->acc, forceWG, forceRun)
#define SPH_FORCE(S, D, T)                                                                                        \
    do {                                                                                                          \
    if (walk == Walk::CellWave && mixed && h->cwList)                                                             \
        SPH_FORCE_LIST(S, D, T, true);                                                                            \
    else if (walk == Walk::CellWave && mixed)                                                                     \
        SPH_FORCE_MASK(S, D, T, true);                                                                            \
    else if (walk == Walk::CellWave && h->cwList)                                                                 \
        SPH_FORCE_LIST(S, D, T, false);                                                                           \
    else if (walk == Walk::CellWave)                                                                              \
        SPH_FORCE_MASK(S, D, T, false);                                                                           \
    else if (walk == Walk::Tiles)                                                                                 \
        hipLaunchKernelGGL((k_force_tile<S, D, T>), gridX, block, 0, h->stream, h->dBlocks, h->numBlocks,         \
                           h->dScenes, h->pos, h->vel, h->tab[c], h->spos, h->stab, h->mask, h->ids[c], h->tris,  \
                           dt, h->posT, h->velT, h->tab[c ^ 1], h->acc);                                          \
    else if (walk == Walk::Runs)                                                                                  \
        hipLaunchKernelGGL((k_force_runs<S, D, T>), gridX, block, 0, h->stream, h->dBlocks, h->numBlocks,        \
                           h->dScenes, h->pos, h->vel, h->tab[c], h->spos, h->stab, h->mask, h->ids[c], h->tris,  \
                           dt, h->posT, h->velT, h->tab[c ^ 1], h->acc);                                          \
    else if (h->pmap)                                                                                             \
        hipLaunchKernelGGL((k_force<S, D, T, true>), grid, block, 0, h->stream, h->dBlocks, h->dScenes, h->pos,   \
                           h->vel, h->pres, h->tab[c], h->spos, h->stab, h->pmap, h->ids[c], h->tris, dt, h->posT, \
                           h->velT, h->tab[c ^ 1], h->acc);                                                       \
    else                                                                                                          \
        hipLaunchKernelGGL((k_force<S, D, T, false>), grid, block, 0, h->stream, h->dBlocks, h->dScenes, h->pos,  \
                           h->vel, h->pres, h->tab[c], h->spos, h->stab, nullptr, h->ids[c], h->tris, dt, h->posT,\
                           h->velT, h->tab[c ^ 1], h->acc);                                                       \
    } while (0)
    if (h->hasStatic) {
        if (dbg) {
            if (h->hasTris) SPH_FORCE(true, true, true); else SPH_FORCE(true, true, false);
        } else {
            if (h->hasTris) SPH_FORCE(true, false, true); else SPH_FORCE(true, false, false);
        }
    } else {
        if (dbg) {
            if (h->hasTris) SPH_FORCE(false, true, true); else SPH_FORCE(false, true, false);
        } else {
            if (h->hasTris) SPH_FORCE(false, false, true); else SPH_FORCE(false, false, false);
        }
    }
#undef SPH_FORCE
#undef SPH_FORCE_LIST
#undef SPH_FORCE_MASK
    SPH_HIP(h, hipGetLastError());
    if ((rc = end())) return rc;

    // Scan the histogram; clear the table read this step for the next histogram.
    if ((rc = begin(SPH_STAGE_SCAN))) return rc;
    hipLaunchKernelGGL(k_scan, dim3(h->numTiles), dim3(kScanThreads), 0, h->stream, h->tab[c ^ 1], h->tab[c],
                       (uint32_t)h->L, h->status[c], h->status[c ^ 1], h->ticket, h->numTiles,
                       h->nanCnt + (size_t)c * h->nScenes, h->nScenes);
    SPH_HIP(h, hipGetLastError());
    if ((rc = end())) return rc;

    // Move the state into bucket order.
    if ((rc = begin(SPH_STAGE_SCATTER))) return rc;
    // With the cell-wave walk and walls, the same launch also computes the next
    // step's launch order (workgroups near walls first) on eight more workgroups.
    if (walk == Walk::CellWave && h->hasStatic) {
        hipLaunchKernelGGL(k_scatter_order, dim3(h->numBlocks + 8u), block, 0, h->stream, h->dBlocks, h->numBlocks,
                           h->dScenes, h->posT, h->velT, h->ids[c], h->tab[c ^ 1], h->pos, h->vel, h->ids[c ^ 1],
                           h->heavy, h->tileOrder, h->off16);
    } else if (h->pmap && !h->reorderNow) {
        // Particle-map mode between reorders: memory order kept, map refilled.
        hipLaunchKernelGGL(k_fill_map, grid, block, 0, h->stream, h->dBlocks, h->dScenes, h->posT, h->velT,
                           h->ids[c], h->tab[c ^ 1], h->pos, h->vel, h->ids[c ^ 1], h->pmap);
    } else {
        hipLaunchKernelGGL(k_scatter, grid, block, 0, h->stream, h->dBlocks, h->dScenes, h->posT, h->velT,
                           h->ids[c], h->tab[c ^ 1], h->pos, h->vel, h->ids[c ^ 1], h->pmap, h->off16);
    }
    SPH_HIP(h, hipGetLastError());
    if ((rc = end())) return rc;
    return SPH_OK;
}

int capture_graph(sph_engine* h, int parity, float dt) {
    const int gi = 2 * (h->reorderNow ? 1 : 0) + parity;
    if (h->graph[gi]) {
        SPH_HIP(h, hipGraphExecDestroy(h->graph[gi]));
        h->graph[gi] = nullptr;
    }
    const int saved = h->cur;
    h->cur = parity;
    SPH_HIP(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step(h, dt, true);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    h->cur = saved;
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) {
        h->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
        return SPH_E_HIP;
    }
    e = hipGraphInstantiate(&h->graph[gi], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        h->err = std::string("hipGraphInstantiate: ") + hipGetErrorString(e);
        h->graph[gi] = nullptr;
        return SPH_E_HIP;
    }
    h->graphDt[gi] = dt;
    return SPH_OK;
}

bool valid(sph_handle h) { return h != nullptr; }

// Device address of each GPU's copy of `scratch`, resolved once per device with
// that device current.
constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
float* gScratch[kMaxDevices] = {};

int noalloc_scratch(sph_engine* h) {
    if (h->device >= kMaxDevices) {
        h->err = "device ordinal above the library's table of per-device arrays";
        return SPH_E_INVALID;
    }
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!gScratch[h->device]) {
        void* p = nullptr;
        SPH_HIP(h, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        gScratch[h->device] = static_cast<float*>(p);
    }
    h->scratch = gScratch[h->device];
    return SPH_OK;
}

}  // namespace

extern "C" {

const char* sph_system_implementation_name(void) { return "HIP"; }

int sph_create(sph_handle* out, int device) {
    if (!out) return SPH_E_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return SPH_E_NO_DEVICE;
    if (device < 0 || device >= count) return SPH_E_INVALID;
    sph_engine* h = new sph_engine;
    h->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        if (h->stream) (void)hipStreamDestroy(h->stream);
        delete h;
        return SPH_E_HIP;
    }
    const int rc = noalloc_scratch(h);
    if (rc) {
        (void)hipStreamDestroy(h->stream);
        delete h;
        return rc;
    }
    *out = h;
    return SPH_OK;
}

int sph_destroy(sph_handle h) {
    if (!valid(h)) return SPH_E_INVALID;
    (void)hipSetDevice(h->device);
    release(h);
    for (auto e : h->eventPool) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return SPH_OK;
}

int sph_clear(sph_handle h) {
    if (!valid(h)) return SPH_E_INVALID;
    SPH_HIP(h, hipSetDevice(h->device));
    release(h);
    return SPH_OK;
}

const char* sph_last_error(sph_handle h) { return h ? h->err.c_str() : "invalid handle"; }

int sph_initialize(sph_handle h, const sph_params* params, const float* dyn_pos, const float* dyn_vel, uint64_t n,
                   const float* static_pos, uint64_t n_s, uint32_t flags) {
    const uint64_t dOff[2] = {0, n};
    const uint64_t sOff[2] = {0, n_s};
    return sph_batch_initialize(h, 1, params, dOff, dyn_pos, dyn_vel, sOff, static_pos, flags);
}

int sph_batch_initialize(sph_handle h, uint32_t n_scenes, const sph_params* params, const uint64_t* dyn_offsets,
                         const float* dyn_pos, const float* dyn_vel, const uint64_t* static_offsets,
                         const float* static_pos, uint32_t flags) {
    if (!valid(h)) return SPH_E_INVALID;
    if (n_scenes == 0 || !params || !dyn_offsets || !static_offsets) {
        h->err = "sph_batch_initialize: no scenes or missing offsets/params";
        return SPH_E_INVALID;
    }
    SPH_HIP(h, hipSetDevice(h->device));
    release(h);  // Initialize starts with Clear (SimulationEngineCPU.cpp:398)
    h->err.clear();
    h->flags = flags;
    const uint64_t N = dyn_offsets[n_scenes], Ns = static_offsets[n_scenes];
    if ((N && !dyn_pos) || (Ns && !static_pos)) {
        h->err = "sph_batch_initialize: missing particle arrays";
        return SPH_E_INVALID;
    }
    for (uint32_t s = 0; s < n_scenes; ++s) {
        if (dyn_offsets[s + 1] < dyn_offsets[s] || static_offsets[s + 1] < static_offsets[s] ||
            !(params[s].maxInteractionDistance > 0.f) || params[s].gasConstant == 0.f) {
            h->err = "sph_batch_initialize: bad offsets or parameters for a scene";
            return SPH_E_INVALID;
        }
    }
    uint64_t L = 0;
    for (uint32_t s = 0; s < n_scenes; ++s) L += 2 * (dyn_offsets[s + 1] - dyn_offsets[s]) + 1;
    if (N >= (1ull << 28) || L >= (1ull << 32) - kScanTile || Ns >= (1ull << 28)) {
        h->err = "sph_batch_initialize: more than 2^28 - 1 dynamic or static particles";
        return SPH_E_INVALID;
    }
    h->nScenes = n_scenes;
    h->N = N;
    h->Ns = Ns;
    h->L = L;

    std::vector<float4> hPos(N), hVel(N), hSpos(Ns);
    std::vector<uint32_t> hIds(N), hOrig(N), hTab(L), hStab, hStaticScene(Ns), hWallMap, hNan(n_scenes, 0u);
    std::vector<uint32_t> hWallCells, hWallList;
    h->hStaticIds.assign(Ns, 0);
    h->hScenes.resize(n_scenes);
    h->hSceneBlocks.assign(n_scenes + 1, 0);
    h->hBlocks.clear();
    uint64_t tBegin = 0, stBegin = 0;
    std::vector<uint32_t> keys, table, slotOf;
    for (uint32_t s = 0; s < n_scenes; ++s) {
        const sph_params& P = params[s];
        SceneDev& S = h->hScenes[s];
        memset(&S, 0, sizeof(S));
        S.h = P.maxInteractionDistance;
        S.h2 = S.h * S.h;
        S.mass = P.particleMass;
        S.k = P.gasConstant;
        S.rho0 = P.restDensity;
        S.mu = P.viscosity;
        S.gx = P.gravityX;
        S.gy = P.gravityY;
        S.gz = P.gravityZ;
        S.C = kernel_constant(S.h);                     // SimulationEngineCPU.cpp:406
        S.selfDensity = S.mass * d0_at_zero(S.h) * S.C;  // SimulationEngineCPU.cpp:407
        S.h2wide = S.h2 * 1.0001f;
        const uint64_t p0 = dyn_offsets[s], pn = dyn_offsets[s + 1] - p0;
        const uint64_t s0 = static_offsets[s], sn = static_offsets[s + 1] - s0;
        S.pBegin = (uint32_t)p0;
        S.pCount = (uint32_t)pn;
        S.tBegin = (uint32_t)tBegin;
        S.M = (uint32_t)(2 * pn);
        S.magicM = pn ? fastmod_magic(S.M) : 0;
        S.sBegin = (uint32_t)s0;
        S.sCount = (uint32_t)sn;
        S.stBegin = (uint32_t)stBegin;
        S.Ms = (uint32_t)sn;
        S.magicMs = sn ? fastmod_magic(S.Ms) : 0;
        // Walls: SimulationEngine.h:79-114, M_s = N_s (SimulationEngineCPU.cpp:468).
        if (sn) {
            keys.resize(sn);
            for (uint64_t i = 0; i < sn; ++i) keys[i] = key_of(static_pos + 3 * (s0 + i), S.h, S.Ms);
            counting_sort(keys, S.Ms, table, slotOf);
            for (uint64_t i = 0; i < sn; ++i) {
                const float* q = static_pos + 3 * (s0 + i);
                hSpos[s0 + slotOf[i]] = make_float4(q[0], q[1], q[2], 0.f);
                h->hStaticIds[s0 + slotOf[i]] = (uint32_t)(s0 + i);
                hStaticScene[s0 + slotOf[i]] = s;
            }
            for (uint64_t b = 0; b <= sn; ++b) hStab.push_back((uint32_t)(s0 + table[b]));
            stBegin += sn + 1;
            // Wall-cell bitmap, 32 bits per wall particle rounded up to a power of
            // two in words: the cell-wave passes skip the static buckets of
            // neighbour cells that cannot hold a wall within h.
            uint64_t words = 32;
            while (words < sn) words *= 2;
            S.wmBegin = (uint32_t)hWallMap.size();
            S.wmMask = (uint32_t)(words - 1);
            hWallMap.resize(hWallMap.size() + words, 0u);
            for (uint64_t i = 0; i < sn; ++i) {
                const float* q = static_pos + 3 * (s0 + i);
                const uint32_t hc = wall_hash(morton3(cell_coord(q[0], S.h), cell_coord(q[1], S.h), cell_coord(q[2], S.h)));
                hWallMap[S.wmBegin + ((hc >> 5) & S.wmMask)] |= 1u << (hc & 31u);
            }
            build_wall_lists(S, hSpos.data(), hStab.data() + S.stBegin, hWallCells, hWallList);
        }
        // Dynamic: M = 2N (SimulationEngineCPU.cpp:486-487), one-thread order.
        if (pn) {
            keys.resize(pn);
            for (uint64_t i = 0; i < pn; ++i) keys[i] = key_of(dyn_pos + 3 * (p0 + i), S.h, S.M);
            counting_sort(keys, S.M, table, slotOf);
            for (uint64_t i = 0; i < pn; ++i) {
                const uint64_t g = p0 + slotOf[i];
                const float* q = dyn_pos + 3 * (p0 + i);
                hPos[g] = make_float4(q[0], q[1], q[2], 0.f);
                const float* w = dyn_vel ? dyn_vel + 3 * (p0 + i) : nullptr;
                float keyBits;
                memcpy(&keyBits, &keys[i], sizeof(float));
                hVel[g] = make_float4(w ? w[0] : 0.f, w ? w[1] : 0.f, w ? w[2] : 0.f, keyBits);
                hIds[g] = (uint32_t)g;
                hOrig[g] = (uint32_t)(p0 + i);
            }
            for (uint64_t b = 0; b <= S.M; ++b) hTab[tBegin + b] = (uint32_t)(p0 + table[b]);
            for (uint64_t i = 0; i < pn; ++i) {
                const float* q = dyn_pos + 3 * (p0 + i);
                if (keys[i] == 0u && (std::isnan(q[0]) || std::isnan(q[1]) || std::isnan(q[2]))) ++hNan[s];
            }
        } else {
            hTab[tBegin] = (uint32_t)p0;
        }
        tBegin += 2 * pn + 1;
        // Workgroups of this scene.
        h->hSceneBlocks[s] = (uint32_t)h->hBlocks.size();
        for (uint64_t b = 0; b < pn; b += kBlock) {
            BlockDesc bd;
            bd.scene = s;
            bd.begin = (uint32_t)(p0 + b);
            bd.end = (uint32_t)(p0 + std::min<uint64_t>(pn, b + kBlock));
            bd.pad = 0;
            h->hBlocks.push_back(bd);
        }
    }
    h->hSceneBlocks[n_scenes] = (uint32_t)h->hBlocks.size();
    h->Ls = hStab.size();
    h->numBlocks = (uint32_t)h->hBlocks.size();
    h->numTiles = (uint32_t)((L + kScanTile - 1) / kScanTile);
    h->hasStatic = Ns > 0;

    int rc;
#define SPH_ALLOC(p, n)                       \
    if ((rc = dalloc(h, &(p), (n)))) return rc;
    SPH_ALLOC(h->dScenes, n_scenes);
    SPH_ALLOC(h->dBlocks, std::max<uint64_t>(h->numBlocks, 1));
    SPH_ALLOC(h->dSceneBlocks, n_scenes + 1);
    SPH_ALLOC(h->pos, std::max<uint64_t>(N, 1));
    SPH_ALLOC(h->vel, std::max<uint64_t>(N, 1));
    SPH_ALLOC(h->posT, std::max<uint64_t>(N, 1));
    SPH_ALLOC(h->velT, std::max<uint64_t>(N, 1));
    SPH_ALLOC(h->pres, std::max<uint64_t>(N, 1));
    SPH_ALLOC(h->ids[0], std::max<uint64_t>(N, 1));
    SPH_ALLOC(h->ids[1], std::max<uint64_t>(N, 1));
    SPH_ALLOC(h->origOfSlot, std::max<uint64_t>(N, 1));
    SPH_ALLOC(h->tab[0], L);
    SPH_ALLOC(h->tab[1], L);
    SPH_ALLOC(h->status[0], h->numTiles);
    SPH_ALLOC(h->status[1], h->numTiles);
    SPH_ALLOC(h->ticket, 1);
    SPH_ALLOC(h->spos, std::max<uint64_t>(Ns, 1));
    SPH_ALLOC(h->stab, std::max<uint64_t>(h->Ls, 1));
    h->wallWords = hWallMap.size();
    SPH_ALLOC(h->wallMap, std::max<uint64_t>(h->wallWords, 1));
    SPH_ALLOC(h->wallCells, std::max<uint64_t>(hWallCells.size(), 1));
    SPH_ALLOC(h->wallList, std::max<uint64_t>(hWallList.size(), 1));
    SPH_ALLOC(h->tileOrder, std::max<uint64_t>(h->numBlocks, 1));
    SPH_ALLOC(h->heavy, std::max<uint64_t>(h->numBlocks, 1));
    if (Ns) {
        SPH_ALLOC(h->wallSum, std::max<uint64_t>(N, 1));
        SPH_ALLOC(h->wallZ, std::max<uint64_t>(N, 1));
    }
    SPH_ALLOC(h->staticScene, std::max<uint64_t>(Ns, 1));
    SPH_ALLOC(h->nanCnt, 2ull * n_scenes);
    SPH_ALLOC(h->consPartial, 8ull * std::max<uint32_t>(h->numBlocks, 1));
    SPH_ALLOC(h->consOut, 8ull * n_scenes);
    SPH_ALLOC(h->valBits, 1);
    SPH_ALLOC(h->valIdx, 8ull * n_scenes);
    static_assert(kMaskWords == kTileMaskWords && kMaskWords == kCwMaskWords,
                  "every mask format uses 8 words per particle");
    if (flags & SPH_FLAG_PARTICLE_MAP) SPH_ALLOC(h->pmap, std::max<uint64_t>(N, 1));
    if (walk_of(flags) != Walk::Cells) {
        SPH_ALLOC(h->mask, kMaskWords * std::max<uint64_t>(N, 1));
        if (kViscLists) SPH_ALLOC(h->visc, std::max<uint64_t>(N, 1));
        // Lists while they stay in the caches (4 B per neighbour), else masks (32 B per particle).
        h->cwList = (flags & SPH_FLAG_CW_LISTS) || (!(flags & SPH_FLAG_CW_MASKS) && N <= SPH_CW_LIST_MAX);
        if (!h->shapeSet) {
            // Measured on MI355X: over lists 4 + 4 waves per workgroup; over masks the
            // force pass (which rebuilds the cell groups) gains 3% from 8 waves, the
            // density pass 1% from 2 (dam-break-8m).
            h->densWaves = h->cwList ? 4u : 2u;
            h->forceWaves = h->cwList ? 4u : 8u;
        }
        if (walk_of(flags) == Walk::CellWave && (flags & SPH_FLAG_FP16_STORED))
            SPH_ALLOC(h->off16, std::max<uint64_t>(N, 1));
        if (h->cwList) {
            SPH_ALLOC(h->nbr, (uint64_t)std::max<uint32_t>(h->numBlocks, 1) * kWaves * kListCap * 64u);
            SPH_ALLOC(h->nbrCount, std::max<uint64_t>(N, 1));
        } else {
            SPH_ALLOC(h->extraMask, std::max<uint64_t>(N, 1));
        }
    }
    if (flags & SPH_FLAG_RECORD_DEBUG) {
        SPH_ALLOC(h->dens, std::max<uint64_t>(N, 1));
        SPH_ALLOC(h->acc, std::max<uint64_t>(N, 1));
    }
#undef SPH_ALLOC
    SPH_HIP(h, hipMemcpy(h->dScenes, h->hScenes.data(), sizeof(SceneDev) * n_scenes, hipMemcpyHostToDevice));
    if (h->numBlocks)
        SPH_HIP(h, hipMemcpy(h->dBlocks, h->hBlocks.data(), sizeof(BlockDesc) * h->numBlocks, hipMemcpyHostToDevice));
    SPH_HIP(h, hipMemcpy(h->dSceneBlocks, h->hSceneBlocks.data(), sizeof(uint32_t) * (n_scenes + 1),
                         hipMemcpyHostToDevice));
    if (N) {
        SPH_HIP(h, hipMemcpy(h->pos, hPos.data(), sizeof(float4) * N, hipMemcpyHostToDevice));
        SPH_HIP(h, hipMemcpy(h->vel, hVel.data(), sizeof(float4) * N, hipMemcpyHostToDevice));
        SPH_HIP(h, hipMemcpy(h->ids[0], hIds.data(), sizeof(uint32_t) * N, hipMemcpyHostToDevice));
        // The initial sort moved the state into bucket order (SimulationEngineCPU.cpp:55-131):
        // the particle map starts as the identity (hIds holds slot k at k).
        if (h->pmap) SPH_HIP(h, hipMemcpy(h->pmap, hIds.data(), sizeof(uint32_t) * N, hipMemcpyHostToDevice));
        SPH_HIP(h, hipMemcpy(h->origOfSlot, hOrig.data(), sizeof(uint32_t) * N, hipMemcpyHostToDevice));
        SPH_HIP(h, hipMemset(h->pres, 0, sizeof(float) * N));
        if ((rc = refresh_offsets(h))) return rc;
    }
    SPH_HIP(h, hipMemcpy(h->tab[0], hTab.data(), sizeof(uint32_t) * L, hipMemcpyHostToDevice));
    hNan.resize(2ull * n_scenes, 0u);  // parity 1 starts at zero
    SPH_HIP(h, hipMemcpy(h->nanCnt, hNan.data(), sizeof(uint32_t) * hNan.size(), hipMemcpyHostToDevice));
    SPH_HIP(h, hipMemset(h->tab[1], 0, sizeof(uint32_t) * L));
    SPH_HIP(h, hipMemset(h->status[0], 0, sizeof(unsigned long long) * h->numTiles));
    SPH_HIP(h, hipMemset(h->status[1], 0, sizeof(unsigned long long) * h->numTiles));
    SPH_HIP(h, hipMemset(h->ticket, 0, sizeof(uint32_t)));
    {
        std::vector<uint32_t> identity(std::max<uint64_t>(h->numBlocks, 1));
        for (size_t t = 0; t < identity.size(); ++t) identity[t] = (uint32_t)t;
        SPH_HIP(h, hipMemcpy(h->tileOrder, identity.data(), sizeof(uint32_t) * identity.size(), hipMemcpyHostToDevice));
        SPH_HIP(h, hipMemset(h->heavy, 0, sizeof(uint32_t) * identity.size()));
    }
    if (h->mask) {
        const DensityOut o{h->heavy, h->nbrCount, h->mask, h->extraMask, h->wallSum, h->wallZ, h->dens, h->visc,
                           h->wallCells, h->wallList};
        if ((rc = dalloc(h, &h->densOut, 1))) return rc;
        SPH_HIP(h, hipMemcpy(h->densOut, &o, sizeof(o), hipMemcpyHostToDevice));
    }
    if (Ns) {
        SPH_HIP(h, hipMemcpy(h->spos, hSpos.data(), sizeof(float4) * Ns, hipMemcpyHostToDevice));
        SPH_HIP(h, hipMemcpy(h->stab, hStab.data(), sizeof(uint32_t) * h->Ls, hipMemcpyHostToDevice));
        SPH_HIP(h, hipMemcpy(h->wallMap, hWallMap.data(), sizeof(uint32_t) * h->wallWords, hipMemcpyHostToDevice));
        if (!hWallCells.empty())
            SPH_HIP(h, hipMemcpy(h->wallCells, hWallCells.data(), sizeof(uint32_t) * hWallCells.size(),
                                 hipMemcpyHostToDevice));
        if (!hWallList.empty())
            SPH_HIP(h, hipMemcpy(h->wallList, hWallList.data(), sizeof(uint32_t) * hWallList.size(),
                                 hipMemcpyHostToDevice));
        SPH_HIP(h, hipMemcpy(h->staticScene, hStaticScene.data(), sizeof(uint32_t) * Ns, hipMemcpyHostToDevice));
    }
    // Validation index lists: dynamic first/last entries per scene, then static.
    std::vector<uint32_t> vi(8ull * n_scenes);
    for (uint32_t s = 0; s < n_scenes; ++s) {
        const SceneDev& S = h->hScenes[s];
        vi[0 * n_scenes + s] = S.tBegin;
        vi[1 * n_scenes + s] = S.pBegin;
        vi[2 * n_scenes + s] = S.tBegin + S.M;
        vi[3 * n_scenes + s] = S.pBegin + S.pCount;
        vi[4 * n_scenes + s] = S.sCount ? S.stBegin : 0;
        vi[5 * n_scenes + s] = S.sCount ? S.sBegin : (h->Ls ? hStab[0] : 0);
        vi[6 * n_scenes + s] = S.sCount ? S.stBegin + S.Ms : 0;
        vi[7 * n_scenes + s] = S.sCount ? S.sBegin + S.sCount : (h->Ls ? hStab[0] : 0);
    }
    SPH_HIP(h, hipMemcpy(h->valIdx, vi.data(), sizeof(uint32_t) * vi.size(), hipMemcpyHostToDevice));
    SPH_HIP(h, hipDeviceSynchronize());
    h->cur = 0;
    h->simTime = 0.0;
    h->initialized = true;
    return SPH_OK;
}

int sph_set_triangles(sph_handle h, uint32_t scene, const float* tri9, uint64_t count) {
    if (!valid(h)) return SPH_E_INVALID;
    if (!h->initialized || scene >= h->nScenes) {
        h->err = "sph_set_triangles: engine not initialized or scene out of range";
        return SPH_E_STATE;
    }
    if (count && !tri9) return SPH_E_INVALID;
    SPH_HIP(h, hipSetDevice(h->device));
    SPH_HIP(h, hipStreamSynchronize(h->stream));
    // Rebuild the concatenated triangle list with this scene's triangles replaced.
    std::vector<float> all;
    std::vector<SceneDev>& S = h->hScenes;
    for (uint32_t s = 0; s < h->nScenes; ++s) {
        const uint32_t oldBegin = S[s].triBegin, oldCount = S[s].triCount;
        const uint32_t newBegin = (uint32_t)(all.size() / 9);
        if (s == scene) {
            all.insert(all.end(), tri9, tri9 + 9 * count);
            S[s].triCount = (uint32_t)count;
        } else {
            all.insert(all.end(), h->hTris.begin() + 9ull * oldBegin, h->hTris.begin() + 9ull * (oldBegin + oldCount));
        }
        S[s].triBegin = newBegin;
    }
    h->hTris.swap(all);
    h->nTris = h->hTris.size() / 9;
    h->hasTris = h->nTris > 0;
    // The mesh buffer: every scene's triangles, then every scene's bin tables.
    std::vector<uint32_t> tables;
    for (uint32_t s = 0; s < h->nScenes; ++s)
        build_mesh_bins(h->hTris.data() + 9ull * S[s].triBegin, S[s].triCount, S[s], tables, h->hTris.size());
    if (h->hTris.size() + tables.size() >= (1ull << 32)) {
        h->err = "sph_set_triangles: mesh tables above 2^32 entries";
        return SPH_E_INVALID;
    }
    std::vector<float> mesh(h->hTris);
    mesh.resize(h->hTris.size() + tables.size());
    if (!tables.empty()) memcpy(mesh.data() + h->hTris.size(), tables.data(), sizeof(uint32_t) * tables.size());
    dfree(h->tris);
    int rc;
    if ((rc = dalloc(h, &h->tris, std::max<uint64_t>(mesh.size(), 1)))) return rc;
    if (!mesh.empty())
        SPH_HIP(h, hipMemcpy(h->tris, mesh.data(), sizeof(float) * mesh.size(), hipMemcpyHostToDevice));
    SPH_HIP(h, hipMemcpy(h->dScenes, S.data(), sizeof(SceneDev) * h->nScenes, hipMemcpyHostToDevice));
    for (auto& g : h->graph) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    return SPH_OK;
}

int sph_set_reorder_interval(sph_handle h, float interval) {
    if (!valid(h)) return SPH_E_INVALID;
    if (!h->initialized) {
        h->err = "sph_set_reorder_interval: engine not initialized";
        return SPH_E_STATE;
    }
    h->reorderInterval = interval;
    return SPH_OK;
}

int sph_update(sph_handle h, float dt, uint32_t steps) {
    if (!valid(h)) return SPH_E_INVALID;
    if (!h->initialized) {
        h->err = "Updating an uninitialized SPH system";  // SimulationEngineCPU.cpp:422-426
        return SPH_E_STATE;
    }
    if (h->N == 0 || steps == 0) return SPH_OK;  // SimulationEngineCPU.cpp:428-429
    SPH_HIP(h, hipSetDevice(h->device));
    const bool useGraph = (h->flags & SPH_FLAG_HIP_GRAPH) && !h->profiling;
    // Particle-map mode: every step of this update reorders when the interval has
    // elapsed (SimulationEngineCPU.cpp:446), checked once per update.
    h->reorderNow = h->pmap && h->reorderElapsed > h->reorderInterval;
    for (uint32_t s = 0; s < steps; ++s) {
        if (useGraph) {
            const int c = h->cur, gi = 2 * (h->reorderNow ? 1 : 0) + c;
            if (!h->graph[gi] || h->graphDt[gi] != dt) {
                int rc = capture_graph(h, c, dt);
                if (rc) return rc;
            }
            SPH_HIP(h, hipGraphLaunch(h->graph[gi], h->stream));
        } else {
            int rc = enqueue_step(h, dt, false);
            if (rc) return rc;
        }
        h->cur ^= 1;
    }
    h->debugValid = (h->flags & SPH_FLAG_RECORD_DEBUG) != 0;
    h->simTime += (double)dt * steps;  // SimulationEngineCPU.cpp:453
    if (h->pmap) {  // SimulationEngineCPU.cpp:455-457: once per update
        h->reorderElapsed += dt;
        if (h->reorderElapsed > h->reorderInterval) h->reorderElapsed = 0.f;
    }
    h->reorderNow = false;
    return SPH_OK;
}

int sph_synchronize(sph_handle h) {
    if (!valid(h)) return SPH_E_INVALID;
    SPH_HIP(h, hipSetDevice(h->device));
    SPH_HIP(h, hipStreamSynchronize(h->stream));
    return SPH_OK;
}

int sph_read(sph_handle h, float* pos, float* vel, float* pressure) {
    if (!valid(h)) return SPH_E_INVALID;
    if (!h->initialized) {
        h->err = "sph_read: engine not initialized";
        return SPH_E_STATE;
    }
    if (h->N == 0) return SPH_OK;
    SPH_HIP(h, hipSetDevice(h->device));
    const uint64_t N = h->N;
    int rc = ensure_staging(h);
    if (rc) return rc;
    float* d = h->rbDev;
    hipLaunchKernelGGL(k_unsort, dim3((uint32_t)((N + 255) / 256)), dim3(256), 0, h->stream, h->pos, h->vel,
                       h->ids[h->cur], h->origOfSlot, (uint32_t)N, d, d + 3 * N, d + 6 * N, nullptr, nullptr, h->pmap);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && pos) e = hipMemcpy(pos, d, sizeof(float) * 3 * N, hipMemcpyDeviceToHost);
    if (e == hipSuccess && vel) e = hipMemcpy(vel, d + 3 * N, sizeof(float) * 3 * N, hipMemcpyDeviceToHost);
    if (e == hipSuccess && pressure) e = hipMemcpy(pressure, d + 6 * N, sizeof(float) * N, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        h->err = std::string("sph_read: ") + hipGetErrorString(e);
        return SPH_E_HIP;
    }
    return SPH_OK;
}

int sph_write(sph_handle h, const float* pos, const float* vel) {
    if (!valid(h)) return SPH_E_INVALID;
    if (!h->initialized) {
        h->err = "sph_write: engine not initialized";
        return SPH_E_STATE;
    }
    if (h->N == 0 || (!pos && !vel)) return SPH_OK;
    SPH_HIP(h, hipSetDevice(h->device));
    const uint64_t N = h->N;
    int rc = ensure_staging(h);
    if (rc) return rc;
    // The staging may still feed an asynchronous readback: wait for the stream first.
    SPH_HIP(h, hipStreamSynchronize(h->stream));
    float* d = h->rbDev;
    if (pos) SPH_HIP(h, hipMemcpy(d, pos, sizeof(float) * 3 * N, hipMemcpyHostToDevice));
    if (vel) SPH_HIP(h, hipMemcpy(d + 3 * N, vel, sizeof(float) * 3 * N, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_write, dim3((uint32_t)((N + 255) / 256)), dim3(256), 0, h->stream, h->pos, h->vel,
                       h->ids[h->cur], h->origOfSlot, (uint32_t)N, pos ? d : nullptr, vel ? d + 3 * N : nullptr);
    SPH_HIP(h, hipGetLastError());
    if (pos && (rc = refresh_offsets(h))) return rc;
    SPH_HIP(h, hipStreamSynchronize(h->stream));
    return SPH_OK;
}

int sph_read_noalloc(sph_handle h, float* pos, float* vel, float* pressure) {
    if (!valid(h)) return SPH_E_INVALID;
    if (!h->initialized) {
        h->err = "sph_read_noalloc: engine not initialized";
        return SPH_E_STATE;
    }
    if (h->N > SPH_NOALLOC_MAX_PARTICLES) {
        h->err = "sph_read_noalloc: " + std::to_string(h->N) + " particles, more than SPH_NOALLOC_MAX_PARTICLES (" +
                 std::to_string(SPH_NOALLOC_MAX_PARTICLES) + ")";
        return SPH_E_INVALID;
    }
    if (h->N == 0) return SPH_OK;
    SPH_HIP(h, hipSetDevice(h->device));
    // k_unsort writes every entry of [0, 7 N): nothing needs clearing first.
    const uint64_t N = h->N;
    float* d = h->scratch;
    hipLaunchKernelGGL(k_unsort, dim3((uint32_t)((N + 255) / 256)), dim3(256), 0, h->stream, h->pos, h->vel,
                       h->ids[h->cur], h->origOfSlot, (uint32_t)N, d, d + 3 * N, d + 6 * N, nullptr, nullptr, h->pmap);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && pos) e = hipMemcpy(pos, d, sizeof(float) * 3 * N, hipMemcpyDeviceToHost);
    if (e == hipSuccess && vel) e = hipMemcpy(vel, d + 3 * N, sizeof(float) * 3 * N, hipMemcpyDeviceToHost);
    if (e == hipSuccess && pressure) e = hipMemcpy(pressure, d + 6 * N, sizeof(float) * N, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        h->err = std::string("sph_read_noalloc: ") + hipGetErrorString(e);
        return SPH_E_HIP;
    }
    return SPH_OK;
}

int sph_readback_enqueue(sph_handle h, uint32_t* slot) {
    if (!valid(h) || !slot) return SPH_E_INVALID;
    if (!h->initialized) {
        h->err = "sph_readback_enqueue: engine not initialized";
        return SPH_E_STATE;
    }
    SPH_HIP(h, hipSetDevice(h->device));
    const uint64_t N = h->N;
    const uint64_t words = 8 * N;  // pos 3N, vel 3N, pressure N, key N
    if (!h->rbEvent[0] || h->rbN != N || (N && !h->rbHost[0])) {  // first use after initialize
        int rc = ensure_staging(h);
        if (rc) return rc;
        for (uint32_t k = 0; k < sph_engine::kReadbackSlots; ++k) {
            if (h->rbHost[k]) (void)hipHostFree(h->rbHost[k]);
            h->rbHost[k] = nullptr;
            h->rbFilled[k] = false;
            if (N) SPH_HIP(h, hipHostMalloc(&h->rbHost[k], sizeof(float) * words, hipHostMallocDefault));
            if (!h->rbEvent[k]) SPH_HIP(h, hipEventCreateWithFlags(&h->rbEvent[k], hipEventDisableTiming));
        }
        h->rbN = N;
        h->rbNext = 0;
    }
    const uint32_t k = h->rbNext;
    if (N) {
        float* d = h->rbDev;
        hipLaunchKernelGGL(k_unsort, dim3((uint32_t)((N + 255) / 256)), dim3(256), 0, h->stream, h->pos, h->vel,
                           h->ids[h->cur], h->origOfSlot, (uint32_t)N, d, d + 3 * N, d + 6 * N,
                           reinterpret_cast<uint32_t*>(d + 7 * N), reinterpret_cast<uint32_t*>(d + 8 * N), h->pmap);
        SPH_HIP(h, hipGetLastError());
        SPH_HIP(h, hipMemcpyAsync(h->rbHost[k], d, sizeof(float) * words, hipMemcpyDeviceToHost, h->stream));
    }
    SPH_HIP(h, hipEventRecord(h->rbEvent[k], h->stream));
    h->rbFilled[k] = true;
    h->rbTime[k] = h->simTime;
    h->rbNext = (k + 1) % sph_engine::kReadbackSlots;
    *slot = k;
    return SPH_OK;
}

int sph_readback_query(sph_handle h, uint32_t slot, int* ready) {
    if (!valid(h) || !ready || slot >= sph_engine::kReadbackSlots) return SPH_E_INVALID;
    if (!h->rbFilled[slot]) {
        h->err = "sph_readback_query: slot holds no readback";
        return SPH_E_STATE;
    }
    SPH_HIP(h, hipSetDevice(h->device));
    const hipError_t e = hipEventQuery(h->rbEvent[slot]);
    if (e == hipSuccess) {
        *ready = 1;
    } else if (e == hipErrorNotReady) {
        *ready = 0;
    } else {
        h->err = std::string("sph_readback_query: ") + hipGetErrorString(e);
        return SPH_E_HIP;
    }
    return SPH_OK;
}

int sph_readback_lock(sph_handle h, uint32_t slot, const float** pos, const float** vel, const float** pressure,
                      const uint32_t** keys, uint64_t* n, double* sim_time) {
    if (!valid(h) || slot >= sph_engine::kReadbackSlots) return SPH_E_INVALID;
    if (!h->rbFilled[slot]) {
        h->err = "sph_readback_lock: slot holds no readback";
        return SPH_E_STATE;
    }
    SPH_HIP(h, hipSetDevice(h->device));
    SPH_HIP(h, hipEventSynchronize(h->rbEvent[slot]));
    const uint64_t N = h->rbN;
    const float* b = static_cast<const float*>(h->rbHost[slot]);
    if (pos) *pos = b;
    if (vel) *vel = b ? b + 3 * N : nullptr;
    if (pressure) *pressure = b ? b + 6 * N : nullptr;
    if (keys) *keys = b ? reinterpret_cast<const uint32_t*>(b + 7 * N) : nullptr;
    if (n) *n = N;
    if (sim_time) *sim_time = h->rbTime[slot];
    return SPH_OK;
}

namespace {
// keys (original order) and/or bucket ids (sorted order) of the current state.
int read_keys_ids(sph_engine* h, uint32_t* keys, uint32_t* bucketIds) {
    if (!h->initialized) {
        h->err = "engine not initialized";
        return SPH_E_STATE;
    }
    if (h->N == 0) return SPH_OK;
    SPH_HIP(h, hipSetDevice(h->device));
    const uint64_t N = h->N;
    int rc = ensure_staging(h);
    if (rc) return rc;
    float* d = h->rbDev;
    uint32_t* dK = reinterpret_cast<uint32_t*>(d + 7 * N);
    uint32_t* dI = reinterpret_cast<uint32_t*>(d + 8 * N);
    hipLaunchKernelGGL(k_unsort, dim3((uint32_t)((N + 255) / 256)), dim3(256), 0, h->stream, h->pos, h->vel,
                       h->ids[h->cur], h->origOfSlot, (uint32_t)N, d, d + 3 * N, d + 6 * N, dK, dI, h->pmap);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && keys) e = hipMemcpy(keys, dK, sizeof(uint32_t) * N, hipMemcpyDeviceToHost);
    if (e == hipSuccess && bucketIds) e = hipMemcpy(bucketIds, dI, sizeof(uint32_t) * N, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        h->err = std::string("read: ") + hipGetErrorString(e);
        return SPH_E_HIP;
    }
    return SPH_OK;
}
}  // namespace

int sph_read_keys(sph_handle h, uint32_t* keys) {
    if (!valid(h) || !keys) return SPH_E_INVALID;
    return read_keys_ids(h, keys, nullptr);
}

int sph_read_bucket_ids(sph_handle h, uint32_t* ids) {
    if (!valid(h) || !ids) return SPH_E_INVALID;
    return read_keys_ids(h, nullptr, ids);
}

int sph_read_cell_starts(sph_handle h, uint32_t* starts) {
    if (!valid(h) || !starts) return SPH_E_INVALID;
    if (!h->initialized) {
        h->err = "sph_read_cell_starts: engine not initialized";
        return SPH_E_STATE;
    }
    SPH_HIP(h, hipSetDevice(h->device));
    SPH_HIP(h, hipStreamSynchronize(h->stream));
    SPH_HIP(h, hipMemcpy(starts, h->tab[h->cur], sizeof(uint32_t) * h->L, hipMemcpyDeviceToHost));
    for (uint32_t s = 0; s < h->nScenes; ++s) {
        const SceneDev& S = h->hScenes[s];
        for (uint64_t b = 0; b <= S.M; ++b) starts[S.tBegin + b] -= S.pBegin;
    }
    return SPH_OK;
}

int sph_read_static(sph_handle h, uint32_t* starts, uint32_t* ids) {
    if (!valid(h)) return SPH_E_INVALID;
    if (!h->initialized) {
        h->err = "sph_read_static: engine not initialized";
        return SPH_E_STATE;
    }
    SPH_HIP(h, hipSetDevice(h->device));
    if (starts && h->Ls) {
        SPH_HIP(h, hipMemcpy(starts, h->stab, sizeof(uint32_t) * h->Ls, hipMemcpyDeviceToHost));
        for (uint32_t s = 0; s < h->nScenes; ++s) {
            const SceneDev& S = h->hScenes[s];
            if (!S.sCount) continue;
            for (uint64_t b = 0; b <= S.Ms; ++b) starts[S.stBegin + b] -= S.sBegin;
        }
    }
    if (ids && h->Ns) memcpy(ids, h->hStaticIds.data(), sizeof(uint32_t) * h->Ns);
    return SPH_OK;
}

int sph_read_debug(sph_handle h, float* density, float* accel) {
    if (!valid(h)) return SPH_E_INVALID;
    if (!(h->flags & SPH_FLAG_RECORD_DEBUG) || !h->debugValid) {
        h->err = "sph_read_debug: needs SPH_FLAG_RECORD_DEBUG and at least one step";
        return SPH_E_STATE;
    }
    SPH_HIP(h, hipSetDevice(h->device));
    const uint64_t N = h->N;
    int rc = ensure_staging(h);
    if (rc) return rc;
    float *dD = h->rbDev, *dA = h->rbDev + N;
    // The debug arrays are in the order of the last step's input: ids[cur ^ 1].
    hipLaunchKernelGGL(k_unsort_debug, dim3((uint32_t)((N + 255) / 256)), dim3(256), 0, h->stream, h->dens, h->acc,
                       h->ids[h->cur ^ 1], h->origOfSlot, (uint32_t)N, dD, dA);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && density) e = hipMemcpy(density, dD, sizeof(float) * N, hipMemcpyDeviceToHost);
    if (e == hipSuccess && accel) e = hipMemcpy(accel, dA, sizeof(float) * 3 * N, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        h->err = std::string("sph_read_debug: ") + hipGetErrorString(e);
        return SPH_E_HIP;
    }
    return SPH_OK;
}

int sph_conserved(sph_handle h, double* out) {
    if (!valid(h) || !out) return SPH_E_INVALID;
    if (!h->initialized) {
        h->err = "sph_conserved: engine not initialized";
        return SPH_E_STATE;
    }
    SPH_HIP(h, hipSetDevice(h->device));
    if (h->numBlocks)
        hipLaunchKernelGGL(k_conserved_partial, dim3(h->numBlocks), dim3(kBlock), 0, h->stream, h->dBlocks, h->pos,
                           h->vel, h->consPartial);
    hipLaunchKernelGGL(k_conserved_final, dim3(h->nScenes), dim3(64), 0, h->stream, h->dSceneBlocks, h->dScenes,
                       h->consPartial, h->consOut);
    SPH_HIP(h, hipGetLastError());
    SPH_HIP(h, hipStreamSynchronize(h->stream));
    SPH_HIP(h, hipMemcpy(out, h->consOut, sizeof(double) * 8 * h->nScenes, hipMemcpyDeviceToHost));
    return SPH_OK;
}

int sph_validate(sph_handle h, uint32_t* bits) {
    if (!valid(h)) return SPH_E_INVALID;
    if (!h->initialized) {
        h->err = "sph_validate: engine not initialized";
        return SPH_E_STATE;
    }
    SPH_HIP(h, hipSetDevice(h->device));
    SPH_HIP(h, hipMemsetAsync(h->valBits, 0, sizeof(uint32_t), h->stream));
    const uint32_t ns = h->nScenes;
    if (h->numBlocks)
        hipLaunchKernelGGL(k_validate_particles, dim3(h->numBlocks), dim3(kBlock), 0, h->stream, h->dBlocks,
                           h->dScenes, h->pos, h->vel, h->tab[h->cur], h->pmap, h->valBits);
    {
        const uint64_t n = std::max<uint64_t>(h->L, ns);
        hipLaunchKernelGGL(k_validate_table, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, h->stream,
                           h->tab[h->cur], (uint32_t)h->L, h->valIdx, h->valIdx + ns, h->valIdx + 2 * ns,
                           h->valIdx + 3 * ns, ns, 0u, h->valBits);
    }
    if (h->Ns) {
        // Static table: same checks, reported under SPH_V_STATIC.
        const uint64_t n = std::max<uint64_t>(h->Ls, ns);
        hipLaunchKernelGGL(k_validate_table, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, h->stream, h->stab,
                           (uint32_t)h->Ls, h->valIdx + 4 * ns, h->valIdx + 5 * ns, h->valIdx + 6 * ns,
                           h->valIdx + 7 * ns, ns, 4u, h->valBits);
        hipLaunchKernelGGL(k_validate_static, dim3((uint32_t)((h->Ns + 255) / 256)), dim3(256), 0, h->stream,
                           h->dScenes, ns, h->staticScene, h->spos, h->stab, (uint32_t)h->Ns, h->valBits);
    }
    SPH_HIP(h, hipGetLastError());
    uint32_t b = 0;
    SPH_HIP(h, hipMemcpyAsync(&b, h->valBits, sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
    SPH_HIP(h, hipStreamSynchronize(h->stream));
    // Static table bits were shifted by 4 into 0x40..0x200; fold them into SPH_V_STATIC.
    uint32_t out = b & 0x3Fu;
    if (b & ~0x3Fu) out |= SPH_V_STATIC;
    if (bits) *bits = out;
    return out ? SPH_E_VALIDATION : SPH_OK;
}

float sph_simulation_time(sph_handle h) { return h ? (float)h->simTime : 0.f; }

uint64_t sph_particle_count(sph_handle h) { return h ? h->N : 0; }
uint64_t sph_static_particle_count(sph_handle h) { return h ? h->Ns : 0; }
uint32_t sph_scene_count(sph_handle h) { return h ? h->nScenes : 0; }
uint64_t sph_hash_map_size(sph_handle h, uint32_t scene) {
    if (!h || scene >= h->nScenes) return 0;
    return h->hScenes[scene].M;
}

int sph_set_profiling(sph_handle h, uint32_t stage_mask) {
    if (!valid(h)) return SPH_E_INVALID;
    SPH_HIP(h, hipSetDevice(h->device));
    SPH_HIP(h, hipStreamSynchronize(h->stream));
    int rc = collect_events(h);
    if (rc) return rc;
    h->profiling = stage_mask & (SPH_PROFILE_ALL | SPH_PROFILE_ROTATE);
    h->profStep = 0;
    for (int s = 0; s < SPH_STAGE_COUNT; ++s) {
        h->stageMs[s] = 0;
        h->stageLaunches[s] = 0;
    }
    return SPH_OK;
}

int sph_stage_times(sph_handle h, double* ms, uint64_t* launches) {
    if (!valid(h)) return SPH_E_INVALID;
    SPH_HIP(h, hipSetDevice(h->device));
    int rc = collect_events(h);
    if (rc) return rc;
    for (int s = 0; s < SPH_STAGE_COUNT; ++s) {
        if (ms) ms[s] = h->stageMs[s];
        if (launches) launches[s] = h->stageLaunches[s];
    }
    return SPH_OK;
}

namespace {
// Both ways for the cell-wave walk to hand neighbours to the force pass (lists,
// or masks with the crowded buckets' record), so the autotuner can time each:
// allocates whichever the engine lacks and refreshes the density pass's output
// pointers.
int ensure_cw_handoffs(sph_engine* h) {
    int rc;
    if (!h->nbr) {
        if ((rc = dalloc(h, &h->nbr, (uint64_t)std::max<uint32_t>(h->numBlocks, 1) * kWaves * kListCap * 64u)))
            return rc;
        if ((rc = dalloc(h, &h->nbrCount, std::max<uint64_t>(h->N, 1)))) return rc;
    }
    if (!h->extraMask && (rc = dalloc(h, &h->extraMask, std::max<uint64_t>(h->N, 1)))) return rc;
    const DensityOut o{h->heavy, h->nbrCount, h->mask, h->extraMask, h->wallSum, h->wallZ, h->dens, h->visc,
                       h->wallCells, h->wallList};
    SPH_HIP(h, hipMemcpy(h->densOut, &o, sizeof(o), hipMemcpyHostToDevice));
    return SPH_OK;
}
}  // namespace

int sph_autotune(sph_handle h, float dt, uint32_t trial_steps, uint32_t* chosen) {
    if (!valid(h)) return SPH_E_INVALID;
    if (!h->initialized) {
        h->err = "sph_autotune: engine not initialized";
        return SPH_E_STATE;
    }
    const uint32_t walkBits = SPH_FLAG_CELL_WALK | SPH_FLAG_RUN_WALK | SPH_FLAG_TILE_WALK;
    const uint32_t handoffBits = SPH_FLAG_CW_LISTS | SPH_FLAG_CW_MASKS;
    auto handoff = [&](uint32_t walk, bool lists) {
        return walk == 0u ? (lists ? SPH_FLAG_CW_LISTS : SPH_FLAG_CW_MASKS) : 0u;
    };
    if (h->N == 0 || !h->mask) {  // nothing to time, or the engine was set up for the cell walk only
        if (chosen)
            *chosen = walk_of(h->flags) == Walk::Cells ? SPH_FLAG_CELL_WALK
                                                       : (h->flags & walkBits) | handoff(h->flags & walkBits, h->cwList);
        return SPH_OK;
    }
    SPH_HIP(h, hipSetDevice(h->device));
    {
        const int arc = ensure_cw_handoffs(h);
        if (arc) return arc;
    }
    const uint32_t warm = 2, timed = std::max<uint32_t>(2u, (trial_steps + 1u) & ~1u);  // even: same parity after
    const uint64_t N = h->N, L = h->L;
    const int c = h->cur;
    const uint32_t savedProfiling = h->profiling;
    const uint32_t savedFlags = h->flags;
    const bool savedList = h->cwList;
    const uint32_t savedShape[2] = {h->densWaves, h->forceWaves};
    h->profiling = 0;
    float4 *bPos = nullptr, *bVel = nullptr;
    uint32_t *bIds = nullptr, *bTab = nullptr, *bNan = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SPH_OK;
    auto finish = [&](int code) {
        dfree(bPos);
        dfree(bVel);
        dfree(bIds);
        dfree(bTab);
        dfree(bNan);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        h->profiling = savedProfiling;
        return code;
    };
    // Snapshot of the state a step starts from: tab[c] holds the cell starts,
    // tab[c ^ 1] is all zero, the NaN counts of parity c are this state's and
    // those of parity c ^ 1 are zero, and the scan's status arrays are in the
    // state parity c expects after any even number of steps.
    const uint64_t S2 = 2ull * h->nScenes;
    if ((rc = dalloc(h, &bPos, N)) || (rc = dalloc(h, &bVel, N)) || (rc = dalloc(h, &bIds, N)) ||
        (rc = dalloc(h, &bTab, L)) || (rc = dalloc(h, &bNan, S2)))
        return finish(rc);
    auto copy = [&](void* dst, const void* src, uint64_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, h->stream);
    };
    auto restore = [&]() -> hipError_t {
        hipError_t e;
        if ((e = copy(h->pos, bPos, N * sizeof(float4))) != hipSuccess) return e;
        if ((e = copy(h->vel, bVel, N * sizeof(float4))) != hipSuccess) return e;
        if ((e = copy(h->ids[c], bIds, N * sizeof(uint32_t))) != hipSuccess) return e;
        if ((e = copy(h->tab[c], bTab, L * sizeof(uint32_t))) != hipSuccess) return e;
        if ((e = copy(h->nanCnt, bNan, S2 * sizeof(uint32_t))) != hipSuccess) return e;
        if (h->off16) {  // the offset records of the restored positions
            hipLaunchKernelGGL(k_offsets, dim3(h->numBlocks), dim3(kBlock), 0, h->stream, h->dBlocks, h->dScenes,
                               h->pos, h->off16);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        return hipMemsetAsync(h->tab[c ^ 1], 0, L * sizeof(uint32_t), h->stream);
    };
    hipError_t e = copy(bPos, h->pos, N * sizeof(float4));
    if (e == hipSuccess) e = copy(bVel, h->vel, N * sizeof(float4));
    if (e == hipSuccess) e = copy(bIds, h->ids[c], N * sizeof(uint32_t));
    if (e == hipSuccess) e = copy(bTab, h->tab[c], L * sizeof(uint32_t));
    if (e == hipSuccess) e = copy(bNan, h->nanCnt, S2 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipEventCreateWithFlags(&e0, hipEventDisableSystemFence);  // timing only
    if (e == hipSuccess) e = hipEventCreateWithFlags(&e1, hipEventDisableSystemFence);
    if (e != hipSuccess) {
        h->err = std::string("sph_autotune: ") + hipGetErrorString(e);
        return finish(SPH_E_HIP);
    }
    // A trial that fails leaves the engine as configured: the starting parity, the
    // configured variant and shape, and the state of the snapshot.
    auto abandon = [&](int code) {
        h->cur = c;
        h->flags = savedFlags;
        h->cwList = savedList;
        h->densWaves = savedShape[0];
        h->forceWaves = savedShape[1];
        (void)hipStreamSynchronize(h->stream);
        (void)restore();
        (void)hipStreamSynchronize(h->stream);
        return finish(code);
    };
    // One trial of the engine's current variant and shape from the snapshot: warm-up
    // steps, then the timed ones; the state is restored afterwards.
    auto trial = [&](float& ms) -> int {
        for (uint32_t s = 0; s < warm + timed; ++s) {
            if (s == warm && (e = hipEventRecord(e0, h->stream)) != hipSuccess) break;
            if ((rc = enqueue_step(h, dt, false))) return rc;
            h->cur ^= 1;
        }
        if (e == hipSuccess) e = hipEventRecord(e1, h->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        h->cur = c;
        if (e == hipSuccess) e = restore();
        if (e != hipSuccess) {
            h->err = std::string("sph_autotune: ") + hipGetErrorString(e);
            return SPH_E_HIP;
        }
        return SPH_OK;
    };
    // The engine's configured variant first, then the others among the cell-wave
    // walk with lists and with masks and the Morton runs. The bucket tiles, many
    // times slower than the cell-wave walk on every workload measured, are timed
    // only when they are the configured walk. A variant replaces the best so far
    // only when it is more than 3% faster: the trial steps are few, and the
    // configured choice reflects the whole run (the lists' advantage at 1M grows
    // as the fluid compresses).
    uint32_t candidates[4] = {0u, 0u, SPH_FLAG_RUN_WALK, SPH_FLAG_TILE_WALK};
    bool candLists[4] = {true, false, false, false};
    int numCandidates = 3;
    {
        const uint32_t cw = h->flags & walkBits;
        int cur = cw == SPH_FLAG_RUN_WALK ? 2 : cw == SPH_FLAG_TILE_WALK ? 3 : (h->cwList ? 0 : 1);
        if (cur == 3) numCandidates = 4;
        std::swap(candidates[0], candidates[cur]);
        std::swap(candLists[0], candLists[cur]);
    }
    const uint32_t base = h->flags & ~(walkBits | handoffBits);
    uint32_t best = h->flags & walkBits;
    bool bestLists = h->cwList;
    float bestMs = 0.f;
    bool first = true;
    for (int ci = 0; ci < numCandidates; ++ci) {
        const uint32_t w = candidates[ci];
        h->flags = base | w;
        if (w == 0u) h->cwList = candLists[ci];
        float ms = 0.f;
        if ((rc = trial(ms))) return abandon(rc);
        if (first || ms < 0.97f * bestMs) {
            best = w;
            if (w == 0u) bestLists = candLists[ci];
            bestMs = ms;
            first = false;
        }
    }
    h->cwList = bestLists;
    h->flags = base | best | handoff(best, bestLists);
    // The cell-wave walk's launch shape, when that walk won: waves per workgroup of
    // the density pass, then of the force pass; each replaces the configured one
    // only when more than 3% faster. The results are the same whatever the shape.
    if (best == 0u) {
        float curMs = 0.f;
        if ((rc = trial(curMs))) return abandon(rc);
        const uint32_t waves[4] = {1u, 2u, 4u, 8u};
        for (int pass = 0; pass < 2; ++pass) {
            uint32_t& field = pass == 0 ? h->densWaves : h->forceWaves;
            uint32_t bestW = field;
            for (uint32_t w : waves) {
                if (w == savedShape[pass]) continue;
                field = w;
                float ms = 0.f;
                if ((rc = trial(ms))) return abandon(rc);
                if (ms < 0.97f * curMs) {
                    curMs = ms;
                    bestW = w;
                }
            }
            field = bestW;
        }
    }
    for (auto& g : h->graph)  // captured steps may hold another walk's kernels or shape
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    h->debugValid = false;
    e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        h->err = std::string("sph_autotune: ") + hipGetErrorString(e);
        return finish(SPH_E_HIP);
    }
    if (chosen) *chosen = best | handoff(best, bestLists);
    return finish(SPH_OK);
}

int sph_set_launch_shape(sph_handle h, uint32_t density_waves, uint32_t force_waves) {
    if (!valid(h)) return SPH_E_INVALID;
    auto wavesOk = [](uint32_t w) { return w == 1u || w == 2u || w == 4u || w == 8u; };
    if (!wavesOk(density_waves) || !wavesOk(force_waves)) {
        h->err = "sph_set_launch_shape: waves per workgroup must be 1, 2, 4 or 8";
        return SPH_E_INVALID;
    }
    SPH_HIP(h, hipSetDevice(h->device));
    SPH_HIP(h, hipStreamSynchronize(h->stream));
    h->densWaves = density_waves;
    h->forceWaves = force_waves;
    h->shapeSet = true;
    for (auto& g : h->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    return SPH_OK;
}

int sph_launch_shape(sph_handle h, uint32_t* density_waves, uint32_t* force_waves) {
    if (!valid(h)) return SPH_E_INVALID;
    if (density_waves) *density_waves = h->densWaves;
    if (force_waves) *force_waves = h->forceWaves;
    return SPH_OK;
}

int sph_copy_probe(sph_handle h, uint64_t bytes, uint32_t repeats) {
    if (!valid(h) || bytes < 16) return SPH_E_INVALID;
    SPH_HIP(h, hipSetDevice(h->device));
    const uint64_t n16 = bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    int rc;
    if ((rc = dalloc(h, &a, n16)) || (rc = dalloc(h, &b, n16))) {
        dfree(a);
        dfree(b);
        return rc;
    }
    hipError_t e = hipMemsetAsync(a, 1, n16 * 16, h->stream);
    for (uint32_t r = 0; r < repeats && e == hipSuccess; ++r) {
        hipLaunchKernelGGL(k_copy_probe, dim3(4096), dim3(256), 0, h->stream, a, b, n16);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    dfree(a);
    dfree(b);
    if (e != hipSuccess) {
        h->err = std::string("sph_copy_probe: ") + hipGetErrorString(e);
        return SPH_E_HIP;
    }
    return SPH_OK;
}

}  // extern "C"

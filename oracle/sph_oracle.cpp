// SPH CPU oracle: TEST INFRASTRUCTURE ONLY.
//
// A restatement, in portable C++17 with std::thread, of the reference's
// multithreaded CPU engine (Murk42/SPHFluidSimulation):
//   SPHFluidSimulation/source/SPH/Kernels/SPHFunctions.cpp            ("SF")
//   SPHFluidSimulation/source/SPH/SimulationEngines/SimulationEngineCPU.cpp ("CPU")
//   SPHFluidSimulation/include/SPH/Core/SimulationEngine.h            ("SE.h")
//   SPHFluidSimulation/source/SPH/Concurrency/ThreadParallelTaskManager.cpp
// The reference cannot be compiled in this image (it needs the un-vendored
// BlazeEngine headers and MSVC), so this file is the checker: tests/ compare the
// HIP engine against it, and the benchmark times it as the CPU baseline
// ("port"). Nothing on the product path loads this library.
//
// Everything is templated on the arithmetic type R: R = float is the faithful
// FP32 oracle (same operation order as the reference, compiled with
// -ffp-contract=off so no FMA is formed), R = double is the FP64 oracle used to
// bound the FP32 error.
//
// How each function is pinned (the tests that check it, none of which retype
// the formula under test):
//   cell_of          tests/test_oracle_math.py: numpy float32 floor-division,
//                    negative coordinates wrap to 2^32 - k.
//   spread3/morton   tests/test_oracle_math.py: bit-by-bit interleave loop over
//                    all 1024 values per axis and random triples.
//   kernel_constant, D0/D1/D2
//                    tests/test_oracle_math.py: C * integral of W over the ball
//                    equals pi/3.1415 (pins the 3.1415 constant), D1 = dD0/dr and
//                    D2 = d^2 D0/dr^2 by finite differences, r >= h gives 0.
//   random_direction tests/test_oracle_math.py: unit length, C++ modf sign rule
//                    (parity of the exact values with the reference is unpinned:
//                    it depends on the platform's sinf).
//   counting sorts   tests/test_oracle_engine.py: brute-force histogram and
//                    exclusive scan with sentinel N, the validators of
//                    SE.h:115-282, within-bucket order of a 1-thread init.
//   update_pressure  tests/test_oracle_engine.py: O(N^2) numpy density with the
//                    27-cell bucket-visit rule (double visits on colliding
//                    buckets), FP32 against FP64.
//   update_dynamics  tests/test_oracle_engine.py: closed-form accelerations of a
//                    particle pair (pressure and viscosity, p > 0 and p < 0) and
//                    of a particle beside one wall (|p_i|, the wall summed once
//                    per neighbour-cell visit), an all-pairs numpy acceleration
//                    weighted by bucket-visit counts (hash sizes 2N, 7, 1000),
//                    free fall, momentum balance, Galilean invariance, zero
//                    static viscosity, FP32 against FP64.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace sphoracle {

using u32 = uint32_t;
using u64 = uint64_t;

template <class R>
struct V3 {
    R x, y, z;
};
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline R dot3(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

struct U3 {
    u32 x, y, z;
};

// SF:65-76, SF:99-102. One IEEE division per axis, floor, then a conversion
// through int32 so that negative cells wrap modulo 2^32 (the OpenCL build's
// as_uint(convert_int_rtn(...)); the C++ build's static_cast<uint> of a negative
// float is undefined, so the defined wrap is the rule).
template <class R>
inline U3 cell_of(V3<R> p, R h) {
    return {(u32)(int32_t)std::floor(p.x / h), (u32)(int32_t)std::floor(p.y / h),
            (u32)(int32_t)std::floor(p.z / h)};
}

// SF:79-86 (ExpandBits): bit k of the 10-bit input moves to bit 3k.
inline u32 spread3(u32 v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// SF:88-97 and SF:103-106 (Morton3D, GetHash): x on bit 3k, y on 3k+1, z on 3k+2.
// Only the low 10 bits of each cell coordinate take part, so cells alias every
// 1024 per axis.
inline u32 morton(u32 cx, u32 cy, u32 cz) { return spread3(cx) | (spread3(cy) << 1) | (spread3(cz) << 2); }

// SF:115-118. The reference writes 3.1415f, not pi; kept on purpose.
template <class R>
inline R kernel_constant(R h) {
    return R(15) / (R(3.1415) * h * h * h * h * h * h);
}
// SF:119-125
template <class R>
inline R D0(R r, R h) {
    if (r >= h) return 0;
    return (h - r) * (h - r) * (h - r);
}
// SF:126-132
template <class R>
inline R D1(R r, R h) {
    if (r >= h) return 0;
    return R(-3) * (h - r) * (h - r);
}
// SF:133-139
template <class R>
inline R D2(R r, R h) {
    if (r >= h) return 0;
    return R(6) * (h - r);
}

// SF:140-144 with the C++ build's modf (keeps the sign of its argument).
template <class R>
inline R noise(R x) {
    R ip = 0;
    return std::modf(std::sin(x * R(112.9898)) * R(43758.5453), &ip);
}
// SF:145-155. Noise is called twice with the same argument, as in the reference.
template <class R>
inline V3<R> random_direction(R x) {
    R theta = noise(x) * R(2) * R(3.1415);
    R z = noise(x) * R(2) - R(1);
    R s = std::sin(theta);
    R c = std::cos(theta);
    R z2 = std::sqrt(R(1) - z * z);
    return {z2 * c, z2 * s, z};
}

template <class R>
struct Triangle {
    V3<R> p1, p2, p3;
};

// SF:156-201 (ClosestPointOnTriangle), region tests in the same order.
template <class R>
inline V3<R> closest_point_on_triangle(V3<R> p, V3<R> a, V3<R> b, V3<R> c) {
    const V3<R> ab = b - a, ac = c - a, ap = p - a;
    const R d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    if (d1 <= R(0) && d2 <= R(0)) return a;
    const V3<R> bp = p - b;
    const R d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= R(0) && d4 <= d3) return b;
    const V3<R> cp = p - c;
    const R d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 >= R(0) && d5 <= d6) return c;
    const R vc = d1 * d4 - d3 * d2;
    if (vc <= R(0) && d1 >= R(0) && d3 <= R(0)) {
        const R v = d1 / (d1 - d3);
        return a + ab * v;
    }
    const R vb = d5 * d2 - d1 * d6;
    if (vb <= R(0) && d2 >= R(0) && d6 <= R(0)) {
        const R v = d2 / (d2 - d6);
        return a + ac * v;
    }
    const R va = d3 * d6 - d5 * d4;
    if (va <= R(0) && (d4 - d3) >= R(0) && (d5 - d6) >= R(0)) {
        const R v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return b + (c - b) * v;
    }
    const R denom = R(1) / (va + vb + vc);
    const R v = vb * denom, w = vc * denom;
    return a + ab * v + ac * w;
}
// SF:202-223: range 2.5, amplitude 50 * (1/t - 1)^2 with t = dist / 2.5.
template <class R>
inline V3<R> triangle_force(V3<R> point, const Triangle<R>& t) {
    const R range = R(2.5);
    V3<R> toTri = closest_point_on_triangle(point, t.p1, t.p2, t.p3) - point;
    R dist = std::sqrt(dot3(toTri, toTri));
    V3<R> dir = (-toTri) / dist;
    if (dist > range) return {0, 0, 0};
    R tt = R(1) / (dist / range) - R(1);
    return dir * (tt * tt) * R(50);
}

// Particles.h:9-20: 32-byte dynamic record, 16-byte static record (R = float).
template <class R>
struct DynP {
    V3<R> pos;
    R pressure;
    V3<R> vel;
    u32 hash;
};
template <class R>
struct StatP {
    V3<R> pos;
    R pressure;
};

// Particles.h:23-41 / SF:23-41
template <class R>
struct Params {
    R mass, k, elasticity, mu, gx, gy, gz, rho0, h, selfDensity, C;
};

using AtomicTable = std::vector<std::atomic<u32>>;

// SF:358-501 (UpdateParticlePressure). densityOut is an oracle-only probe.
template <class R>
void update_pressure(u64 i, u64 M, u64 ns, u64 Ms, const DynP<R>* in, DynP<R>* out, const std::atomic<u32>* hashMap,
                     const u32* particleMap, const StatP<R>* st, const std::atomic<u32>* stHash, const Params<R>& P,
                     R* densityOut) {
    const DynP<R>* self = in + i;
    const V3<R> x = self->pos;
    const U3 c = cell_of(x, P.h);
    const U3 b = {c.x - 1u, c.y - 1u, c.z - 1u};
    const U3 e = {c.x + 2u, c.y + 2u, c.z + 2u};
    R sumDyn = 0, sumSt = 0;
    for (u32 ox = b.x; ox != e.x; ++ox)
        for (u32 oy = b.y; oy != e.y; ++oy)
            for (u32 oz = b.z; oz != e.z; ++oz) {
                const u32 code = morton(ox, oy, oz);
                u32 bm = (u32)(code % M);
                u32 s = hashMap[bm], t = hashMap[bm + 1];
                for (u32 k = s; k < t; ++k) {
                    const DynP<R>* o = in + particleMap[k];
                    if (o == self) continue;
                    V3<R> d = o->pos - x;
                    R r2 = dot3(d, d);
                    if (r2 > P.h * P.h) continue;
                    R r = std::sqrt(r2);
                    sumDyn += D0(r, P.h);
                }
                if (ns == 0) continue;
                bm = (u32)(code % Ms);
                s = stHash[bm];
                t = stHash[bm + 1];
                for (u32 k = s; k < t; ++k) {
                    V3<R> d = st[k].pos - x;
                    R r2 = dot3(d, d);
                    if (r2 > P.h * P.h) continue;
                    R r = std::sqrt(r2);
                    sumSt += D0(r, P.h);
                }
            }
    R rho = P.selfDensity + (sumDyn * P.mass + sumSt * P.mass) * P.C;
    R p = P.k * (rho - P.rho0);
    out[i].pressure = p;
    if (densityOut) densityOut[i] = rho;
}

// SF:503-656 (UpdateParticleDynamics). accelOut is an oracle-only probe.
template <class R>
void update_dynamics(u64 i, u64 M, u64 ns, u64 Ms, const DynP<R>* in, DynP<R>* out, const std::atomic<u32>* hashMap,
                     const u32* particleMap, const StatP<R>* st, const std::atomic<u32>* stHash, R dt,
                     const Params<R>& P, const Triangle<R>* tris, u64 ntris, R* accelOut) {
    V3<R> x = in[i].pos;
    const R p = out[i].pressure;
    V3<R> v = in[i].vel;
    U3 c = cell_of(x, P.h);
    const U3 b = {c.x - 1u, c.y - 1u, c.z - 1u};
    const U3 e = {c.x + 2u, c.y + 2u, c.z + 2u};
    V3<R> Fp{0, 0, 0}, Fv{0, 0, 0}, Fps{0, 0, 0}, Fvs{0, 0, 0};
    const R rho = p / P.k + P.rho0;  // SF:544, density rebuilt from pressure
    for (u32 ox = b.x; ox != e.x; ++ox)
        for (u32 oy = b.y; oy != e.y; ++oy)
            for (u32 oz = b.z; oz != e.z; ++oz) {
                const u32 code = morton(ox, oy, oz);
                u32 bm = (u32)(code % M);
                u32 s = hashMap[bm], t = hashMap[bm + 1];
                for (u32 k = s; k < t; ++k) {
                    const u32 j = particleMap[k];
                    if (j == i) continue;
                    const R pj = out[j].pressure;  // this step's pressure pass (SF:564)
                    const V3<R> xj = in[j].pos, vj = in[j].vel;
                    V3<R> d = xj - x;
                    R r2 = dot3(d, d);
                    if (r2 > P.h * P.h) continue;
                    R r = std::sqrt(r2);
                    if (r2 == 0 || r == 0)
                        d = random_direction((R)i);
                    else
                        d = d / r;
                    Fp = Fp + d * (p + pj) * D1(r, P.h);
                    Fv = Fv + (vj - v) * D2(r, P.h);
                }
                if (ns == 0) continue;
                bm = (u32)(code % Ms);
                s = stHash[bm];
                t = stHash[bm + 1];
                for (u32 k = s; k < t; ++k) {
                    V3<R> d = st[k].pos - x;
                    R r2 = dot3(d, d);
                    if (r2 > P.h * P.h) continue;
                    R r = std::sqrt(r2);
                    if (r2 == 0 || r == 0)
                        d = random_direction((R)i);
                    else
                        d = d / r;
                    Fps = Fps + d * std::fabs(p) * D1(r, P.h);
                    Fvs = Fvs + (-v) * D2(r, P.h);
                }
            }
    Fp = Fp * (P.mass / (R(2) * rho) * P.C);
    Fv = Fv * (P.mu * P.mass * P.C);
    Fps = Fps * (P.mass / (R(2) * rho) * P.C * R(1));
    Fvs = Fvs * (P.mu * P.mass * P.C * R(0));  // SF:629, static viscosity switched off
    V3<R> F{0, 0, 0};
    F = F + Fp;
    F = F + Fv;
    F = F + Fps;
    F = F + Fvs;
    for (u64 t = 0; t < ntris; ++t) F = F + triangle_force(x, tris[t]);
    V3<R> a = F / rho + V3<R>{P.gx, P.gy, P.gz};
    if (accelOut) {
        accelOut[3 * i + 0] = a.x;
        accelOut[3 * i + 1] = a.y;
        accelOut[3 * i + 2] = a.z;
    }
    v = v + a * dt;  // SF:641-642, symplectic Euler
    x = x + v * dt;
    c = cell_of(x, P.h);
    const u32 hash = (u32)(morton(c.x, c.y, c.z) % M);
    out[i].pos = x;
    out[i].pressure = p;
    out[i].vel = v;
    out[i].hash = hash;
}

// ThreadParallelTaskManager.cpp:47-56: a mutex + condition-variable barrier
// that counts arrivals twice so that a fast thread cannot lap a slow one.
class Barrier {
  public:
    explicit Barrier(u64 n) : n_(n) {}
    void sync() {
        if (n_ <= 1) return;
        std::unique_lock<std::mutex> lk(m_);
        ++c1_;
        cv_.notify_all();
        cv_.wait(lk, [&] { return c1_ % n_ == 0; });
        ++c2_;
        cv_.notify_all();
        cv_.wait(lk, [&] { return c2_ % n_ == 0; });
    }

  private:
    u64 n_;
    u64 c1_ = 0, c2_ = 0;
    std::mutex m_;
    std::condition_variable cv_;
};

// A fixed pool of worker threads that run one task at a time on all threads
// (the reference's SimulationThreadFunc loop, without the 3-deep task queue:
// run() blocks until the task is done, so no update is ever dropped).
class Pool {
  public:
    explicit Pool(u64 n) : n_(n ? n : 1), barrier_(n ? n : 1) {
        for (u64 t = 1; t < n_; ++t) workers_.emplace_back([this, t] { loop(t); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> lk(m_);
            exit_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (auto& w : workers_) w.join();
    }
    u64 size() const { return n_; }
    Barrier& barrier() { return barrier_; }
    void run(const std::function<void(u64, u64)>& f) {
        {
            std::lock_guard<std::mutex> lk(m_);
            task_ = &f;
            pending_ = n_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        f(0, n_);
        std::unique_lock<std::mutex> lk(m_);
        done_cv_.wait(lk, [&] { return pending_ == 0; });
        task_ = nullptr;
    }

  private:
    void loop(u64 t) {
        u64 seen = 0;
        for (;;) {
            const std::function<void(u64, u64)>* f = nullptr;
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                if (exit_) return;
                f = task_;
            }
            (*f)(t, n_);
            {
                std::lock_guard<std::mutex> lk(m_);
                --pending_;
            }
            done_cv_.notify_all();
        }
    }
    u64 n_;
    Barrier barrier_;
    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(u64, u64)>* task_ = nullptr;
    u64 pending_ = 0, gen_ = 0;
    bool exit_ = false;
};

enum : int {
    FLAG_RECORD_DEBUG = 1,       // keep per-particle density and acceleration of the last step
    FLAG_PARALLEL_PARTIAL_SUM = 2,  // CPU:256-276 instead of the serial scan (P16)
};

template <class R>
class Engine {
  public:
    Params<R> P{};
    u64 n = 0, ns = 0, M = 0, Ms = 0;
    std::vector<DynP<R>> buf[3];  // the 3-buffer ring of CPU:491
    int cur = 0;
    AtomicTable hashMap, stHash;
    std::vector<u32> particleMap;
    std::vector<StatP<R>> st;
    std::vector<u32> origOfSlot, stOrigOfSlot;
    std::vector<Triangle<R>> tris;
    std::unique_ptr<Pool> pool;
    std::vector<std::atomic<u32>> groupSums;
    int flags = 0;
    float reorderInterval = FLT_MAX;
    float reorderElapsed = 0;
    double simTime = 0;
    std::vector<R> dens, acc;

    // CPU:396-419 (Initialize) with CPU:460-504. Parameters come in as the
    // scene's ParticleBehaviourParameters; C and selfDensity are derived here
    // in float exactly as CPU:406-407 does, then widened for R = double.
    Engine(const float* params, const float* pos, const float* vel, u64 n_, const float* spos, u64 ns_, u64 hashSize,
           int threads, int flags_)
        : n(n_), ns(ns_), flags(flags_) {
        P.mass = params[0];
        P.k = params[1];
        P.elasticity = params[2];
        P.mu = params[3];
        P.gx = params[4];
        P.gy = params[5];
        P.gz = params[6];
        P.rho0 = params[7];
        P.h = params[8];
        P.C = kernel_constant<R>(P.h);
        P.selfDensity = P.mass * D0<R>(R(0), P.h) * P.C;
        pool.reset(new Pool((u64)std::max(threads, 1)));
        if (flags & FLAG_PARALLEL_PARTIAL_SUM) {
            groupSums = std::vector<std::atomic<u32>>(pool->size());
        }
        init_static(spos);
        init_dynamic(pos, vel, hashSize);
    }

    // SE.h:79-114 + CPU:460-477: serial counting sort of the walls with M_s = N_s.
    // The scan runs over all N_s + 1 entries (the last one stays 0 before the
    // scan, so it ends as N_s); the scatter walks the input in order with a
    // pre-decrement, so equal keys land in descending input order.
    void init_static(const float* spos) {
        if (ns == 0) return;
        Ms = ns;
        stHash = AtomicTable(ns + 1);
        for (auto& a : stHash) a.store(0);
        std::vector<StatP<R>> in(ns);
        std::vector<u32> keys(ns);
        for (u64 i = 0; i < ns; ++i) {
            in[i].pos = {(R)spos[3 * i], (R)spos[3 * i + 1], (R)spos[3 * i + 2]};
            in[i].pressure = 0;
            U3 c = cell_of(in[i].pos, P.h);
            keys[i] = (u32)(morton(c.x, c.y, c.z) % Ms);
            ++stHash[keys[i]];
        }
        u32 sum = 0;
        for (auto& a : stHash) {
            sum += a.load();
            a.store(sum);
        }
        st.resize(ns);
        stOrigOfSlot.resize(ns);
        for (u64 i = 0; i < ns; ++i) {
            u32 slot = --stHash[keys[i]];
            st[slot] = in[i];
            stOrigOfSlot[slot] = (u32)i;
        }
    }

    // CPU:478-504 + CPU:55-131 (CalculateHashAndParticleMap).
    void init_dynamic(const float* pos, const float* vel, u64 hashSize) {
        if (n == 0) return;
        M = hashSize ? hashSize : 2 * n;
        hashMap = AtomicTable(M + 1);
        for (auto& a : hashMap) a.store(0);
        hashMap[M].store((u32)n);
        particleMap.assign(n, 0);
        for (auto& b : buf) b.assign(n, DynP<R>{});
        origOfSlot.assign(n, 0);
        DynP<R>* in = buf[0].data();
        for (u64 i = 0; i < n; ++i) {
            in[i].pos = {(R)pos[3 * i], (R)pos[3 * i + 1], (R)pos[3 * i + 2]};
            in[i].pressure = 0;
            in[i].vel = vel ? V3<R>{(R)vel[3 * i], (R)vel[3 * i + 1], (R)vel[3 * i + 2]} : V3<R>{0, 0, 0};
            in[i].hash = 0;
        }
        DynP<R>* outp = buf[1].data();
        Barrier& bar = pool->barrier();
        pool->run([&](u64 t, u64 T) {
            const u64 begin = n * t / T, end = n * (t + 1) / T;
            bar.sync();
            // SF:321-331 (ComputeDynamicParticlesHashAndPrepareHashMap)
            for (u64 i = begin; i < end; ++i) {
                U3 c = cell_of(in[i].pos, P.h);
                u32 key = (u32)(morton(c.x, c.y, c.z) % M);
                in[i].hash = key;
                hashMap[key]++;
            }
            bar.sync();
            if (t == 0) {  // CPU:98-103, serial inclusive scan over the first M entries
                u32 sum = 0;
                for (u64 b = 0; b < M; ++b) {
                    sum += hashMap[b].load();
                    hashMap[b].store(sum);
                }
            }
            bar.sync();
            // SF:332-345 (ReorderDynamicParticlesAndFinishHashMap)
            for (u64 i = begin; i < end; ++i) {
                u32 key = in[i].hash;
                u64 slot = (u64)(hashMap[key]-- - 1);
                particleMap[i] = (u32)i;
                outp[slot] = in[i];
                origOfSlot[slot] = (u32)i;
            }
            bar.sync();
        });
        cur = 1;
    }

    // CPU:420-459 (Update) + CPU:132-349 (SimulateParticlesTimeStep), run to
    // completion (the reference enqueues it on its workers; the caller here
    // always waits, which is what a timed run or a parity test needs).
    int update(float dt, u32 steps) {
        if (n == 0) return 0;
        const bool reorder = reorderElapsed > reorderInterval;
        if ((flags & FLAG_RECORD_DEBUG) && dens.size() != n) {
            dens.assign(n, R(0));
            acc.assign(3 * n, R(0));
        }
        Barrier& bar = pool->barrier();
        DynP<R>* inP = nullptr;
        DynP<R>* outP = nullptr;
        DynP<R>* ordP = nullptr;
        std::vector<u32> origNext;
        pool->run([&](u64 t, u64 T) {
            for (u32 step = 0; step < steps; ++step) {
                if (t == 0) {
                    inP = buf[cur].data();
                    cur = (cur + 1) % 3;
                    outP = buf[cur].data();
                    ordP = nullptr;
                }
                const u64 begin = n * t / T, end = n * (t + 1) / T;
                const u64 hb = M * t / T, he = M * (t + 1) / T;
                const bool last = step + 1 == steps;
                R* dOut = (last && (flags & FLAG_RECORD_DEBUG)) ? dens.data() : nullptr;
                R* aOut = (last && (flags & FLAG_RECORD_DEBUG)) ? acc.data() : nullptr;
                bar.sync();
                for (u64 i = begin; i < end; ++i)
                    update_pressure<R>(i, M, ns, Ms, inP, outP, hashMap.data(), particleMap.data(), st.data(),
                                       stHash.data(), P, dOut);
                bar.sync();
                for (u64 i = begin; i < end; ++i)
                    update_dynamics<R>(i, M, ns, Ms, inP, outP, hashMap.data(), particleMap.data(), st.data(),
                                       stHash.data(), (R)dt, P, tris.data(), tris.size(), aOut);
                bar.sync();
                for (u64 b = hb; b < he; ++b) hashMap[b].store(0);
                if (!groupSums.empty()) groupSums[t].store(0);
                bar.sync();
                for (u64 i = begin; i < end; ++i) hashMap[outP[i].hash]++;
                bar.sync();
                if (groupSums.empty()) {
                    if (t == 0) {  // CPU:245-253, serial scan on thread 0 (P16)
                        u32 sum = 0;
                        for (u64 b = 0; b < M; ++b) {
                            sum += hashMap[b].load();
                            hashMap[b].store(sum);
                        }
                    }
                } else {  // CPU:256-276
                    u32 sum = 0;
                    for (u64 b = hb; b < he; ++b) {
                        sum += hashMap[b].load();
                        hashMap[b].store(sum);
                    }
                    for (u64 g = t; g + 1 < T; ++g) groupSums[g] += sum;
                    bar.sync();
                    if (t != 0) {
                        u32 add = groupSums[t - 1].load();
                        for (u64 b = hb; b < he; ++b) hashMap[b] += add;
                    }
                }
                if (t == 0 && reorder) {
                    cur = (cur + 1) % 3;
                    ordP = buf[cur].data();
                    origNext.assign(n, 0);
                }
                bar.sync();
                if (reorder) {  // SF:332-345
                    for (u64 i = begin; i < end; ++i) {
                        u32 key = outP[i].hash;
                        u64 slot = (u64)(hashMap[key]-- - 1);
                        particleMap[i] = (u32)i;
                        ordP[slot] = outP[i];
                        origNext[slot] = origOfSlot[i];
                    }
                } else {  // SF:346-356 (FillDynamicParticleMapAndFinishHashMap)
                    for (u64 i = begin; i < end; ++i) {
                        u32 key = outP[i].hash;
                        u32 slot = hashMap[key]-- - 1;
                        particleMap[slot] = (u32)i;
                    }
                }
                bar.sync();
                if (t == 0 && reorder) {
                    if (dOut) permute_debug(origNext);
                    origOfSlot.swap(origNext);
                }
                bar.sync();
            }
        });
        simTime += (double)dt * steps;
        reorderElapsed += dt;
        if (reorderElapsed > reorderInterval) reorderElapsed = 0;
        return 0;
    }

    // Debug arrays are indexed by the slot a particle had during the step; when
    // the step also reordered memory, move them to the new slots.
    void permute_debug(const std::vector<u32>& origNext) {
        std::vector<u32> slotOfOrig(n);
        for (u64 s = 0; s < n; ++s) slotOfOrig[origNext[s]] = (u32)s;
        std::vector<R> d2(n), a2(3 * n);
        for (u64 s = 0; s < n; ++s) {
            u32 ns2 = slotOfOrig[origOfSlot[s]];
            d2[ns2] = dens[s];
            for (int k = 0; k < 3; ++k) a2[3 * ns2 + k] = acc[3 * s + k];
        }
        dens.swap(d2);
        acc.swap(a2);
    }

    const DynP<R>* current() const { return buf[cur].data(); }
};

struct Handle {
    int fp64 = 0;
    std::unique_ptr<Engine<float>> f;
    std::unique_ptr<Engine<double>> d;
};

}  // namespace sphoracle

using namespace sphoracle;

extern "C" {

// params: the 11 floats of ParticleBehaviourParameters (Particles.h:23-41);
// entries 9 and 10 (selfDensity, smoothingKernelConstant) are ignored and
// derived. hash_size = 0 selects the CPU engine's rule M = 2N (CPU:486).
void* sph_oracle_create(const float* params, const float* pos, const float* vel, uint64_t n, const float* spos,
                        uint64_t ns, uint64_t hash_size, int threads, int fp64, int flags) {
    Handle* H = new Handle;
    H->fp64 = fp64;
    if (fp64)
        H->d.reset(new Engine<double>(params, pos, vel, n, spos, ns, hash_size, threads, flags));
    else
        H->f.reset(new Engine<float>(params, pos, vel, n, spos, ns, hash_size, threads, flags));
    return H;
}

void sph_oracle_destroy(void* h) { delete (Handle*)h; }

int sph_oracle_set_triangles(void* h, const float* tri9, uint64_t count) {
    Handle* H = (Handle*)h;
    auto load = [&](auto& E) {
        using R = decltype(E.P.h);
        E.tris.resize(count);
        for (uint64_t t = 0; t < count; ++t) {
            const float* q = tri9 + 9 * t;
            E.tris[t].p1 = {(R)q[0], (R)q[1], (R)q[2]};
            E.tris[t].p2 = {(R)q[3], (R)q[4], (R)q[5]};
            E.tris[t].p3 = {(R)q[6], (R)q[7], (R)q[8]};
        }
        return 0;
    };
    return H->fp64 ? load(*H->d) : load(*H->f);
}

int sph_oracle_set_reorder_interval(void* h, float interval) {
    Handle* H = (Handle*)h;
    if (H->fp64)
        H->d->reorderInterval = interval;
    else
        H->f->reorderInterval = interval;
    return 0;
}

int sph_oracle_update(void* h, float dt, uint32_t steps) {
    Handle* H = (Handle*)h;
    return H->fp64 ? H->d->update(dt, steps) : H->f->update(dt, steps);
}

uint64_t sph_oracle_hash_size(void* h) {
    Handle* H = (Handle*)h;
    return H->fp64 ? H->d->M : H->f->M;
}

double sph_oracle_simulation_time(void* h) {
    Handle* H = (Handle*)h;
    return H->fp64 ? H->d->simTime : H->f->simTime;
}

// State by original particle index (the order of the arrays given to create).
int sph_oracle_read(void* h, double* pos, double* vel, double* pressure) {
    Handle* H = (Handle*)h;
    auto rd = [&](auto& E) {
        const auto* P = E.current();
        for (uint64_t s = 0; s < E.n; ++s) {
            uint32_t o = E.origOfSlot[s];
            if (pos) {
                pos[3 * o] = P[s].pos.x;
                pos[3 * o + 1] = P[s].pos.y;
                pos[3 * o + 2] = P[s].pos.z;
            }
            if (vel) {
                vel[3 * o] = P[s].vel.x;
                vel[3 * o + 1] = P[s].vel.y;
                vel[3 * o + 2] = P[s].vel.z;
            }
            if (pressure) pressure[o] = P[s].pressure;
        }
        return 0;
    };
    return H->fp64 ? rd(*H->d) : rd(*H->f);
}

// OfflineCPUParticleBufferManager::LockWrite of the current buffer
// (OfflineCPUParticleBufferManager.cpp:68-74) and a caller writing positions
// and/or velocities (original order; either may be null) into its records: the
// stored hash, the pressure, the hash map and particleMap stay as they are, so
// the next update walks the buckets of the last step (CPU:132-349).
int sph_oracle_write(void* h, const float* pos, const float* vel) {
    Handle* H = (Handle*)h;
    auto wr = [&](auto& E) {
        using R = decltype(E.P.h);
        auto* P = E.buf[E.cur].data();
        for (uint64_t s = 0; s < E.n; ++s) {
            const uint32_t o = E.origOfSlot[s];
            if (pos) P[s].pos = {(R)pos[3 * o], (R)pos[3 * o + 1], (R)pos[3 * o + 2]};
            if (vel) P[s].vel = {(R)vel[3 * o], (R)vel[3 * o + 1], (R)vel[3 * o + 2]};
        }
        return 0;
    };
    return H->fp64 ? wr(*H->d) : wr(*H->f);
}

// Stored bucket key (velocityAndHash.w) by original particle index.
int sph_oracle_read_keys(void* h, uint32_t* keys) {
    Handle* H = (Handle*)h;
    auto rd = [&](auto& E) {
        const auto* P = E.current();
        for (uint64_t s = 0; s < E.n; ++s) keys[E.origOfSlot[s]] = P[s].hash;
        return 0;
    };
    return H->fp64 ? rd(*H->d) : rd(*H->f);
}

// The M + 1 entry dynamic table (bucket starts, sentinel N).
int sph_oracle_read_hash_map(void* h, uint32_t* table) {
    Handle* H = (Handle*)h;
    auto rd = [&](auto& E) {
        for (uint64_t b = 0; b <= E.M; ++b) table[b] = E.hashMap[b].load();
        return 0;
    };
    return H->fp64 ? rd(*H->d) : rd(*H->f);
}

// particleMap expressed as original particle indices: entry k is the original
// index of the k-th particle in bucket order.
int sph_oracle_read_bucket_ids(void* h, uint32_t* ids) {
    Handle* H = (Handle*)h;
    auto rd = [&](auto& E) {
        for (uint64_t k = 0; k < E.n; ++k) ids[k] = E.origOfSlot[E.particleMap[k]];
        return 0;
    };
    return H->fp64 ? rd(*H->d) : rd(*H->f);
}

// Static table (M_s + 1 entries) and the original wall index in sorted order.
int sph_oracle_read_static(void* h, uint32_t* table, uint32_t* ids) {
    Handle* H = (Handle*)h;
    auto rd = [&](auto& E) {
        if (table)
            for (uint64_t b = 0; b <= E.Ms && E.ns; ++b) table[b] = E.stHash[b].load();
        if (ids)
            for (uint64_t s = 0; s < E.ns; ++s) ids[s] = E.stOrigOfSlot[s];
        return 0;
    };
    return H->fp64 ? rd(*H->d) : rd(*H->f);
}

// Density and acceleration of the last step of the last update (needs
// FLAG_RECORD_DEBUG), by original particle index.
int sph_oracle_read_debug(void* h, double* density, double* accel) {
    Handle* H = (Handle*)h;
    auto rd = [&](auto& E) {
        if (E.dens.size() != E.n) return -2;
        for (uint64_t s = 0; s < E.n; ++s) {
            uint32_t o = E.origOfSlot[s];
            if (density) density[o] = E.dens[s];
            if (accel)
                for (int k = 0; k < 3; ++k) accel[3 * o + k] = E.acc[3 * s + k];
        }
        return 0;
    };
    return H->fp64 ? rd(*H->d) : rd(*H->f);
}

// FP64 {mass, px, py, pz, kinetic energy, cx, cy, cz} of the current state.
int sph_oracle_conserved(void* h, double* out) {
    Handle* H = (Handle*)h;
    auto rd = [&](auto& E) {
        const auto* P = E.current();
        double m = (double)E.P.mass, M0 = 0, px = 0, py = 0, pz = 0, ke = 0, cx = 0, cy = 0, cz = 0;
        for (uint64_t s = 0; s < E.n; ++s) {
            double vx = P[s].vel.x, vy = P[s].vel.y, vz = P[s].vel.z;
            M0 += m;
            px += m * vx;
            py += m * vy;
            pz += m * vz;
            ke += 0.5 * m * (vx * vx + vy * vy + vz * vz);
            cx += m * (double)P[s].pos.x;
            cy += m * (double)P[s].pos.y;
            cz += m * (double)P[s].pos.z;
        }
        out[0] = M0;
        out[1] = px;
        out[2] = py;
        out[3] = pz;
        out[4] = ke;
        out[5] = M0 > 0 ? cx / M0 : 0;
        out[6] = M0 > 0 ? cy / M0 : 0;
        out[7] = M0 > 0 ? cz / M0 : 0;
        return 0;
    };
    return H->fp64 ? rd(*H->d) : rd(*H->f);
}

// ---- free functions for the known-answer tests ----

void sph_oracle_cells(const float* pos, uint64_t n, float h, uint32_t* cells) {
    for (uint64_t i = 0; i < n; ++i) {
        U3 c = cell_of(V3<float>{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}, h);
        cells[3 * i] = c.x;
        cells[3 * i + 1] = c.y;
        cells[3 * i + 2] = c.z;
    }
}

uint32_t sph_oracle_morton(uint32_t x, uint32_t y, uint32_t z) { return morton(x, y, z); }

void sph_oracle_keys(const float* pos, uint64_t n, float h, uint64_t M, uint32_t* keys) {
    for (uint64_t i = 0; i < n; ++i) {
        U3 c = cell_of(V3<float>{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}, h);
        keys[i] = (uint32_t)(morton(c.x, c.y, c.z) % M);
    }
}

// out = {D0(r,h), D1(r,h), D2(r,h), C(h)} in float.
void sph_oracle_kernels(float r, float h, float* out) {
    out[0] = D0<float>(r, h);
    out[1] = D1<float>(r, h);
    out[2] = D2<float>(r, h);
    out[3] = kernel_constant<float>(h);
}

void sph_oracle_random_direction(float x, float* out) {
    V3<float> d = random_direction<float>(x);
    out[0] = d.x;
    out[1] = d.y;
    out[2] = d.z;
}

float sph_oracle_noise(float x) { return noise<float>(x); }

void sph_oracle_triangle_force(const float* p, const float* tri9, float* out) {
    Triangle<float> t{{tri9[0], tri9[1], tri9[2]}, {tri9[3], tri9[4], tri9[5]}, {tri9[6], tri9[7], tri9[8]}};
    V3<float> f = triangle_force<float>(V3<float>{p[0], p[1], p[2]}, t);
    out[0] = f.x;
    out[1] = f.y;
    out[2] = f.z;
}

}  // extern "C"

// SimulationEngineHIP: the reference's engine interface, implemented over the C ABI.
//
// Mirrors, for the step path of Murk42/SPHFluidSimulation:
//   SPH::SimulationEngine        SPHFluidSimulation/include/SPH/Core/SimulationEngine.h:35-48
//   SPH::ParticleBufferManager   SPHFluidSimulation/include/SPH/Core/ParticleBufferManager.h:8-25
//   SPH::DynamicParticle / StaticParticle / ParticleBehaviourParameters
//                                SPHFluidSimulation/include/SPH/Core/Particles.h:9-41
// A caller of the reference (CPUSimulation / GPUSimulation in
// FluidSimulationProject/source/Scenes/SimulationVisualisationScene/
// SimulationVisualisationScene.cpp:8-124) switches engines by constructing
// SimulationEngineHIP instead of SimulationEngineCPU/GPU; Initialize, Update,
// GetSimulationTime and SystemImplementationName keep their meaning. The
// reference's SceneBlueprint generates particles from JSON; here the scene is
// handed over as arrays (SceneArrays), the output of that generation step.
//
// Header-only; link against sphfluidsimulation_amd/libsph_hip.so.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../sph_hip.h"

namespace SPH {

struct Vec3f {
    float x, y, z;
};

// Particles.h:9-17: the 32-byte record the reference's buffers hold.
struct DynamicParticle {
    Vec3f position;
    float pressure;
    Vec3f velocity;
    uint32_t hash;
};
static_assert(sizeof(DynamicParticle) == 32, "DynamicParticle must stay 32 bytes");

// Particles.h:19-23
struct StaticParticle {
    Vec3f position;
    float pressure;
};
static_assert(sizeof(StaticParticle) == 16, "StaticParticle must stay 16 bytes");

using ParticleBehaviourParameters = sph_params;  // Particles.h:23-41, same field order

// What SceneBlueprint::GenerateLayerParticles("dynamic" / "static") produces
// (SceneBlueprint.h:23-51), plus GetSystemParameters().particleBehaviourParameters.
struct SceneArrays {
    ParticleBehaviourParameters parameters{};
    std::vector<Vec3f> dynamicPositions;
    std::vector<Vec3f> dynamicVelocities;  // empty: zero, as the reference's generators leave it
    std::vector<Vec3f> staticPositions;
    std::vector<float> triangles;          // 9 floats per triangle (SceneBlueprint::GetMesh)
};

// ParticleBufferManager.h:8-25, reduced to what a reader of an offline engine
// uses: the particle count and a read lock on the current buffer. LockRead
// returns the current state as 32-byte records in original particle order
// (the reference returns memory order, which it never exposes as an order).
class ParticleBufferManager {
  public:
    virtual ~ParticleBufferManager() = default;
    virtual uint64_t GetParticleCount() const = 0;
    virtual uint64_t GetParticleSize() const { return sizeof(DynamicParticle); }
    virtual const std::vector<DynamicParticle>& LockRead() = 0;
    virtual void FlushAllOperations() = 0;
};

// SimulationEngine.h:35-48
class SimulationEngine {
  public:
    virtual ~SimulationEngine() = default;
    virtual void Clear() = 0;
    virtual void Initialize(const SceneArrays& scene) = 0;
    virtual void Update(float dt, uint32_t simulationStepCount) = 0;
    virtual const char* SystemImplementationName() = 0;
    virtual ParticleBufferManager* GetDynamicParticlesBufferManager() = 0;
    virtual float GetSimulationTime() = 0;
};

class SimulationEngineHIP;

// Buffer manager of SimulationEngineHIP: the HIP form of the reference's
// renderable GPU buffer manager (RenderableGPUParticleBufferManager.cpp:190-270,
// OpenCLResourceLock.cpp:73-163). Advance() enqueues, behind the steps already
// enqueued, a copy of the state into the next of three pinned host buffers and
// returns at once; LockRead() waits for the newest copy (enqueuing one first if
// the state changed since) and returns it as 32-byte records; TryLockRead() is
// the non-blocking form. A locked view stays valid until its slot comes round
// again, three Advance() calls later.
class HIPAsyncBufferManager final : public ParticleBufferManager {
  public:
    explicit HIPAsyncBufferManager(sph_handle h) : h_(h) {}
    uint64_t GetParticleCount() const override { return sph_particle_count(h_); }

    void Advance() {
        uint32_t slot = 0;
        check(sph_readback_enqueue(h_, &slot));
        slot_ = static_cast<int>(slot);
        stale_ = false;
    }
    const std::vector<DynamicParticle>& LockRead() override {
        if (slot_ < 0 || stale_) Advance();
        return view(static_cast<uint32_t>(slot_));
    }
    // Returns false (and leaves `out` alone) while the newest copy is in flight.
    bool TryLockRead(const std::vector<DynamicParticle>*& out) {
        if (slot_ < 0 || stale_) Advance();
        int ready = 0;
        check(sph_readback_query(h_, static_cast<uint32_t>(slot_), &ready));
        if (!ready) return false;
        out = &view(static_cast<uint32_t>(slot_));
        return true;
    }
    double LockedSimulationTime() const { return lockedTime_; }
    void FlushAllOperations() override { check(sph_synchronize(h_)); }
    void MarkStale() { stale_ = true; }  // the engine enqueued steps
    void Reset() {
        slot_ = -1;
        stale_ = true;
    }

  private:
    const std::vector<DynamicParticle>& view(uint32_t slot) {
        const float *pos = nullptr, *vel = nullptr, *p = nullptr;
        const uint32_t* keys = nullptr;
        uint64_t n = 0;
        check(sph_readback_lock(h_, slot, &pos, &vel, &p, &keys, &n, &lockedTime_));
        records_.resize(n);
        for (uint64_t i = 0; i < n; ++i)
            records_[i] = DynamicParticle{{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]},
                                          p[i],
                                          {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]},
                                          keys[i]};
        return records_;
    }
    void check(int rc) const {
        if (rc != SPH_OK) throw std::runtime_error(std::string("sph: ") + sph_last_error(h_));
    }
    sph_handle h_;
    int slot_ = -1;
    bool stale_ = true;
    double lockedTime_ = 0.0;
    std::vector<DynamicParticle> records_;
};

// The engine. Errors from the device raise std::runtime_error (the reference
// logs them); Update before Initialize is a no-op with a warning in the
// reference (SimulationEngineCPU.cpp:422-426) and raises here.
class SimulationEngineHIP final : public SimulationEngine {
  public:
    explicit SimulationEngineHIP(int device = 0, uint32_t flags = 0) : flags_(flags) {
        if (sph_create(&h_, device) != SPH_OK) throw std::runtime_error("sph_create: no HIP device");
        manager_ = new HIPAsyncBufferManager(h_);
    }
    ~SimulationEngineHIP() override {
        delete manager_;
        sph_destroy(h_);
    }
    SimulationEngineHIP(const SimulationEngineHIP&) = delete;
    SimulationEngineHIP& operator=(const SimulationEngineHIP&) = delete;

    void Clear() override {
        check(sph_clear(h_));
        manager_->Reset();
    }

    void Initialize(const SceneArrays& scene) override {
        const uint64_t n = scene.dynamicPositions.size();
        const uint64_t ns = scene.staticPositions.size();
        const float* vel = scene.dynamicVelocities.empty()
                               ? nullptr
                               : reinterpret_cast<const float*>(scene.dynamicVelocities.data());
        check(sph_initialize(h_, &scene.parameters, reinterpret_cast<const float*>(scene.dynamicPositions.data()),
                             vel, n, ns ? reinterpret_cast<const float*>(scene.staticPositions.data()) : nullptr, ns,
                             flags_));
        if (!scene.triangles.empty())
            check(sph_set_triangles(h_, 0, scene.triangles.data(), scene.triangles.size() / 9));
        manager_->Reset();
    }

    void Update(float dt, uint32_t steps) override {
        check(sph_update(h_, dt, steps));
        manager_->MarkStale();
    }
    const char* SystemImplementationName() override { return sph_system_implementation_name(); }
    ParticleBufferManager* GetDynamicParticlesBufferManager() override { return manager_; }
    HIPAsyncBufferManager* GetAsyncBufferManager() { return manager_; }
    float GetSimulationTime() override { return sph_simulation_time(h_); }

    sph_handle handle() const { return h_; }

  private:
    void check(int rc) {
        if (rc != SPH_OK) throw std::runtime_error(std::string("sph: ") + sph_last_error(h_));
    }
    sph_handle h_ = nullptr;
    uint32_t flags_ = 0;
    HIPAsyncBufferManager* manager_ = nullptr;
};

}  // namespace SPH

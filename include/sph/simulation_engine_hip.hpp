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

// Readback manager of SimulationEngineHIP (RenderableGPUParticleBufferManager.cpp:256-270
// reads the current buffer back to the host; this one does so on LockRead).
class HIPReadbackBufferManager final : public ParticleBufferManager {
  public:
    explicit HIPReadbackBufferManager(sph_handle h) : h_(h) {}
    uint64_t GetParticleCount() const override { return sph_particle_count(h_); }
    const std::vector<DynamicParticle>& LockRead() override {
        const uint64_t n = GetParticleCount();
        pos_.resize(3 * n);
        vel_.resize(3 * n);
        p_.resize(n);
        keys_.resize(n);
        records_.resize(n);
        if (n) {
            check(sph_read(h_, pos_.data(), vel_.data(), p_.data()));
            check(sph_read_keys(h_, keys_.data()));
        }
        for (uint64_t i = 0; i < n; ++i)
            records_[i] = DynamicParticle{{pos_[3 * i], pos_[3 * i + 1], pos_[3 * i + 2]},
                                          p_[i],
                                          {vel_[3 * i], vel_[3 * i + 1], vel_[3 * i + 2]},
                                          keys_[i]};
        return records_;
    }
    void FlushAllOperations() override { check(sph_synchronize(h_)); }

  private:
    void check(int rc) {
        if (rc != SPH_OK) throw std::runtime_error(std::string("sph: ") + sph_last_error(h_));
    }
    sph_handle h_;
    std::vector<float> pos_, vel_, p_;
    std::vector<uint32_t> keys_;
    std::vector<DynamicParticle> records_;
};

// The engine. Errors from the device raise std::runtime_error (the reference
// logs them); Update before Initialize is a no-op with a warning in the
// reference (SimulationEngineCPU.cpp:422-426) and raises here.
class SimulationEngineHIP final : public SimulationEngine {
  public:
    explicit SimulationEngineHIP(int device = 0, uint32_t flags = 0) : flags_(flags) {
        if (sph_create(&h_, device) != SPH_OK) throw std::runtime_error("sph_create: no HIP device");
        manager_ = new HIPReadbackBufferManager(h_);
    }
    ~SimulationEngineHIP() override {
        delete manager_;
        sph_destroy(h_);
    }
    SimulationEngineHIP(const SimulationEngineHIP&) = delete;
    SimulationEngineHIP& operator=(const SimulationEngineHIP&) = delete;

    void Clear() override { check(sph_clear(h_)); }

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
    }

    void Update(float dt, uint32_t steps) override { check(sph_update(h_, dt, steps)); }
    const char* SystemImplementationName() override { return sph_system_implementation_name(); }
    ParticleBufferManager* GetDynamicParticlesBufferManager() override { return manager_; }
    float GetSimulationTime() override { return sph_simulation_time(h_); }

    sph_handle handle() const { return h_; }

  private:
    void check(int rc) {
        if (rc != SPH_OK) throw std::runtime_error(std::string("sph: ") + sph_last_error(h_));
    }
    sph_handle h_ = nullptr;
    uint32_t flags_ = 0;
    HIPReadbackBufferManager* manager_ = nullptr;
};

}  // namespace SPH

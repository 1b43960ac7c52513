// Declaration of the engine interface of
// SPHFluidSimulation/include/SPH/Core/SimulationEngine.h:13-48:
// ParticleSimulationParameters and the virtual members of SimulationEngine. Its
// static helpers (GetCell, GetHash, the smoothing kernels, the static sort and the
// debug validators, SimulationEngine.h:50-76) are left out: no engine overrides
// them and the drop-in engine does not call them.
#pragma once
#include "SPH/Core/Particles.h"

namespace SPH {
class SceneBlueprint;

struct ParticleSimulationParameters {
    ParticleBehaviourParameters particleBehaviourParameters;
    Map<String, String> otherParameters;

    bool ParseParameter(StringView name, float& value) const;
    bool ParseParameter(StringView name, bool& value) const;
};

class ParticleBufferManager;

class SimulationEngine {
  public:
    virtual ~SimulationEngine() {}

    virtual void Clear() = 0;
    virtual void Initialize(SceneBlueprint& scene, ParticleBufferManager& dynamicParticlesBufferManager,
                            ParticleBufferManager& staticParticlesBufferManager) = 0;
    virtual void Update(float dt, uint simulationStepCount) = 0;

    virtual StringView SystemImplementationName() = 0;
    virtual ParticleBufferManager* GetDynamicParticlesBufferManager() = 0;
    virtual ParticleBufferManager* GetStaticParticlesBufferManager() = 0;

    virtual float GetSimulationTime() = 0;
};
}  // namespace SPH

// Declaration of the members of SPHFluidSimulation/include/SPH/Core/SceneBlueprint.h:9-59
// that engines call (GetSystemParameters, GetMesh, GenerateLayerParticles). The
// reference fills a blueprint from a scene file (LoadScene) and its particle-set
// generators; here the test fills each layer's positions directly (SetLayer,
// SetParameters, SetOtherParameter, test-only), and GenerateLayerParticles
// appends a layer's positions to `particles` as the reference's does.
#pragma once
#include "SPH/Core/ParticleBufferManager.h"
#include "SPH/Core/SimulationEngine.h"

namespace SPH {
class SceneBlueprint {
  public:
    ParticleSimulationParameters GetSystemParameters() const { return systemParameters; }
    const Graphics::BasicIndexedMesh& GetMesh() const { return mesh; }

    template <typename Particle>
    void GenerateLayerParticles(StringView layerName, Array<Particle>& particles) {
        const std::vector<Vec3f>* layer = layers.Find(String(layerName.str()));
        if (!layer) return;
        const uintMem base = particles.Count();
        particles.Resize(base + layer->size());
        for (uintMem i = 0; i < layer->size(); ++i) particles[base + i].position = (*layer)[i];
    }

    // test-only
    void SetLayer(StringView layerName, std::vector<Vec3f> positions) {
        layers.Insert(String(layerName.str()), std::move(positions));
    }
    void SetParameters(const ParticleBehaviourParameters& p) { systemParameters.particleBehaviourParameters = p; }
    void SetOtherParameter(StringView name, StringView value) {
        systemParameters.otherParameters.Insert(String(name.str()), String(value.str()));
    }
    Graphics::BasicIndexedMesh& MutableMesh() { return mesh; }

  private:
    ParticleSimulationParameters systemParameters;
    Map<String, std::vector<Vec3f>> layers;
    Graphics::BasicIndexedMesh mesh;
};
}  // namespace SPH

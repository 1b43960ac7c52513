// Declarations of SPHFluidSimulation/include/SPH/Core/Particles.h:9-41 (same
// members, order and layout), for tests/cpp/reference_binding.cpp.
#pragma once
#include "blaze_min.h"

using namespace Blaze;

namespace SPH {
struct DynamicParticle {
    Vec3f position;
    float pressure;
    Vec3f velocity;
    uint32 hash;
};
struct StaticParticle {
    Vec3f position;
    float pressure;
};
struct ParticleBehaviourParameters {
    float particleMass = 0.0f;
    float gasConstant = 0.0f;
    float elasticity = 0.0f;
    float viscosity = 0.0f;
    float gravityX = 0.0f;
    float gravityY = 0.0f;
    float gravityZ = 0.0f;
    float restDensity = 0.0f;
    float maxInteractionDistance = 0.0f;
    float selfDensity = 0.0f;
    float smoothingKernelConstant = 0.0f;
};
}  // namespace SPH

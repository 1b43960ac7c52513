#pragma once
// Drop-in file for the reference tree (SPHFluidSimulation/include/SPH/SimulationEngines/),
// beside SimulationEngineCPU.h and SimulationEngineGPU.h. Needs -I<this repo>/include for
// sph_hip.h and links libsph_hip.so (the MI355X engine behind the C ABI).
#include "SPH/Core/SimulationEngine.h"
#include "SPH/Core/ParticleBufferManager.h"
#include "SPH/Core/SceneBlueprint.h"
#include "SPH/ParticleBufferManagers/HIPParticleBufferManager.h"
#include "sph_hip.h"
using namespace Blaze;

namespace SPH
{
	// The third engine behind SPH::SimulationEngine (SimulationEngine.h:35-48), beside
	// SimulationEngineCPU (SimulationEngineCPU.cpp:351-504) and SimulationEngineGPU:
	// the step runs on one HIP device. Initialize takes the scene and the two buffer
	// managers as the other engines do (SimulationEngineCPU.cpp:396-419): it generates
	// the "dynamic" and "static" layers, reads the scene's reorderTimeInterval, hands
	// the particles to the device and Allocates the managers with them (3 buffers for
	// the dynamic particles, 1 for the static ones, SimulationEngineCPU.cpp:470, 491).
	// The managers must be HIPParticleBufferManager, as SimulationEngineGPU needs its
	// OpenCL managers. Update is asynchronous, like the other engines', and never drops
	// an update. Errors are logged, as the reference's engines log theirs.
	class SimulationEngineHIP : public SimulationEngine
	{
	public:
		// flags: SPH_FLAG_* of sph_hip.h (hipGraph replay, neighbour-pass variant, mixed precision).
		SimulationEngineHIP(int device = 0, uint32 flags = 0);
		~SimulationEngineHIP();

		void Clear() override;
		void Initialize(SceneBlueprint& scene, ParticleBufferManager& dynamicParticlesBufferManager, ParticleBufferManager& staticParticlesBufferManager) override;
		void Update(float dt, uint simulationSteps) override;

		StringView SystemImplementationName() override { return "HIP"; }
		ParticleBufferManager* GetDynamicParticlesBufferManager() override { return dynamicParticlesBufferManager; }
		ParticleBufferManager* GetStaticParticlesBufferManager() override { return staticParticlesBufferManager; }

		float GetSimulationTime() override;

		// The engine's C-ABI handle (sph_hip.h), for the calls beyond SimulationEngine
		// (sph_autotune, sph_conserved, sph_validate, sph_stage_times).
		sph_handle Handle() const { return handle; }
	private:
		sph_handle handle;
		uint32 flags;
		bool initialized;
		HIPParticleBufferManager* dynamicParticlesBufferManager;
		HIPParticleBufferManager* staticParticlesBufferManager;
	};
}

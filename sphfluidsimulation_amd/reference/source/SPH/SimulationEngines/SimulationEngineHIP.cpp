// Drop-in file for the reference tree (SPHFluidSimulation/source/SPH/SimulationEngines/).
#include "pch.h"
#include "SPH/Core/SceneBlueprint.h"
#include "SPH/SimulationEngines/SimulationEngineHIP.h"

#include <cfloat>
#include <cstring>
#include <vector>

namespace SPH
{
	static_assert(sizeof(ParticleBehaviourParameters) == sizeof(sph_params), "ParticleBehaviourParameters is the 11 floats of sph_params");

	SimulationEngineHIP::SimulationEngineHIP(int device, uint32 flags)
		: handle(nullptr), flags(flags), initialized(false),
		dynamicParticlesBufferManager(nullptr), staticParticlesBufferManager(nullptr)
	{
		if (sph_create(&handle, device) != SPH_OK)
		{
			handle = nullptr;
			Debug::Logger::LogFatal("SPH Library", "SimulationEngineHIP: no HIP device");
		}
	}
	SimulationEngineHIP::~SimulationEngineHIP()
	{
		Clear();
		if (handle != nullptr)
			sph_destroy(handle);
	}
	void SimulationEngineHIP::Clear()
	{
		if (dynamicParticlesBufferManager != nullptr)
		{
			dynamicParticlesBufferManager->FlushAllOperations();
			dynamicParticlesBufferManager->BindEngine(nullptr, true);
			dynamicParticlesBufferManager = nullptr;
		}
		if (staticParticlesBufferManager != nullptr)
		{
			staticParticlesBufferManager->FlushAllOperations();
			staticParticlesBufferManager->BindEngine(nullptr, false);
			staticParticlesBufferManager = nullptr;
		}
		if (handle != nullptr)
			sph_clear(handle);
		initialized = false;
	}
	void SimulationEngineHIP::Initialize(SceneBlueprint& scene, ParticleBufferManager& dynamicManager, ParticleBufferManager& staticManager)
	{
		Clear();
		if (handle == nullptr)
		{
			Debug::Logger::LogFatal("SPH Library", "SimulationEngineHIP: no HIP device");
			return;
		}
		HIPParticleBufferManager* dyn = dynamic_cast<HIPParticleBufferManager*>(&dynamicManager);
		HIPParticleBufferManager* st = dynamic_cast<HIPParticleBufferManager*>(&staticManager);
		if (dyn == nullptr || st == nullptr)
		{
			Debug::Logger::LogFatal("SPH Library", "SimulationEngineHIP needs HIPParticleBufferManager buffer managers");
			return;
		}

		ParticleSimulationParameters parameters = scene.GetSystemParameters();
		float reorderTimeInterval = FLT_MAX;
		parameters.ParseParameter("reorderTimeInterval", reorderTimeInterval);

		Array<DynamicParticle> dynamicParticles;
		Array<StaticParticle> staticParticles;
		scene.GenerateLayerParticles("dynamic", dynamicParticles);
		scene.GenerateLayerParticles("static", staticParticles);

		const uintMem n = dynamicParticles.Count(), ns = staticParticles.Count();
		std::vector<float> pos(3 * n), vel(3 * n), spos(3 * ns);
		for (uintMem i = 0; i < n; ++i)
		{
			pos[3 * i] = dynamicParticles[i].position.x;
			pos[3 * i + 1] = dynamicParticles[i].position.y;
			pos[3 * i + 2] = dynamicParticles[i].position.z;
			vel[3 * i] = dynamicParticles[i].velocity.x;
			vel[3 * i + 1] = dynamicParticles[i].velocity.y;
			vel[3 * i + 2] = dynamicParticles[i].velocity.z;
		}
		for (uintMem i = 0; i < ns; ++i)
		{
			spos[3 * i] = staticParticles[i].position.x;
			spos[3 * i + 1] = staticParticles[i].position.y;
			spos[3 * i + 2] = staticParticles[i].position.z;
		}
		sph_params params;
		memcpy(&params, &parameters.particleBehaviourParameters, sizeof(params));
		if (sph_initialize(handle, &params, pos.data(), vel.data(), n, ns ? spos.data() : nullptr, ns, flags) != SPH_OK)
		{
			Debug::Logger::LogError("SPH Library", "SimulationEngineHIP: initialize failed");
			Debug::Logger::LogError("SPH Library", sph_last_error(handle));
			return;
		}
		if (sph_set_reorder_interval(handle, reorderTimeInterval) != SPH_OK)
		{
			Debug::Logger::LogError("SPH Library", sph_last_error(handle));
			return;
		}

		// SimulationEngineCPU.cpp:418: the scene's boundary mesh.
		auto triangles = scene.GetMesh().CreateTriangleArray();
		if (!triangles.Empty())
		{
			std::vector<float> tri9(9 * triangles.Count());
			for (uintMem t = 0; t < triangles.Count(); ++t)
			{
				const Vec3f* p[3] = { &triangles[t].p1, &triangles[t].p2, &triangles[t].p3 };
				for (int k = 0; k < 3; ++k)
				{
					tri9[9 * t + 3 * k] = p[k]->x;
					tri9[9 * t + 3 * k + 1] = p[k]->y;
					tri9[9 * t + 3 * k + 2] = p[k]->z;
				}
			}
			if (sph_set_triangles(handle, 0, tri9.data(), triangles.Count()) != SPH_OK)
			{
				Debug::Logger::LogError("SPH Library", sph_last_error(handle));
				return;
			}
		}

		dyn->Allocate(sizeof(DynamicParticle), n, dynamicParticles.Ptr(), 3);
		st->Allocate(sizeof(StaticParticle), ns, staticParticles.Ptr(), 1);
		dyn->BindEngine(handle, true);
		st->BindEngine(handle, false);
		dynamicParticlesBufferManager = dyn;
		staticParticlesBufferManager = st;
		initialized = true;
	}
	void SimulationEngineHIP::Update(float dt, uint simulationSteps)
	{
		if (!initialized)
		{
			Debug::Logger::LogWarning("Client", "Updating an uninitialized SimulationEngineHIP");
			return;
		}
		if (dynamicParticlesBufferManager->GetParticleCount() == 0)
			return;
		if (sph_update(handle, dt, simulationSteps) != SPH_OK)
		{
			Debug::Logger::LogError("SPH Library", sph_last_error(handle));
			return;
		}
		dynamicParticlesBufferManager->MarkStale();
	}
	float SimulationEngineHIP::GetSimulationTime()
	{
		return handle != nullptr ? sph_simulation_time(handle) : 0.0f;
	}
}

#pragma once
// Drop-in file for the reference tree (SPHFluidSimulation/include/SPH/ParticleBufferManagers/).
// Needs -I<this repo>/include for sph_hip.h and links libsph_hip.so.
#include <vector>

#include "SPH/Core/ParticleBufferManager.h"
#include "sph_hip.h"

namespace SPH
{
	// ParticleBufferManager (SPHFluidSimulation/include/SPH/Core/ParticleBufferManager.h:8-25)
	// of SimulationEngineHIP. The particle state lives on the GPU; this manager holds
	// its host side: GetBufferCount() buffers of records in the reference's layout
	// (DynamicParticle or StaticParticle, Particles.h:9-20), in the order the engine
	// was given the particles.
	//   Allocate   as OfflineCPUParticleBufferManager.cpp:27-48: the buffers, with the
	//              given records copied into the first.
	//   Advance    the next buffer becomes current; for the dynamic particles of a bound
	//              engine it also enqueues, behind the steps already enqueued, a copy of
	//              the engine's state into it through the engine's ring of pinned buffers
	//              (RenderableGPUParticleBufferManager.cpp:190-270), and returns at once.
	//   LockRead   makes the current buffer hold the engine's latest state (advancing
	//              first if the engine stepped since the last copy), waits for that copy's
	//              event (as OpenCLResourceLock.cpp:73-163 waits for its cl_events) and
	//              returns the records. A non-null signalEvent receives null: the records
	//              are ready when the lock returns, as with an OpenCL lock that has no
	//              write pending.
	//   LockWrite  the same; unlocking uploads the records' positions and velocities to
	//              the engine (sph_write), so its next Update starts from them. The static
	//              particles are sorted once at Initialize: a write to them is not uploaded
	//              and is reported as an error.
	//   FlushAllOperations  waits for the engine's stream.
	// Locking a buffer that is already locked, or unlocking one that is not, logs a
	// fatal error and is ignored, as the reference's CPU lock does (CPUResourceLock.cpp:37-46).
	class HIPParticleBufferManager : public ParticleBufferManager
	{
	public:
		HIPParticleBufferManager();
		~HIPParticleBufferManager() override;

		void Clear() override;
		void Advance() override;

		void Allocate(uintMem newParticleSize, uintMem newParticleCount, void* particles, uintMem newBufferCount) override;

		uintMem GetBufferCount() const override;
		uintMem GetParticleCount() const override;
		uintMem GetParticleSize() const override;

		ResourceLockGuard LockRead(void* signalEvent) override;
		ResourceLockGuard LockWrite(void* signalEvent) override;

		void FlushAllOperations() override;

		// For SimulationEngineHIP: the engine whose dynamic (or static) particles the
		// buffers hold, or nullptr; and that it has enqueued steps since.
		void BindEngine(sph_handle engine, bool dynamicParticles);
		void MarkStale();

	private:
		struct BufferLock
		{
			HIPParticleBufferManager* owner;
			uintMem buffer;
			int state;        // 0 unlocked, 1 read, 2 write
			int pendingSlot;  // engine readback slot still to be copied in, or -1
		};

		ResourceLockGuard Lock(void* signalEvent, int state);
		void Refresh();
		void Receive(BufferLock& lock);
		static void UnlockRead(ArrayView<void*> waitEvents, void* userData);
		static void UnlockWrite(ArrayView<void*> waitEvents, void* userData);
		char* BufferPtr(uintMem buffer) { return storage.data() + buffer * particleSize * particleCount; }

		std::vector<char> storage;
		std::vector<BufferLock> locks;
		uintMem currentBuffer;
		uintMem particleSize;
		uintMem particleCount;
		sph_handle engine;
		bool dynamic;
		bool stale;
	};
}

// Drop-in file for the reference tree (SPHFluidSimulation/source/SPH/ParticleBufferManagers/).
#include "pch.h"
#include "SPH/ParticleBufferManagers/HIPParticleBufferManager.h"

#include <cstring>

namespace SPH
{
	// A failed engine call: what was attempted, then the engine's own message.
	static void LogEngineError(const char* what, sph_handle engine)
	{
		Debug::Logger::LogError("SPH Library", what);
		Debug::Logger::LogError("SPH Library", sph_last_error(engine));
	}

	HIPParticleBufferManager::HIPParticleBufferManager()
		: currentBuffer(0), particleSize(0), particleCount(0), engine(nullptr), dynamic(true), stale(false)
	{
	}
	HIPParticleBufferManager::~HIPParticleBufferManager()
	{
		Clear();
	}
	void HIPParticleBufferManager::Clear()
	{
		FlushAllOperations();
		storage.clear();
		locks.clear();
		currentBuffer = 0;
		particleSize = 0;
		particleCount = 0;
		engine = nullptr;
		stale = false;
	}
	void HIPParticleBufferManager::Allocate(uintMem newParticleSize, uintMem newParticleCount, void* particles, uintMem newBufferCount)
	{
		Clear();
		if (newBufferCount == 0)
		{
			Debug::Logger::LogFatal("SPH Library", "HIPParticleBufferManager::Allocate with no buffers");
			return;
		}
		particleSize = newParticleSize;
		particleCount = newParticleCount;
		storage.assign(particleSize * particleCount * newBufferCount, 0);
		locks.resize(newBufferCount);
		for (uintMem i = 0; i < newBufferCount; ++i)
			locks[i] = BufferLock{ this, i, 0, -1 };
		if (particles != nullptr && particleSize * particleCount != 0)
			memcpy(storage.data(), particles, particleSize * particleCount);
	}
	void HIPParticleBufferManager::BindEngine(sph_handle newEngine, bool dynamicParticles)
	{
		engine = newEngine;
		dynamic = dynamicParticles;
		stale = false;
	}
	void HIPParticleBufferManager::MarkStale()
	{
		stale = true;
	}
	uintMem HIPParticleBufferManager::GetBufferCount() const
	{
		return locks.size();
	}
	uintMem HIPParticleBufferManager::GetParticleCount() const
	{
		return particleCount;
	}
	uintMem HIPParticleBufferManager::GetParticleSize() const
	{
		return particleSize;
	}
	void HIPParticleBufferManager::Advance()
	{
		if (locks.empty())
			return;
		currentBuffer = (currentBuffer + 1) % locks.size();
		if (engine == nullptr || !dynamic || particleCount == 0)
			return;
		uint32_t slot = 0;
		if (sph_readback_enqueue(engine, &slot) != SPH_OK)
		{
			LogEngineError("HIPParticleBufferManager: readback failed", engine);
			return;
		}
		// The engine reuses its slot three enqueues later: a buffer still waiting for
		// that slot would receive a newer state, so it keeps what it holds.
		for (auto& l : locks)
			if (l.pendingSlot == (int)slot)
				l.pendingSlot = -1;
		locks[currentBuffer].pendingSlot = (int)slot;
		stale = false;
	}
	// Converts the engine's readback (positions, velocities, pressures, keys in
	// original order) into the locked buffer's records.
	void HIPParticleBufferManager::Receive(BufferLock& lock)
	{
		if (lock.pendingSlot < 0)
			return;
		if (lock.state != 0)
		{
			// Still held since the ring last came round to it: its records stay as they are.
			Debug::Logger::LogFatal("SPH Library", "HIPParticleBufferManager: a copy landed in a buffer that is still locked");
			lock.pendingSlot = -1;
			return;
		}
		const float *pos = nullptr, *vel = nullptr, *pressure = nullptr;
		const uint32_t* keys = nullptr;
		uint64_t n = 0;
		double time = 0.0;
		const int rc = sph_readback_lock(engine, (uint32_t)lock.pendingSlot, &pos, &vel, &pressure, &keys, &n, &time);
		lock.pendingSlot = -1;
		if (rc != SPH_OK || n != particleCount || particleSize != sizeof(DynamicParticle))
		{
			LogEngineError("HIPParticleBufferManager: readback failed", engine);
			return;
		}
		DynamicParticle* out = (DynamicParticle*)BufferPtr(lock.buffer);
		for (uint64_t i = 0; i < n; ++i)
		{
			out[i].position = Vec3f(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
			out[i].pressure = pressure[i];
			out[i].velocity = Vec3f(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
			out[i].hash = keys[i];
		}
	}
	void HIPParticleBufferManager::Refresh()
	{
		if (engine != nullptr && dynamic && stale)
			Advance();
		Receive(locks[currentBuffer]);
	}
	ResourceLockGuard HIPParticleBufferManager::Lock(void* signalEvent, int state)
	{
		if (signalEvent != nullptr)
			*(void**)signalEvent = nullptr;
		if (locks.empty())
			return ResourceLockGuard();
		// Bring the current buffer up to date first: that may advance the ring, and
		// the lock is on the buffer current after it.
		Refresh();
		BufferLock& lock = locks[currentBuffer];
		if (lock.state != 0)
		{
			Debug::Logger::LogFatal("SPH Library", "HIPParticleBufferManager: the buffer is already locked");
			return ResourceLockGuard();
		}
		lock.state = state;
		return ResourceLockGuard(state == 1 ? UnlockRead : UnlockWrite, BufferPtr(lock.buffer), &lock);
	}
	ResourceLockGuard HIPParticleBufferManager::LockRead(void* signalEvent)
	{
		return Lock(signalEvent, 1);
	}
	ResourceLockGuard HIPParticleBufferManager::LockWrite(void* signalEvent)
	{
		return Lock(signalEvent, 2);
	}
	void HIPParticleBufferManager::UnlockRead(ArrayView<void*> waitEvents, void* userData)
	{
		BufferLock& lock = *(BufferLock*)userData;
		if (lock.state != 1)
		{
			Debug::Logger::LogFatal("SPH Library", "HIPParticleBufferManager: read unlock of a buffer not locked for reading");
			return;
		}
		lock.state = 0;
	}
	void HIPParticleBufferManager::UnlockWrite(ArrayView<void*> waitEvents, void* userData)
	{
		BufferLock& lock = *(BufferLock*)userData;
		HIPParticleBufferManager& m = *lock.owner;
		if (lock.state != 2)
		{
			Debug::Logger::LogFatal("SPH Library", "HIPParticleBufferManager: write unlock of a buffer not locked for writing");
			return;
		}
		lock.state = 0;
		if (m.engine == nullptr || m.particleCount == 0)
			return;
		if (!m.dynamic)
		{
			Debug::Logger::LogError("SPH Library", "SimulationEngineHIP sorts its static particles once at Initialize; the write is not uploaded");
			return;
		}
		std::vector<float> pos(3 * m.particleCount), vel(3 * m.particleCount);
		const DynamicParticle* in = (const DynamicParticle*)m.BufferPtr(lock.buffer);
		for (uintMem i = 0; i < m.particleCount; ++i)
		{
			pos[3 * i] = in[i].position.x;
			pos[3 * i + 1] = in[i].position.y;
			pos[3 * i + 2] = in[i].position.z;
			vel[3 * i] = in[i].velocity.x;
			vel[3 * i + 1] = in[i].velocity.y;
			vel[3 * i + 2] = in[i].velocity.z;
		}
		if (sph_write(m.engine, pos.data(), vel.data()) != SPH_OK)
			LogEngineError("HIPParticleBufferManager: write failed", m.engine);
	}
	void HIPParticleBufferManager::FlushAllOperations()
	{
		if (engine != nullptr && sph_synchronize(engine) != SPH_OK)
			LogEngineError("HIPParticleBufferManager: synchronize failed", engine);
	}
}

// Definitions behind the mirror declarations in tests/cpp/refmirror/ (the
// reference defines these in SPHFluidSimulation/source/SPH/Core/ResourceLockGuard.cpp
// and ParticleSimulationEngine.cpp:6-36, and BlazeEngine's logger), written for the
// test build with the documented behaviour: a guard unlocks once and reports a
// destruction while still locked; ParseParameter converts a named scene parameter.
#include <cstdio>
#include <cstdlib>

#include "pch.h"
#include "SPH/Core/ResourceLockGuard.h"
#include "SPH/Core/SimulationEngine.h"

namespace Blaze {
int Debug::Logger::errors = 0;
static void log_line(const char* level, StringView source, StringView message) {
    std::fprintf(stderr, "[%s] %s: %s\n", level, source.str().c_str(), message.str().c_str());
}
void Debug::Logger::LogInfo(StringView s, StringView m) { log_line("info", s, m); }
void Debug::Logger::LogWarning(StringView s, StringView m) { log_line("warning", s, m); }
void Debug::Logger::LogError(StringView s, StringView m) {
    ++errors;
    log_line("error", s, m);
}
void Debug::Logger::LogFatal(StringView s, StringView m) {
    ++errors;
    log_line("fatal", s, m);
}
Array<Graphics::BasicIndexedMesh::Triangle> Graphics::BasicIndexedMesh::CreateTriangleArray() const {
    Array<Triangle> out(triangles.size());
    for (uintMem i = 0; i < triangles.size(); ++i) out[i] = triangles[i];
    return out;
}
}  // namespace Blaze

namespace SPH {
ResourceLockGuard::ResourceLockGuard() : unlockFunction(nullptr), resource(nullptr), userData(nullptr) {}
ResourceLockGuard::ResourceLockGuard(UnlockFunction f, void* r, void* u) : unlockFunction(f), resource(r), userData(u) {}
ResourceLockGuard::ResourceLockGuard(ResourceLockGuard&& o) noexcept
    : unlockFunction(o.unlockFunction), resource(o.resource), userData(o.userData) {
    o.unlockFunction = nullptr;
    o.resource = nullptr;
    o.userData = nullptr;
}
ResourceLockGuard::~ResourceLockGuard() {
    if (unlockFunction) Debug::Logger::LogError("SPH Library", "lock guard destroyed while locked");
}
void ResourceLockGuard::Unlock(ArrayView<void*> waitEvents) {
    if (unlockFunction) unlockFunction(waitEvents, userData);
    unlockFunction = nullptr;
    resource = nullptr;
}
void* ResourceLockGuard::GetResource() { return resource; }
const ResourceLockGuard& ResourceLockGuard::operator=(ResourceLockGuard&& o) noexcept {
    unlockFunction = o.unlockFunction;
    resource = o.resource;
    userData = o.userData;
    o.unlockFunction = nullptr;
    o.resource = nullptr;
    o.userData = nullptr;
    return *this;
}

bool ParticleSimulationParameters::ParseParameter(StringView name, float& value) const {
    const String* v = otherParameters.Find(String(name.str()));
    if (!v) return false;
    char* end = nullptr;
    const float f = std::strtof(v->str().c_str(), &end);
    if (end == v->str().c_str()) {
        Debug::Logger::LogWarning("Client", "scene parameter is not a number");
        return false;
    }
    value = f;
    return true;
}
bool ParticleSimulationParameters::ParseParameter(StringView name, bool& value) const {
    const String* v = otherParameters.Find(String(name.str()));
    if (!v) return false;
    if (v->str() == "true" || v->str() == "1") {
        value = true;
        return true;
    }
    if (v->str() == "false" || v->str() == "0") {
        value = false;
        return true;
    }
    return false;
}
}  // namespace SPH

// Drives sph_hip::Engine (include/sph/simulation_engine_hip.hpp) the way the
// reference's application drives its engines (SimulationVisualisationScene.cpp:
// Initialize once, Update(dt, 1) per frame, read the particles), on a small
// dam break given on stdin as: n, then n lines "x y z", then n_s, then n_s lines,
// then 9 parameter floats. Prints the final state in original order.
#include <cstdio>
#include <exception>

#include "sph/simulation_engine_hip.hpp"

int main() {
    try {
        sph_hip::SceneArrays scene;
        unsigned long long n = 0, ns = 0;
        if (scanf("%llu", &n) != 1) return 2;
        scene.dynamicPositions.resize(n);
        for (auto& p : scene.dynamicPositions)
            if (scanf("%f %f %f", &p.x, &p.y, &p.z) != 3) return 2;
        if (scanf("%llu", &ns) != 1) return 2;
        scene.staticPositions.resize(ns);
        for (auto& p : scene.staticPositions)
            if (scanf("%f %f %f", &p.x, &p.y, &p.z) != 3) return 2;
        float* f = &scene.parameters.particleMass;
        for (int k = 0; k < 9; ++k)
            if (scanf("%f", f + k) != 1) return 2;
        int frames = 0;
        float dt = 0;
        if (scanf("%d %f", &frames, &dt) != 2) return 2;

        sph_hip::Engine e(0);
        e.Initialize(scene);
        // Frames as the reference's app runs them, with the asynchronous
        // readback a renderer would use: enqueue a copy each frame, consume it
        // when it has landed without waiting for the GPU.
        sph_hip::AsyncReadback* async = e.GetReadback();
        unsigned landed = 0;
        for (int i = 0; i < frames; ++i) {
            e.Update(dt, 1);
            async->Advance();
            const std::vector<sph_hip::Record>* view = nullptr;
            if (async->TryLockRead(view) && view->size() == n) ++landed;
        }
        sph_hip::AsyncReadback* m = e.GetReadback();
        m->FlushAllOperations();
        const auto& recs = m->LockRead();
        fprintf(stderr, "frames with a landed copy: %u of %d\n", landed, frames);
        printf("%s %llu %.9g\n", e.SystemImplementationName(), (unsigned long long)m->GetParticleCount(),
               (double)e.GetSimulationTime());
        for (const auto& r : recs)
            printf("%.9g %.9g %.9g %.9g %.9g %.9g %.9g %u\n", r.position.x, r.position.y, r.position.z, r.pressure,
                   r.velocity.x, r.velocity.y, r.velocity.z, r.hash);
        return 0;
    } catch (const std::exception& ex) {
        fprintf(stderr, "%s\n", ex.what());
        return 1;
    }
}

/*
 * sph_hip.h — C ABI of the MI355X (gfx950) SPH step engine.
 *
 * This is the drop-in boundary for the step hot path of Murk42/SPHFluidSimulation:
 * the reference's engines sit behind the C++ interface SPH::SimulationEngine
 * (SPHFluidSimulation/include/SPH/Core/SimulationEngine.h:35-48) and exchange
 * particle data through SPH::ParticleBufferManager
 * (SPHFluidSimulation/include/SPH/Core/ParticleBufferManager.h:8-25). Every
 * entry point below names the reference interface it replaces. The C++ adapter
 * that implements SPH::SimulationEngine on top of these calls is
 * include/sph/simulation_engine_hip.hpp.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative SPH_E_* code on error;
 *     nothing throws across the ABI. sph_last_error() returns the message of the
 *     last failure on a handle.
 *   - One engine per handle; a handle is not thread-safe (the reference is driven
 *     from one thread, the application's main loop).
 *   - Device memory belongs to the handle. Pointers passed in are host pointers,
 *     read during the call only.
 *   - Positions and velocities are packed xyz float triples. "Original order" is
 *     the order of the arrays given to sph_initialize / sph_batch_initialize.
 */
#ifndef SPH_HIP_H
#define SPH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define SPH_OK 0
#define SPH_E_INVALID (-1)       /* bad handle or argument */
#define SPH_E_HIP (-2)           /* a HIP runtime call failed */
#define SPH_E_STATE (-3)         /* call not valid in the current state */
#define SPH_E_VALIDATION (-4)    /* sph_validate found a broken invariant */
#define SPH_E_NO_DEVICE (-5)     /* no HIP device */

/* sph_initialize / sph_batch_initialize flags. */
#define SPH_FLAG_HIP_GRAPH 0x1u     /* replay each step from a captured hipGraph */
#define SPH_FLAG_RECORD_DEBUG 0x2u  /* keep per-particle density and acceleration of the last step */
#define SPH_FLAG_FP16_OFFSETS 0x4u  /* mixed precision (default walk only): the density pass's candidate
                                       filter tests FP16 cell offsets (a proven superset of the h-ball),
                                       every neighbour sum is Kahan-compensated FP32; kernel values
                                       stay FP32, conserved reductions FP64 */
#define SPH_FLAG_FP16_STORED 0x200u /* mixed precision as SPH_FLAG_FP16_OFFSETS (which it implies), the
                                       filter reading per-particle FP16 offsets from the particle's own
                                       cell corner, 8 bytes written by each step's scatter, instead of
                                       forming them from the 16-byte positions */
#define SPH_FLAG_CELL_WALK 0x8u     /* neighbour passes read every candidate from HBM, one cell at a
                                       time: the reference's loop shape, kept for comparison */
#define SPH_FLAG_RUN_WALK 0x10u     /* neighbour passes read candidates from HBM in merged Morton runs */
#define SPH_FLAG_TILE_WALK 0x20u    /* neighbour passes stage each workgroup's buckets in LDS.
                                       Default (no walk flag): cell-group waves, where a wave tests
                                       the candidates of a cell against all of the cell's particles;
                                       the density pass hands the force pass each particle's neighbour
                                       list (up to SPH_CW_LIST_MAX particles in the engine) or its
                                       neighbour mask (more particles, or either flag below) */
#define SPH_FLAG_CW_LISTS 0x40u     /* default walk: neighbour lists whatever the particle count */
#define SPH_FLAG_CW_MASKS 0x80u     /* default walk: neighbour masks whatever the particle count */
#define SPH_CW_LIST_MAX 1572864u    /* particles up to which the default walk uses lists */
#define SPH_FLAG_PARTICLE_MAP 0x100u /* the reference's particle-map mode (SimulationEngineCPU.cpp:280-316):
                                       the state keeps its memory order between reorders and the buckets
                                       list it through a particle map refilled every step (SPHFunctions.cpp:
                                       346-356); a step reorders the state when the reorder interval has
                                       elapsed (sph_set_reorder_interval). Walks cell by cell, as the
                                       reference's loop does. Default (no flag): the state is reordered
                                       every step, the fastest choice on the GPU */

/* sph_validate result bits (the invariants of SimulationEngine.h:115-282). */
#define SPH_V_NONFINITE 0x1u        /* NaN or Inf in position, velocity or pressure */
#define SPH_V_KEY 0x2u              /* stored key != key recomputed from position */
#define SPH_V_TABLE_FIRST 0x4u      /* table[0] != 0 */
#define SPH_V_TABLE_LAST 0x8u       /* table[M] != N */
#define SPH_V_TABLE_ORDER 0x10u     /* table not monotone or above N */
#define SPH_V_BUCKET 0x20u          /* a particle lies outside its bucket's range */
#define SPH_V_STATIC 0x40u          /* one of the above on the static (wall) table */

/* Stages reported by sph_stage_times, in launch order. */
#define SPH_STAGE_DENSITY 0   /* UpdateParticlePressure, SPHFunctions.cpp:358-501 */
#define SPH_STAGE_FORCE 1     /* UpdateParticleDynamics + hash + histogram, SPHFunctions.cpp:503-656, 321-331 */
#define SPH_STAGE_SCAN 2      /* prefix sum, SimulationEngineCPU.cpp:245-253 */
#define SPH_STAGE_SCATTER 3   /* bucket fill / reorder, SPHFunctions.cpp:332-356 */
#define SPH_STAGE_COUNT 4

/* ParticleBehaviourParameters (Particles.h:23-41 / SPHFunctions.cpp:23-41),
 * same 11-float layout. selfDensity and smoothingKernelConstant are outputs:
 * the engine derives them as SimulationEngineCPU.cpp:406-407 does and ignores
 * the values passed in. elasticity is carried but unused, as in the reference. */
typedef struct sph_params {
    float particleMass;
    float gasConstant;
    float elasticity;
    float viscosity;
    float gravityX;
    float gravityY;
    float gravityZ;
    float restDensity;
    float maxInteractionDistance;
    float selfDensity;
    float smoothingKernelConstant;
} sph_params;

typedef struct sph_engine* sph_handle;

/* Constructor of SimulationEngineGPU (SimulationEngineGPU.cpp:32) /
 * SimulationEngineCPU (SimulationEngineCPU.cpp:351): binds the handle to a HIP
 * device and creates its stream. */
int sph_create(sph_handle* out, int device);

/* Destructor (SimulationEngineCPU.cpp:361-364): Clear() and free the handle. */
int sph_destroy(sph_handle h);

/* SimulationEngine::SystemImplementationName (SimulationEngine.h:44):
 * "CPU" / "GPU" in the reference, "HIP" here. */
const char* sph_system_implementation_name(void);

/* SimulationEngine::Initialize (SimulationEngine.h:41; SimulationEngineCPU.cpp:396-504):
 * copies one scene's dynamic particles (n, with velocities; vel may be NULL for
 * zero) and static wall particles (n_s, may be 0) to the device and sorts both
 * into buckets, M = 2n dynamic buckets and M_s = n_s static buckets
 * (SimulationEngineCPU.cpp:468, 486). Within a bucket particles are ordered by
 * descending original index, the order a one-thread reference run produces
 * (SimulationEngineCPU.cpp:55-131, SimulationEngine.h:79-100). */
int sph_initialize(sph_handle h, const sph_params* params, const float* dyn_pos, const float* dyn_vel, uint64_t n,
                   const float* static_pos, uint64_t n_s, uint32_t flags);

/* Batched Initialize (no reference counterpart; the reference runs one scene per
 * engine): n_scenes independent scenes advance together, one launch per stage.
 * dyn_offsets and static_offsets hold n_scenes + 1 prefix offsets into the
 * concatenated dyn_pos/dyn_vel and static_pos arrays; params holds one
 * sph_params per scene. Each scene keeps its own table of 2*N_s + 1 entries.
 * Both initialize calls accept up to 2^28 - 1 dynamic and 2^28 - 1 static
 * particles in total (device offsets of 16-byte records are 32-bit). */
int sph_batch_initialize(sph_handle h, uint32_t n_scenes, const sph_params* params, const uint64_t* dyn_offsets,
                         const float* dyn_pos, const float* dyn_vel, const uint64_t* static_offsets,
                         const float* static_pos, uint32_t flags);

/* Triangle boundary mesh of one scene (SceneBlueprint::GetMesh, SceneBlueprint.h:20;
 * SimulationEngineCPU.cpp:418): count triangles of 9 floats (p1, p2, p3). */
int sph_set_triangles(sph_handle h, uint32_t scene, const float* tri9, uint64_t count);

/* The scene's reorderTimeInterval (SimulationEngineCPU.cpp:355, 400; read from the
 * scene's otherParameters): in SPH_FLAG_PARTICLE_MAP mode, every step of an update
 * whose start finds the time since the last reorder above the interval reorders
 * the state; the time advances by dt once per update and restarts at 0 in the
 * update that takes it above the interval (SimulationEngineCPU.cpp:446, 455-457),
 * so the next update's check finds 0: as in the reference, a positive interval
 * never reorders and a negative one reorders every update. Initialize resets it
 * to FLT_MAX (SimulationEngineCPU.cpp:389-390). No effect without the flag. */
int sph_set_reorder_interval(sph_handle h, float interval);

/* SimulationEngine::Update (SimulationEngine.h:42; SimulationEngineCPU.cpp:420-459):
 * enqueues `steps` steps of size dt on the engine's stream and returns. Unlike
 * the reference's CPU engine it never drops an update. Update before Initialize
 * returns SPH_E_STATE; n == 0 returns SPH_OK and does nothing
 * (SimulationEngineCPU.cpp:422-429). */
int sph_update(sph_handle h, float dt, uint32_t steps);

/* ParticleBufferManager::LockWrite of the current buffer (ParticleBufferManager.h:22,
 * OfflineCPUParticleBufferManager.cpp:68-74) followed by its unlock: replaces the
 * positions and/or velocities (3 floats per particle, original order; either may
 * be NULL to keep it) of the state the next Update starts from. As in the
 * reference, whose engines walk a written buffer through the tables of their last
 * step (SimulationEngineCPU.cpp:132-349), the stored keys, the bucket tables and
 * the pressures stay as they are until that Update rehashes; sph_validate reports
 * SPH_V_KEY for a written particle that left its cell. Synchronous: waits for
 * the steps already enqueued. */
int sph_write(sph_handle h, const float* pos, const float* vel);

/* No counterpart in the reference: sph_read for callers that cannot allocate
 * device memory (serving loops and harnesses that forbid hipMalloc after set-up,
 * code next to HIP graph capture). Same results as sph_read; its staging is a
 * fixed device array of SPH_NOALLOC_MAX_PARTICLES particles (28 bytes each,
 * 28 MiB) that each GPU holds from sph_create until the process ends, so the call
 * allocates nothing. Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (two threads, two handles) race. An engine with more
 * particles returns SPH_E_INVALID. */
#define SPH_NOALLOC_MAX_PARTICLES (1u << 20)
int sph_read_noalloc(sph_handle h, float* pos, float* vel, float* pressure);

/* Asynchronous readback, the HIP form of the reference's renderable buffer
 * manager (RenderableGPUParticleBufferManager.cpp:190-270: a ring of buffers, a
 * read lock that waits on an event, OpenCLResourceLock.cpp:73-163). The engine
 * keeps a ring of 3 pinned host buffers.
 *
 * sph_readback_enqueue enqueues, after every step already enqueued, a copy of the
 * state at that point (original particle order) into the next ring slot, and
 * returns the slot at once; later sph_update calls do not change that copy.
 * sph_readback_query tells whether the copy has landed (*ready = 1) without
 * waiting. sph_readback_lock waits for it and returns pointers into the pinned
 * slot: pos and vel hold 3 floats per particle, pressure and keys one each,
 * sim_time the simulation time of the copy. The pointers stay valid until the
 * slot is enqueued again (three enqueues later) or the engine is cleared.
 * Querying or locking a slot never enqueued returns SPH_E_STATE. */
int sph_readback_enqueue(sph_handle h, uint32_t* slot);
int sph_readback_query(sph_handle h, uint32_t slot, int* ready);
int sph_readback_lock(sph_handle h, uint32_t slot, const float** pos, const float** vel, const float** pressure,
                      const uint32_t** keys, uint64_t* n, double* sim_time);

/* Waits for all enqueued work (ParticleBufferManager::FlushAllOperations,
 * ParticleBufferManager.h:24). */
int sph_synchronize(sph_handle h);

/* ParticleBufferManager::LockRead of the current buffer
 * (OfflineCPUParticleBufferManager.cpp:61-67, RenderableGPUParticleBufferManager.cpp:256-270),
 * returned in original order rather than the reference's memory order.
 * pos and vel take 3 floats per particle, pressure 1; any may be NULL. */
int sph_read(sph_handle h, float* pos, float* vel, float* pressure);

/* Stored bucket key per particle (DynamicParticle::hash, Particles.h:16 /
 * velocityAndHash.w, SPHFunctions.cpp:328, 645), in original order; scene-local. */
int sph_read_keys(sph_handle h, uint32_t* keys);

/* The bucket-start table (the reference's hashMap, SimulationEngineCPU.cpp:486-487):
 * per scene M + 1 scene-local entries, concatenated over scenes. */
int sph_read_cell_starts(sph_handle h, uint32_t* starts);

/* Particles in bucket order (what the reference's particleMap indexes,
 * SPHFunctions.cpp:346-356), as original (global) indices. */
int sph_read_bucket_ids(sph_handle h, uint32_t* ids);

/* Static table (per scene M_s + 1 scene-local entries) and the original
 * (global) wall index of each sorted wall particle; either may be NULL. */
int sph_read_static(sph_handle h, uint32_t* starts, uint32_t* ids);

/* Per-particle density and acceleration (3 floats) of the last step, original
 * order; needs SPH_FLAG_RECORD_DEBUG. */
int sph_read_debug(sph_handle h, float* density, float* accel);

/* FP64 conserved quantities per scene: out[8*s + {0..7}] =
 * {mass, momentum x, y, z, kinetic energy, centre of mass x, y, z}. */
int sph_conserved(sph_handle h, double* out);

/* Runs the reference's debug validators (SimulationEngine.h:115-282) on the
 * device. Returns SPH_OK, or SPH_E_VALIDATION with the SPH_V_* bits in *bits. */
int sph_validate(sph_handle h, uint32_t* bits);

/* SimulationEngine::GetSimulationTime (SimulationEngine.h:48). */
float sph_simulation_time(sph_handle h);

/* SimulationEngine::Clear (SimulationEngine.h:40; SimulationEngineCPU.cpp:365-395):
 * waits for outstanding work and releases all device state. */
int sph_clear(sph_handle h);

/* Message of the last failure on this handle ("" if none). */
const char* sph_last_error(sph_handle h);

/* ParticleBufferManager::GetParticleCount (ParticleBufferManager.h:20) of the
 * dynamic and static managers, summed over scenes. */
uint64_t sph_particle_count(sph_handle h);
uint64_t sph_static_particle_count(sph_handle h);
uint32_t sph_scene_count(sph_handle h);
/* Number of dynamic buckets M of one scene (the table has M + 1 entries). */
uint64_t sph_hash_map_size(sph_handle h, uint32_t scene);

/* Per-stage timing with hipEvents on the engine's stream (the reference's
 * PerformanceProfile, SimulationEngineGPU.cpp:204-273). stage_mask has bit
 * (1 << SPH_STAGE_*) set for every stage whose launches are bracketed by events
 * (0 turns profiling off, SPH_PROFILE_ALL times every stage); sph_stage_times
 * returns the summed milliseconds and launch count per stage since the last
 * sph_set_profiling call. Steps are launched directly, not from hipGraphs,
 * while profiling is on. With SPH_PROFILE_ROTATE also set, each step brackets
 * only one of the masked stages, taking them in turn: every event record costs
 * the stream a few microseconds, so this halves the cost of timing two stages
 * while each is still timed over the same run (on every other step). */
#define SPH_PROFILE_ALL 0xFu
#define SPH_PROFILE_ROTATE 0x10u
int sph_set_profiling(sph_handle h, uint32_t stage_mask);
int sph_stage_times(sph_handle h, double* ms, uint64_t* launches);

/* Autotuner with no counterpart in the reference (whose engines have one fixed
 * neighbour loop, SPHFunctions.cpp:397-404): times the neighbour-pass variants
 * (cell-group waves with neighbour lists and with neighbour masks, Morton runs;
 * the LDS bucket tiles only when they are the configured walk, being many times
 * slower elsewhere) for 2 warm-up steps plus trial_steps (rounded up to an even
 * count, at least 2) steps of dt on the current state, keeps the fastest for
 * later updates (the configured variant unless another is more than 3% faster)
 * and, for the cell-group waves, their launch shape the same way (waves per
 * workgroup of each pass; sph_launch_shape reads it), and writes its flags to
 * *chosen: the walk flag (0, SPH_FLAG_RUN_WALK or SPH_FLAG_TILE_WALK), with
 * SPH_FLAG_CW_LISTS or SPH_FLAG_CW_MASKS added for the cell-group waves. It allocates whichever of
 * the lists and masks the engine lacks and keeps both. The state, the
 * simulation time and the results of later updates (within the float tolerance
 * of the variants' summation orders) are as if it had not been called: each
 * trial runs from a device copy of the state. Synchronous; allocates scratch
 * copies of positions, velocities, ids and the cell-start table. An engine
 * initialized with SPH_FLAG_CELL_WALK keeps that walk. */
int sph_autotune(sph_handle h, float dt, uint32_t trial_steps, uint32_t* chosen);

/* Launch shape of the default walk's neighbour passes, no counterpart in the
 * reference: waves per workgroup of the density pass and of the force pass (1, 2,
 * 4 or 8 waves of 64 lanes, one particle per lane). Every wave keeps its
 * 64-particle slice, so results do not depend on the shape. Until it is set,
 * sph_initialize picks 4, 4 over neighbour lists and 2, 8 over masks; a shape
 * set here is kept by later sph_initialize calls. sph_autotune chooses it with
 * the walk. Cell size is not a tuning parameter:
 * keys and bucket tables must equal the reference's, whose cells are h. */
int sph_set_launch_shape(sph_handle h, uint32_t density_waves, uint32_t force_waves);
int sph_launch_shape(sph_handle h, uint32_t* density_waves, uint32_t* force_waves);

/* Measurement aid with no counterpart in the reference: launches `repeats` plain
 * copies of `bytes` bytes between two device buffers (kernel k_copy_probe), so
 * that HBM byte counters read under a profiler can be scaled against a kernel
 * whose traffic is known exactly. Synchronous. */
int sph_copy_probe(sph_handle h, uint64_t bytes, uint32_t repeats);

#ifdef __cplusplus
}
#endif

#endif /* SPH_HIP_H */

"""SimulationEngineHIP: the Python host over the C ABI (include/sph_hip.h).

It mirrors SPH::SimulationEngine (SPHFluidSimulation/include/SPH/Core/SimulationEngine.h:35-48):
Clear, Initialize, Update, SystemImplementationName and GetSimulationTime keep the
reference's names and meaning; the read* methods stand in for
ParticleBufferManager::LockRead on the current buffer and return numpy arrays in
original particle order. Errors raise SPHError (the reference logs and returns;
the C ABI returns the code and this layer raises it).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .scenes import Scene, concat


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _params(p) -> _lib.sph_params:
    out = _lib.sph_params()
    for name, _ in _lib.sph_params._fields_:
        setattr(out, name, float(getattr(p, name)))
    return out


class SimulationEngineHIP:
    """One engine per GPU stream; not thread-safe (like the reference's engines)."""

    def __init__(self, device: int = 0):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        rc = self._lib.sph_create(ctypes.byref(h), int(device))
        if rc != _lib.SPH_OK:
            raise _lib.SPHError(rc, "sph_create failed (no HIP device?)")
        self._h = h
        self._scene_sizes = []  # (n, n_static, M) per scene
        self.neighbours = None  # the cell-wave hand-off the last autotune chose ("lists" or "masks")

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.sph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        return _lib.check(self._h, rc)

    # -- SimulationEngine interface ----------------------------------------
    @staticmethod
    def SystemImplementationName() -> str:
        return _lib.load().sph_system_implementation_name().decode()

    def Clear(self):
        self._check(self._lib.sph_clear(self._h))
        self._scene_sizes = []

    def Initialize(self, scene, flags: int = 0):
        """Initialize from one Scene or a list of Scenes (batched)."""
        scenes = [scene] if isinstance(scene, Scene) else list(scene)
        params, dyn_off, pos, vel, st_off, spos = concat(scenes)
        arr = (_lib.sph_params * len(params))(*[_params(p) for p in params])
        pos, vel, spos = _f32(pos), _f32(vel), _f32(spos)
        self._check(self._lib.sph_batch_initialize(
            self._h, len(scenes), arr, _ptr(dyn_off, ctypes.c_uint64), _ptr(pos, ctypes.c_float),
            _ptr(vel, ctypes.c_float), _ptr(st_off, ctypes.c_uint64),
            _ptr(spos, ctypes.c_float) if spos.size else None, int(flags)))
        self._scene_sizes = [(sc.n, sc.n_static, int(self._lib.sph_hash_map_size(self._h, s)))
                             for s, sc in enumerate(scenes)]
        self._dyn_off = dyn_off.astype(np.int64)
        self._st_off = st_off.astype(np.int64)
        # A scene file's reorderTimeInterval (SimulationEngineCPU.cpp:400), used in particle-map mode.
        if len(scenes) == 1 and "reorderTimeInterval" in scenes[0].other:
            try:
                self.set_reorder_interval(float(scenes[0].other["reorderTimeInterval"]))
            except ValueError:
                pass  # not a number: the reference warns and keeps its interval
        # Boundary meshes (SimulationEngineCPU.cpp:418: triangles = scene.GetMesh().CreateTriangleArray()).
        for s, sc in enumerate(scenes):
            if getattr(sc, "triangles", None) is not None and len(sc.triangles):
                self.set_triangles(s, sc.triangles)

    def set_reorder_interval(self, interval: float):
        """reorderTimeInterval of SPH_FLAG_PARTICLE_MAP mode (SimulationEngineCPU.cpp:280-316, 446)."""
        self._check(self._lib.sph_set_reorder_interval(self._h, ctypes.c_float(interval)))

    def Update(self, dt: float, steps: int = 1):
        self._check(self._lib.sph_update(self._h, float(dt), int(steps)))

    def GetSimulationTime(self) -> float:
        return float(self._lib.sph_simulation_time(self._h))

    # -- extras -----------------------------------------------------------
    def set_triangles(self, scene: int, tris):
        t = _f32(tris).reshape(-1, 9)
        self._check(self._lib.sph_set_triangles(self._h, int(scene), _ptr(t, ctypes.c_float), t.shape[0]))

    def synchronize(self):
        self._check(self._lib.sph_synchronize(self._h))

    @property
    def n(self) -> int:
        return int(self._lib.sph_particle_count(self._h))

    @property
    def n_static(self) -> int:
        return int(self._lib.sph_static_particle_count(self._h))

    @property
    def n_scenes(self) -> int:
        return int(self._lib.sph_scene_count(self._h))

    def hash_map_size(self, scene: int = 0) -> int:
        return int(self._lib.sph_hash_map_size(self._h, int(scene)))

    def read(self):
        n = self.n
        pos = np.empty((n, 3), np.float32)
        vel = np.empty((n, 3), np.float32)
        p = np.empty(n, np.float32)
        self._check(self._lib.sph_read(self._h, _ptr(pos, ctypes.c_float), _ptr(vel, ctypes.c_float),
                                       _ptr(p, ctypes.c_float)))
        return pos, vel, p

    def write(self, pos=None, vel=None):
        """LockWrite + unlock of the current buffer: replace positions and/or
        velocities (original order) before the next Update; keys and tables are
        rebuilt by that Update, as in the reference."""
        n = self.n
        p = None if pos is None else _f32(pos).reshape(n, 3)
        v = None if vel is None else _f32(vel).reshape(n, 3)
        self._check(self._lib.sph_write(self._h, _ptr(p, ctypes.c_float), _ptr(v, ctypes.c_float)))

    def read_noalloc(self):
        """read() through sph_read_noalloc: staging in the library's fixed device
        array (at most SPH_NOALLOC_MAX_PARTICLES particles), no device allocation."""
        n = self.n
        pos = np.empty((n, 3), np.float32)
        vel = np.empty((n, 3), np.float32)
        p = np.empty(n, np.float32)
        self._check(self._lib.sph_read_noalloc(self._h, _ptr(pos, ctypes.c_float), _ptr(vel, ctypes.c_float),
                                               _ptr(p, ctypes.c_float)))
        return pos, vel, p

    def read_keys(self):
        k = np.empty(self.n, np.uint32)
        self._check(self._lib.sph_read_keys(self._h, _ptr(k, ctypes.c_uint32)))
        return k

    def read_cell_starts(self):
        """List of per-scene tables (M + 1 scene-local entries each)."""
        total = sum(m + 1 for _, _, m in self._scene_sizes)
        t = np.empty(total, np.uint32)
        self._check(self._lib.sph_read_cell_starts(self._h, _ptr(t, ctypes.c_uint32)))
        out, o = [], 0
        for _, _, m in self._scene_sizes:
            out.append(t[o:o + m + 1])
            o += m + 1
        return out

    def read_bucket_ids(self):
        ids = np.empty(self.n, np.uint32)
        self._check(self._lib.sph_read_bucket_ids(self._h, _ptr(ids, ctypes.c_uint32)))
        return ids

    def read_static(self):
        """(list of per-scene static tables, original wall index per sorted wall)."""
        total = sum((ns + 1) if ns else 0 for _, ns, _ in self._scene_sizes)
        t = np.empty(max(total, 1), np.uint32)
        ids = np.empty(max(self.n_static, 1), np.uint32)
        self._check(self._lib.sph_read_static(self._h, _ptr(t, ctypes.c_uint32), _ptr(ids, ctypes.c_uint32)))
        out, o = [], 0
        for _, ns, _ in self._scene_sizes:
            if ns:
                out.append(t[o:o + ns + 1])
                o += ns + 1
            else:
                out.append(np.zeros(0, np.uint32))
        return out, ids[:self.n_static]

    def read_debug(self):
        n = self.n
        d = np.empty(n, np.float32)
        a = np.empty((n, 3), np.float32)
        self._check(self._lib.sph_read_debug(self._h, _ptr(d, ctypes.c_float), _ptr(a, ctypes.c_float)))
        return d, a

    def conserved(self):
        out = np.empty((self.n_scenes, 8), np.float64)
        self._check(self._lib.sph_conserved(self._h, _ptr(out, ctypes.c_double)))
        return out

    def validate(self) -> int:
        bits = ctypes.c_uint32(0)
        rc = self._lib.sph_validate(self._h, ctypes.byref(bits))
        if rc not in (_lib.SPH_OK, _lib.SPH_E_VALIDATION):
            self._check(rc)
        return int(bits.value)

    def set_profiling(self, stages=None, rotate: bool = False):
        """stages: iterable of stage names to time with hipEvents, True for all, None/False for none.
        rotate: bracket one of those stages per step, in turn (SPH_PROFILE_ROTATE)."""
        if stages is True:
            mask = _lib.SPH_PROFILE_ALL
        elif not stages:
            mask = 0
        else:
            mask = 0
            for s in stages:
                mask |= 1 << _lib.STAGES.index(s)
        if mask and rotate:
            mask |= _lib.SPH_PROFILE_ROTATE
        self._check(self._lib.sph_set_profiling(self._h, mask))

    def stage_times(self):
        ms = np.zeros(4, np.float64)
        n = np.zeros(4, np.uint64)
        self._check(self._lib.sph_stage_times(self._h, _ptr(ms, ctypes.c_double), _ptr(n, ctypes.c_uint64)))
        return {name: (float(ms[i]), int(n[i])) for i, name in enumerate(_lib.STAGES)}

    WALK_NAMES = {0: "cellwave", _lib.SPH_FLAG_RUN_WALK: "runs", _lib.SPH_FLAG_TILE_WALK: "tile",
                  _lib.SPH_FLAG_CELL_WALK: "cells"}

    def autotune(self, dt: float, trial_steps: int = 4) -> str:
        """Time every neighbour-pass variant on the current state (the cell-wave
        walk with neighbour lists and with masks, the Morton runs, the bucket
        tiles), keep the fastest; the state is left as it was. Returns the
        chosen walk's name; for the cell-wave walk, `neighbours` then says
        "lists" or "masks"."""
        chosen = ctypes.c_uint32(0)
        self._check(self._lib.sph_autotune(self._h, ctypes.c_float(dt), int(trial_steps), ctypes.byref(chosen)))
        c = int(chosen.value)
        walk = c & (_lib.SPH_FLAG_CELL_WALK | _lib.SPH_FLAG_RUN_WALK | _lib.SPH_FLAG_TILE_WALK)
        self.neighbours = ("lists" if c & _lib.SPH_FLAG_CW_LISTS else
                           "masks" if c & _lib.SPH_FLAG_CW_MASKS else None)
        return self.WALK_NAMES[walk]

    def set_launch_shape(self, density_waves: int = 4, force_waves: int = 4):
        """Waves per workgroup (1, 2, 4 or 8) of the default walk's density and
        force passes (results unchanged)."""
        self._check(self._lib.sph_set_launch_shape(self._h, int(density_waves), int(force_waves)))

    def launch_shape(self):
        """(density waves, force waves) per workgroup."""
        v = [ctypes.c_uint32(0) for _ in range(2)]
        self._check(self._lib.sph_launch_shape(self._h, *(ctypes.byref(x) for x in v)))
        return tuple(int(x.value) for x in v)

    # -- asynchronous readback (RenderableGPUParticleBufferManager.cpp:190-270) --
    def readback_enqueue(self) -> int:
        """Enqueue a copy of the state after the steps enqueued so far into the
        next of 3 pinned slots; returns the slot without waiting."""
        slot = ctypes.c_uint32(0)
        self._check(self._lib.sph_readback_enqueue(self._h, ctypes.byref(slot)))
        return int(slot.value)

    def readback_ready(self, slot: int) -> bool:
        ready = ctypes.c_int(0)
        self._check(self._lib.sph_readback_query(self._h, int(slot), ctypes.byref(ready)))
        return bool(ready.value)

    def readback_lock(self, slot: int):
        """Wait for a slot; returns copies of (pos, vel, pressure, keys, sim_time)."""
        fp = [ctypes.POINTER(ctypes.c_float)() for _ in range(3)]
        kp = ctypes.POINTER(ctypes.c_uint32)()
        n = ctypes.c_uint64(0)
        t = ctypes.c_double(0.0)
        self._check(self._lib.sph_readback_lock(self._h, int(slot), ctypes.byref(fp[0]), ctypes.byref(fp[1]),
                                                ctypes.byref(fp[2]), ctypes.byref(kp), ctypes.byref(n),
                                                ctypes.byref(t)))
        N = int(n.value)
        if N == 0:
            z = np.zeros((0, 3), np.float32)
            return z, z.copy(), np.zeros(0, np.float32), np.zeros(0, np.uint32), float(t.value)
        pos = np.ctypeslib.as_array(fp[0], shape=(N, 3)).copy()
        vel = np.ctypeslib.as_array(fp[1], shape=(N, 3)).copy()
        p = np.ctypeslib.as_array(fp[2], shape=(N,)).copy()
        keys = np.ctypeslib.as_array(kp, shape=(N,)).copy()
        return pos, vel, p, keys, float(t.value)

    def copy_probe(self, nbytes: int, repeats: int = 1):
        """Plain device copy of nbytes (counter calibration under a profiler)."""
        self._check(self._lib.sph_copy_probe(self._h, int(nbytes), int(repeats)))

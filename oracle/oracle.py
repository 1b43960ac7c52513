"""ctypes binding of the CPU oracle (oracle/libsph_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker and by bench.py's cpu_baseline leg. The oracle restates the reference's
multithreaded CPU engine (SimulationEngineCPU.cpp + SPHFunctions.cpp); see the
header of oracle/sph_oracle.cpp for how each of its functions is pinned.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsph_oracle.so")

FLAG_RECORD_DEBUG = 1
FLAG_PARALLEL_PARTIAL_SUM = 2

_P = ctypes.c_void_p
_F32P = ctypes.POINTER(ctypes.c_float)
_F64P = ctypes.POINTER(ctypes.c_double)
_U32P = ctypes.POINTER(ctypes.c_uint32)

_SIGS = {
    "sph_oracle_create": (_P, [_F32P, _F32P, _F32P, ctypes.c_uint64, _F32P, ctypes.c_uint64, ctypes.c_uint64,
                               ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "sph_oracle_destroy": (None, [_P]),
    "sph_oracle_set_triangles": (ctypes.c_int, [_P, _F32P, ctypes.c_uint64]),
    "sph_oracle_set_reorder_interval": (ctypes.c_int, [_P, ctypes.c_float]),
    "sph_oracle_update": (ctypes.c_int, [_P, ctypes.c_float, ctypes.c_uint32]),
    "sph_oracle_hash_size": (ctypes.c_uint64, [_P]),
    "sph_oracle_simulation_time": (ctypes.c_double, [_P]),
    "sph_oracle_read": (ctypes.c_int, [_P, _F64P, _F64P, _F64P]),
    "sph_oracle_write": (ctypes.c_int, [_P, _F32P, _F32P]),
    "sph_oracle_read_keys": (ctypes.c_int, [_P, _U32P]),
    "sph_oracle_read_hash_map": (ctypes.c_int, [_P, _U32P]),
    "sph_oracle_read_bucket_ids": (ctypes.c_int, [_P, _U32P]),
    "sph_oracle_read_static": (ctypes.c_int, [_P, _U32P, _U32P]),
    "sph_oracle_read_debug": (ctypes.c_int, [_P, _F64P, _F64P]),
    "sph_oracle_conserved": (ctypes.c_int, [_P, _F64P]),
    "sph_oracle_cells": (None, [_F32P, ctypes.c_uint64, ctypes.c_float, _U32P]),
    "sph_oracle_morton": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
    "sph_oracle_keys": (None, [_F32P, ctypes.c_uint64, ctypes.c_float, ctypes.c_uint64, _U32P]),
    "sph_oracle_kernels": (None, [ctypes.c_float, ctypes.c_float, _F32P]),
    "sph_oracle_random_direction": (None, [ctypes.c_float, _F32P]),
    "sph_oracle_noise": (ctypes.c_float, [ctypes.c_float]),
    "sph_oracle_triangle_force": (None, [_F32P, _F32P, _F32P]),
}

_lib = None


def load(build_if_missing: bool = True):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH) and build_if_missing:
        subprocess.run(["make", "-s", "libsph_oracle.so"], cwd=HERE, check=True)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a, t):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(t))


class Oracle:
    """The restated reference CPU engine on one scene."""

    def __init__(self, params, pos, vel=None, static_pos=None, hash_size: int = 0, threads: int = 1,
                 fp64: bool = False, flags: int = 0):
        lib = load()
        self._lib = lib
        prm = _f32(params.as_array() if hasattr(params, "as_array") else params)
        pos = _f32(pos).reshape(-1, 3)
        vel = None if vel is None else _f32(vel).reshape(-1, 3)
        spos = np.zeros((0, 3), np.float32) if static_pos is None else _f32(static_pos).reshape(-1, 3)
        self.n = pos.shape[0]
        self.n_static = spos.shape[0]
        self._keep = (prm, pos, vel, spos)
        self._h = lib.sph_oracle_create(_p(prm, ctypes.c_float), _p(pos, ctypes.c_float), _p(vel, ctypes.c_float),
                                        self.n, _p(spos, ctypes.c_float) if self.n_static else None, self.n_static,
                                        int(hash_size), int(threads), 1 if fp64 else 0, int(flags))

    def close(self):
        if self._h:
            self._lib.sph_oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def M(self) -> int:
        return int(self._lib.sph_oracle_hash_size(self._h))

    def set_triangles(self, tris):
        t = _f32(tris).reshape(-1, 9)
        self._lib.sph_oracle_set_triangles(self._h, _p(t, ctypes.c_float), t.shape[0])

    def set_reorder_interval(self, interval: float):
        self._lib.sph_oracle_set_reorder_interval(self._h, float(interval))

    def update(self, dt: float, steps: int = 1):
        rc = self._lib.sph_oracle_update(self._h, float(dt), int(steps))
        assert rc == 0

    def simulation_time(self) -> float:
        return float(self._lib.sph_oracle_simulation_time(self._h))

    def read(self):
        pos = np.empty((self.n, 3))
        vel = np.empty((self.n, 3))
        p = np.empty(self.n)
        self._lib.sph_oracle_read(self._h, _p(pos, ctypes.c_double), _p(vel, ctypes.c_double),
                                  _p(p, ctypes.c_double))
        return pos, vel, p

    def write(self, pos=None, vel=None):
        """LockWrite of the current buffer: positions and/or velocities by original
        index; hash, pressure and tables unchanged until the next update."""
        p = None if pos is None else _f32(pos).reshape(self.n, 3)
        v = None if vel is None else _f32(vel).reshape(self.n, 3)
        self._keep_write = (p, v)
        assert self._lib.sph_oracle_write(self._h, _p(p, ctypes.c_float), _p(v, ctypes.c_float)) == 0

    def read_keys(self):
        k = np.empty(self.n, np.uint32)
        self._lib.sph_oracle_read_keys(self._h, _p(k, ctypes.c_uint32))
        return k

    def read_hash_map(self):
        t = np.empty(self.M + 1, np.uint32)
        self._lib.sph_oracle_read_hash_map(self._h, _p(t, ctypes.c_uint32))
        return t

    def read_bucket_ids(self):
        ids = np.empty(self.n, np.uint32)
        self._lib.sph_oracle_read_bucket_ids(self._h, _p(ids, ctypes.c_uint32))
        return ids

    def read_static(self):
        t = np.empty(self.n_static + 1, np.uint32)
        ids = np.empty(max(self.n_static, 1), np.uint32)
        self._lib.sph_oracle_read_static(self._h, _p(t, ctypes.c_uint32), _p(ids, ctypes.c_uint32))
        return t, ids[:self.n_static]

    def read_debug(self):
        d = np.empty(self.n)
        a = np.empty((self.n, 3))
        rc = self._lib.sph_oracle_read_debug(self._h, _p(d, ctypes.c_double), _p(a, ctypes.c_double))
        assert rc == 0, "oracle was created without FLAG_RECORD_DEBUG or has not stepped"
        return d, a

    def conserved(self):
        out = np.empty(8)
        self._lib.sph_oracle_conserved(self._h, _p(out, ctypes.c_double))
        return out


# ---- free functions ---------------------------------------------------------

def cells(pos, h: float):
    pos = _f32(pos).reshape(-1, 3)
    out = np.empty_like(pos, dtype=np.uint32)
    load().sph_oracle_cells(_p(pos, ctypes.c_float), pos.shape[0], float(h), _p(out, ctypes.c_uint32))
    return out


def morton(x: int, y: int, z: int) -> int:
    return int(load().sph_oracle_morton(x, y, z))


def keys(pos, h: float, M: int):
    pos = _f32(pos).reshape(-1, 3)
    out = np.empty(pos.shape[0], np.uint32)
    load().sph_oracle_keys(_p(pos, ctypes.c_float), pos.shape[0], float(h), int(M), _p(out, ctypes.c_uint32))
    return out


def kernels(r: float, h: float):
    out = np.empty(4, np.float32)
    load().sph_oracle_kernels(float(r), float(h), _p(out, ctypes.c_float))
    return out


def random_direction(x: float):
    out = np.empty(3, np.float32)
    load().sph_oracle_random_direction(float(x), _p(out, ctypes.c_float))
    return out


def noise(x: float) -> float:
    return float(load().sph_oracle_noise(float(x)))


def triangle_force(p, tri9):
    p = _f32(p)
    t = _f32(tri9)
    out = np.empty(3, np.float32)
    load().sph_oracle_triangle_force(_p(p, ctypes.c_float), _p(t, ctypes.c_float), _p(out, ctypes.c_float))
    return out

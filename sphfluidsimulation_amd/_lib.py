"""ctypes binding of the C ABI declared in include/sph_hip.h.

The shared library is built in-tree (sphfluidsimulation_amd/libsph_hip.so) by
sphfluidsimulation_amd/build.py. There is no fallback: if the library is missing
or does not load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsph_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "sph_hip.h")

SPH_OK = 0
SPH_E_INVALID = -1
SPH_E_HIP = -2
SPH_E_STATE = -3
SPH_E_VALIDATION = -4
SPH_E_NO_DEVICE = -5

SPH_FLAG_HIP_GRAPH = 0x1
SPH_FLAG_RECORD_DEBUG = 0x2
SPH_FLAG_FP16_OFFSETS = 0x4
SPH_FLAG_CELL_WALK = 0x8
SPH_FLAG_RUN_WALK = 0x10
SPH_FLAG_TILE_WALK = 0x20
SPH_FLAG_CW_LISTS = 0x40
SPH_FLAG_CW_MASKS = 0x80
SPH_CW_LIST_MAX = 1572864
SPH_FLAG_PARTICLE_MAP = 0x100
SPH_FLAG_FP16_STORED = 0x200
SPH_NOALLOC_MAX_PARTICLES = 1 << 20

SPH_V_NONFINITE = 0x1
SPH_V_KEY = 0x2
SPH_V_TABLE_FIRST = 0x4
SPH_V_TABLE_LAST = 0x8
SPH_V_TABLE_ORDER = 0x10
SPH_V_BUCKET = 0x20
SPH_V_STATIC = 0x40

STAGES = ("density", "force", "scan", "scatter")
SPH_PROFILE_ALL = 0xF
SPH_PROFILE_ROTATE = 0x10


class SPHError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"SPH error {code}: {message}")
        self.code = code


class sph_params(ctypes.Structure):
    _fields_ = [(name, ctypes.c_float) for name in (
        "particleMass", "gasConstant", "elasticity", "viscosity", "gravityX", "gravityY", "gravityZ",
        "restDensity", "maxInteractionDistance", "selfDensity", "smoothingKernelConstant")]


_P = ctypes.c_void_p
_F32P = ctypes.POINTER(ctypes.c_float)
_U32P = ctypes.POINTER(ctypes.c_uint32)
_U64P = ctypes.POINTER(ctypes.c_uint64)
_F64P = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes)
SIGNATURES = {
    "sph_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int]),
    "sph_destroy": (ctypes.c_int, [_P]),
    "sph_system_implementation_name": (ctypes.c_char_p, []),
    "sph_initialize": (ctypes.c_int, [_P, ctypes.POINTER(sph_params), _F32P, _F32P, ctypes.c_uint64, _F32P,
                                      ctypes.c_uint64, ctypes.c_uint32]),
    "sph_batch_initialize": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.POINTER(sph_params), _U64P, _F32P, _F32P,
                                            _U64P, _F32P, ctypes.c_uint32]),
    "sph_set_triangles": (ctypes.c_int, [_P, ctypes.c_uint32, _F32P, ctypes.c_uint64]),
    "sph_update": (ctypes.c_int, [_P, ctypes.c_float, ctypes.c_uint32]),
    "sph_synchronize": (ctypes.c_int, [_P]),
    "sph_read": (ctypes.c_int, [_P, _F32P, _F32P, _F32P]),
    "sph_read_noalloc": (ctypes.c_int, [_P, _F32P, _F32P, _F32P]),
    "sph_write": (ctypes.c_int, [_P, _F32P, _F32P]),
    "sph_set_reorder_interval": (ctypes.c_int, [_P, ctypes.c_float]),
    "sph_set_launch_shape": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32]),
    "sph_launch_shape": (ctypes.c_int, [_P, _U32P, _U32P]),
    "sph_read_keys": (ctypes.c_int, [_P, _U32P]),
    "sph_read_cell_starts": (ctypes.c_int, [_P, _U32P]),
    "sph_read_bucket_ids": (ctypes.c_int, [_P, _U32P]),
    "sph_read_static": (ctypes.c_int, [_P, _U32P, _U32P]),
    "sph_read_debug": (ctypes.c_int, [_P, _F32P, _F32P]),
    "sph_conserved": (ctypes.c_int, [_P, _F64P]),
    "sph_validate": (ctypes.c_int, [_P, _U32P]),
    "sph_simulation_time": (ctypes.c_float, [_P]),
    "sph_clear": (ctypes.c_int, [_P]),
    "sph_last_error": (ctypes.c_char_p, [_P]),
    "sph_particle_count": (ctypes.c_uint64, [_P]),
    "sph_static_particle_count": (ctypes.c_uint64, [_P]),
    "sph_scene_count": (ctypes.c_uint32, [_P]),
    "sph_hash_map_size": (ctypes.c_uint64, [_P, ctypes.c_uint32]),
    "sph_set_profiling": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "sph_stage_times": (ctypes.c_int, [_P, _F64P, _U64P]),
    "sph_copy_probe": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_uint32]),
    "sph_autotune": (ctypes.c_int, [_P, ctypes.c_float, ctypes.c_uint32, _U32P]),
    "sph_readback_enqueue": (ctypes.c_int, [_P, _U32P]),
    "sph_readback_query": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int)]),
    "sph_readback_lock": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.POINTER(_F32P), ctypes.POINTER(_F32P),
                                         ctypes.POINTER(_F32P), ctypes.POINTER(_U32P), _U64P,
                                         ctypes.POINTER(ctypes.c_double)]),
}

_lib = None


def header_symbols(path: str = HEADER_PATH):
    """Function names declared in include/sph_hip.h."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_[a-z0-9_]+)\s*\(", text)))


def load():
    """Load libsph_hip.so (once) and declare its signatures. Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run sphfluidsimulation_amd.build.build_all() "
                          "(or __graft_entry__.build()) first")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(handle, rc: int) -> int:
    if rc != SPH_OK:
        msg = load().sph_last_error(handle)
        raise SPHError(rc, msg.decode() if msg else "")
    return rc

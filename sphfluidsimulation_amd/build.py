"""Build recipe: compiles the gfx950 engine library in-tree and the CPU oracle.

    python -m sphfluidsimulation_amd.build

produces sphfluidsimulation_amd/libsph_hip.so (hipcc, --offload-arch=gfx950)
and oracle/libsph_oracle.so (g++; test infrastructure only).
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libsph_hip.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # No packed FP32 instructions (v_pk_*_f32) anywhere in the device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, cwd=None):
    print("+ " + " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_hip():
    src = os.path.join(CSRC, "sph_engine.hip")
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, src])
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    _run(["make", "-B", "-s", "libsph_oracle.so"], cwd=ORACLE_DIR)
    return os.path.join(ORACLE_DIR, "libsph_oracle.so")


ADAPTER_SRC = os.path.join(ROOT, "tests", "cpp", "adapter_demo.cpp")
ADAPTER_BIN = os.path.join(ROOT, "tests", "cpp", "adapter_demo")


def build_adapter_demo():
    """The C++ SimulationEngineHIP adapter driven like the reference's application
    (a host-only g++ build against libsph_hip.so)."""
    _run(["g++", "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"), "-o", ADAPTER_BIN, ADAPTER_SRC, LIB,
          "-Wl,-rpath," + PKG])
    return ADAPTER_BIN


REFERENCE_DIR = os.path.join(PKG, "reference")
REFERENCE_SRCS = [
    os.path.join(REFERENCE_DIR, "source", "SPH", "SimulationEngines", "SimulationEngineHIP.cpp"),
    os.path.join(REFERENCE_DIR, "source", "SPH", "ParticleBufferManagers", "HIPParticleBufferManager.cpp"),
]
MIRROR_DIR = os.path.join(ROOT, "tests", "cpp", "refmirror")
BINDING_SRC = os.path.join(ROOT, "tests", "cpp", "reference_binding.cpp")
BINDING_BIN = os.path.join(ROOT, "tests", "cpp", "reference_binding")


def reference_binding_cmd(extra=()):
    """g++ line for the drop-in reference engine (sphfluidsimulation_amd/reference/)
    against the declaration mirror of the reference's interface (tests/cpp/refmirror/)."""
    return ["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", "-Wno-unused-parameter", *extra,
            "-I" + MIRROR_DIR, "-I" + os.path.join(REFERENCE_DIR, "include"), "-I" + os.path.join(ROOT, "include"),
            *REFERENCE_SRCS, os.path.join(MIRROR_DIR, "refmirror.cpp"), BINDING_SRC]


def build_reference_binding():
    """The drop-in SimulationEngineHIP driven through the reference's interface
    (a host-only g++ build against libsph_hip.so)."""
    _run(reference_binding_cmd() + ["-o", BINDING_BIN, LIB, "-Wl,-rpath," + PKG])
    return BINDING_BIN


WALK_SRC = os.path.join(ROOT, "tests", "cpp", "walk_check.cpp")
WALK_BIN = os.path.join(ROOT, "tests", "cpp", "walk_check")


def build_walk_check():
    """Host-side check of the device neighbour walk (compiled by hipcc, runs on the CPU)."""
    _run([HIPCC, "-O2", "-std=c++17", "-o", WALK_BIN, WALK_SRC])
    return WALK_BIN


WALL_LISTS_SRC = os.path.join(ROOT, "tests", "cpp", "wall_lists_check.cpp")
WALL_LISTS_BIN = os.path.join(ROOT, "tests", "cpp", "wall_lists_check")


def build_wall_lists_check():
    """Host-side check of the density pass's wall lists (compiled by hipcc, runs on the CPU)."""
    _run([HIPCC, "-O2", "-std=c++17", "-o", WALL_LISTS_BIN, WALL_LISTS_SRC])
    return WALL_LISTS_BIN


def build_all():
    build_hip()
    build_oracle()
    build_adapter_demo()
    build_reference_binding()
    build_walk_check()
    build_wall_lists_check()


if __name__ == "__main__":
    build_all()
    sys.exit(0)

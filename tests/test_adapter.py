"""The C++ SimulationEngineHIP adapter (include/sph/simulation_engine_hip.hpp)
driven as the reference's application drives an engine: Initialize, Update per
frame, read the dynamic buffer manager."""
import os
import subprocess

import numpy as np
import pytest

from sphfluidsimulation_amd import SimulationEngineHIP, scenes
from sphfluidsimulation_amd.build import ADAPTER_BIN, ADAPTER_SRC

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def scene_text(sc, frames, dt):
    lines = [str(sc.n)] + [" ".join(f"{v:.9g}" for v in p) for p in sc.pos.astype(np.float64)]
    lines += [str(sc.n_static)] + [" ".join(f"{v:.9g}" for v in p) for p in sc.static_pos.astype(np.float64)]
    prm = sc.params.as_array()[:9]
    lines.append(" ".join(f"{float(v):.9g}" for v in prm))
    lines.append(f"{frames} {dt:.9g}")
    return "\n".join(lines) + "\n"


def test_adapter_header_compiles(tmp_path):
    """Host-only compile of the adapter against the C ABI header (no GPU needed)."""
    out = tmp_path / "adapter.o"
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                    "-I" + os.path.join(ROOT, "include"), ADAPTER_SRC], check=True)
    assert not out.exists()


@pytest.mark.gpu
def test_adapter_matches_python_host():
    print("[gpu-test] C++ adapter vs Python host", flush=True)
    assert os.path.exists(ADAPTER_BIN), "run __graft_entry__.build() first"
    sc = scenes.dam_break(10, 6, 8, (1.0, 0.6, 0.8), seed=21, shuffle=True, name="adapter")
    r = subprocess.run([ADAPTER_BIN], input=scene_text(sc, 1, 1e-3), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    head, *rows = r.stdout.strip().splitlines()
    name, n, t = head.split()
    assert name == "HIP" and int(n) == sc.n and float(t) == pytest.approx(1e-3)
    vals = np.array([[float(x) for x in row.split()] for row in rows])
    e = SimulationEngineHIP(0)
    e.Initialize(sc)
    e.Update(1e-3, 1)
    pos, vel, p = e.read()
    keys = e.read_keys()
    e.close()
    # one step reads the same sorted input in both hosts: identical results
    np.testing.assert_array_equal(vals[:, 0:3].astype(np.float32), pos)
    np.testing.assert_array_equal(vals[:, 3].astype(np.float32), p)
    np.testing.assert_array_equal(vals[:, 4:7].astype(np.float32), vel)
    np.testing.assert_array_equal(vals[:, 7].astype(np.uint32), keys)

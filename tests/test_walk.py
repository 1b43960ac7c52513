"""The device neighbour walk visits exactly the reference's buckets.

tests/cpp/walk_check.cpp is compiled from the same header as the kernels
(sphfluidsimulation_amd/csrc/sph_math.hpp) and runs on the CPU: for random cells
(around 0, around the 1024 alias, anywhere in uint32) and bucket counts M from 1
to 2^31, the multiset of buckets visited by the 14-run walk equals the 27
buckets of the reference's per-cell loop (SPHFunctions.cpp:397-404), and the
divide-free remainder equals %."""
import os
import subprocess

from sphfluidsimulation_amd.build import WALK_BIN


def test_walk_visits_reference_buckets():
    assert os.path.exists(WALK_BIN), "run __graft_entry__.build() first"
    r = subprocess.run([WALK_BIN], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 mismatches" in r.stdout

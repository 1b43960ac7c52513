"""The device neighbour walk visits exactly the reference's buckets.

tests/cpp/walk_check.cpp is compiled from the same header as the kernels
(sphfluidsimulation_amd/csrc/sph_math.hpp) and runs on the CPU: for random cells
(around 0, around the 1024 alias, anywhere in uint32) and bucket counts M from 1
to 2^31, the multiset of buckets visited by the 14-run walk equals the 27
buckets of the reference's per-cell loop (SPHFunctions.cpp:397-404), and the
divide-free remainder equals %."""
import os
import subprocess

from sphfluidsimulation_amd.build import WALK_BIN, WALL_LISTS_BIN


def test_walk_visits_reference_buckets():
    assert os.path.exists(WALK_BIN), "run __graft_entry__.build() first"
    r = subprocess.run([WALK_BIN], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 mismatches" in r.stdout


def test_wall_lists_keep_what_the_bucket_walk_keeps():
    """tests/cpp/wall_lists_check.cpp, from the kernels' header
    (sphfluidsimulation_amd/csrc/sph_wall_lists.hpp): for wall scenes on and off
    the lattice, far from the origin (to x = 2e5, where an FP32 unit in the last
    place is h / 6), below zero and with colliding static buckets, the walls an
    octant's list keeps under the exact test are the walls the reference's
    27-bucket walk keeps (SPHFunctions.cpp:455-496), in the same order with the
    same duplicates, for particles at random and on cell and half-cell
    boundaries. A reach short of h by 2^-7, or boxes not enlarged for FP32
    rounding, each make it fail."""
    assert os.path.exists(WALL_LISTS_BIN), "run __graft_entry__.build() first"
    r = subprocess.run([WALL_LISTS_BIN], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 mismatches" in r.stdout.splitlines()[-1]


def test_candidate_layout_positions():
    """The cell-wave candidate layout (sph_cellwave.hpp: kSlots, slot_bucket) puts
    slot l % kSlots of bucket l / kSlots at position l of four 64-lane rounds, with
    l / 9 computed as (57 l) >> 9: every position below 27 kSlots maps to one
    (bucket, slot) pair, bucket < 27, and the 27 buckets fit the 256 positions."""
    import re
    hdr = os.path.join(os.path.dirname(__file__), "..", "sphfluidsimulation_amd", "csrc", "sph_cellwave.hpp")
    k = int(re.search(r"#define SPH_CW_SLOTS (\d+)", open(hdr).read()).group(1))
    assert k in (8, 9) and 27 * k <= 256
    seen = set()
    for pos in range(27 * k):
        b = pos >> 3 if k == 8 else (pos * 57) >> 9
        assert b == pos // k and b < 27
        seen.add((b, pos - k * b))
    assert len(seen) == 27 * k and all(s < k for _, s in seen)

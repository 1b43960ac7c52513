"""bench.py's host logic (no GPU): the HBM traffic it reports comes from the
committed rocprofv3 summary of its own step window when there is one."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def _traffic_line(path):
    for line in open(path):
        if line.startswith("traffic_bytes:"):
            return float(line.split(":", 1)[1])
    raise AssertionError(f"no traffic_bytes line in {path}")


def test_traffic_of_the_20_step_window_comes_from_its_own_summary():
    window = os.path.join(ROOT, "profiles", "k_density_cw_dam-break-1m-20-steps_pmc.md")
    whole = os.path.join(ROOT, "profiles", "k_density_cw_dam-break-1m_pmc.md")
    assert bench.load_traffic("k_density_cw", "dam-break-1m", 20) == _traffic_line(window)
    # Any other window falls back to the workload's summary (1000 steps).
    assert bench.load_traffic("k_density_cw", "dam-break-1m", 1000) == _traffic_line(whole)
    assert bench.load_traffic("k_density_cw", "dam-break-1m", 200) == _traffic_line(whole)
    # The early steps move fewer bytes than the whole run.
    assert _traffic_line(window) < _traffic_line(whole)


def test_traffic_of_an_unprofiled_workload_is_null():
    assert bench.load_traffic("k_density_cw", "no-such-workload", 20) is None


def test_valu_busy_of_the_window_and_of_the_run():
    busy20, src20 = bench.load_valu_busy("k_density_cw", "dam-break-1m", 20)
    busy, src = bench.load_valu_busy("k_density_cw", "dam-break-1m", 1000)
    assert src20.endswith("k_density_cw_dam-break-1m-20-steps_sq.md")
    assert src.endswith("k_density_cw_dam-break-1m_sq.md")
    assert 0.0 < busy < busy20 <= 1.0
    assert bench.load_valu_busy("k_density_cw", "no-such-workload", 20) == (None, None)

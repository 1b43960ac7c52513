"""The reference application's simulation profiles (sphfluidsimulation_amd.profiles).

The fixtures are the reference's own profile files
(FluidSimulationProject/assets/simulationProfiles/, copied by
tests/golden/copy_scene_assets.py). The CPU tests check that every file loads,
that each profile's scene has the particle counts of the reference's
generators (FilledBoxParticleSetBlueprint.cpp:37-41,
BoxShellParticleSetBlueprint.cpp:30-41, evaluated here from the files' numbers),
the reorder intervals the experiments step through, the update count of
ProfilingScene.cpp:118 and the report layout of ProfilingScene.cpp:179-251.
The GPU test runs a profile end to end on the HIP engine.
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from sphfluidsimulation_amd.profiles import (ProfileError, load_profiles, parse_profiles, profile_report,
                                             run_profiles, update_count)

PROFILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "app", "assets", "simulationProfiles")

FILES = {
    "increasingBoxSizeProfiles.json": 28,
    "increasingReorderTimeIntervalProfiles.json": 21,
    "increasingStaticParticleCountProfiles.json": 3,
    "systemProfilingProfiles.json": 1,
    "systemVisualisationProfile.json": 1,
    "waveDetailedProfilingProfile.json": 1,
    "waveIncreasingParticleReorderTimeIntervalDetailedProfilingProfiles.json": 6,
    "waveProfile.json": 1,
}


def _filled_box_count(gen) -> int:
    # Vec3<uintMem>(spawnVolumeSize * Pow(particlesPerUnit, 1/3)), single precision.
    lin = np.float32(np.float32(gen["particlesPerUnit"]) ** np.float32(1.0 / 3.0))
    g = [int(np.float32(s) * lin) for s in gen["spawnVolumeSize"]]
    return g[0] * g[1] * g[2]


def _box_shell_count(gen) -> int:
    # Vec3<uintMem>(spawnVolumeSize / particleDistance) + 1, then the six faces
    # without their shared edges and corners counted twice, times two layers.
    gx, gy, gz = (int(np.float32(s) / np.float32(gen["particleDistance"])) + 1 for s in gen["spawnVolumeSize"])
    return (gx * gy + gx * gz + gy * gz - 2 * gx - 2 * gy - 2 * gz + 4) * 2


class _no_warning:
    """Turns warnings into errors inside the block."""

    def __enter__(self):
        import warnings
        self._cm = warnings.catch_warnings()
        self._cm.__enter__()
        warnings.simplefilter("error")

    def __exit__(self, *exc):
        return self._cm.__exit__(*exc)


@pytest.mark.parametrize("name,count", sorted(FILES.items()))
def test_every_reference_profile_file_loads(name, count):
    with pytest.warns(UserWarning) if name == "systemVisualisationProfile.json" else _no_warning():
        profiles = load_profiles(os.path.join(PROFILES, name))
    assert len(profiles) == count
    for p in profiles:
        assert p.step_time > 0 and p.simulation_duration > 0 and p.steps_per_update >= 1
        assert p.reorder_time_interval is None or p.reorder_time_interval >= 0.0


def test_wave_profile_scene_counts_and_parameters():
    path = os.path.join(PROFILES, "systemProfilingProfiles.json")
    with open(path) as fh:
        raw = json.load(fh)["profiles"][0]
    (p,) = load_profiles(path)
    sp = raw["systemParameters"]
    assert p.name == "wave" and p.output_file_path == "outputs/wave-ACER.txt"
    assert p.scene.n == _filled_box_count(sp["dynamicParticleGenerationParameters"]) == 76 * 26 * 26
    assert p.scene.n_static == _box_shell_count(sp["staticParticleGenerationParameters"])
    bp = sp["particleBehaviourParameters"]
    prm = p.scene.params
    assert (prm.particleMass, prm.gasConstant, prm.elasticity, prm.viscosity, prm.restDensity,
            prm.maxInteractionDistance) == (bp["particleMass"], bp["gasConstant"], bp["elasticity"], bp["viscosity"],
                                            bp["restDensity"], bp["maxInteractionDistance"])
    assert (prm.gravityX, prm.gravityY, prm.gravityZ) == tuple(float(g) for g in bp["gravity"])
    # The fluid block lies inside its spawn volume, the walls on the box's faces.
    lo = np.float32(sp["dynamicParticleGenerationParameters"]["spawnVolumeOffset"])
    hi = lo + np.float32(sp["dynamicParticleGenerationParameters"]["spawnVolumeSize"])
    assert np.all(p.scene.pos >= lo - 0.5) and np.all(p.scene.pos <= hi + 0.5)
    assert abs(p.reorder_time_interval - 0.1) < 1e-7
    assert update_count(p) == 20  # 20 s of 0.01 s steps, 100 steps an update


def test_reorder_interval_experiment_steps_through_its_intervals():
    profiles = load_profiles(os.path.join(PROFILES, "increasingReorderTimeIntervalProfiles.json"))
    intervals = [p.reorder_time_interval for p in profiles]
    # The profile names are the intervals the experiment steps through, 0 to 2 s.
    np.testing.assert_allclose(intervals, [float(p.name) for p in profiles], atol=1e-6)
    np.testing.assert_allclose(intervals, np.arange(21) * 0.1, atol=1e-5)
    assert [p.output_file_path for p in profiles][:2] == ["outputs/increasingReorderTimeIntervalProfiles/0.000000s.txt",
                                                          "outputs/increasingReorderTimeIntervalProfiles/0.100000s.txt"]


def test_visualisation_profile_takes_the_applications_frame_step():
    with pytest.warns(UserWarning, match="hashesPerDynamicParticle"):
        (p,) = load_profiles(os.path.join(PROFILES, "systemVisualisationProfile.json"))
    assert (p.step_time, p.steps_per_update, update_count(p)) == (0.01, 1, 100)


def test_update_count_in_single_precision():
    base = {"particleBehaviourParameters": {"particleMass": 1.0, "gasConstant": 1, "elasticity": 0.5,
                                            "viscosity": 0.1, "gravity": [0, -9.81, 0], "restDensity": 1,
                                            "maxInteractionDistance": 1.0},
            "dynamicParticleGenerationParameters": {"generatorType": "FilledBoxParticleGenerator",
                                                    "spawnVolumeSize": [1, 1, 1], "spawnVolumeOffset": [0, 0, 0],
                                                    "particlesPerUnit": 8, "randomOffsetIntensity": 0}}

    def prof(duration, dt, spu):
        return parse_profiles({"profiles": [{"profileName": "x", "outputFilePath": "x.txt", "simulationDuration":
                                             duration, "simulationStepTime": dt, "stepsPerUpdate": spu,
                                             "systemParameters": base}]})[0]
    assert update_count(prof(20, 0.01, 100)) == 20
    assert update_count(prof(1, 0.3, 1)) == 4  # 3.33... updates: the fourth reaches the duration
    assert update_count(prof(0.5, 0.01, 100)) == 1
    p = prof(2, 0.01, 10)
    assert (p.scene.n, p.scene.n_static) == (8, 0)
    assert p.reorder_time_interval is None  # no interval given: the default mode


def test_report_layout():
    (p,) = load_profiles(os.path.join(PROFILES, "systemProfilingProfiles.json"))
    text = profile_report(p, ["HIP"], [[0.1234], [12.5]])
    lines = text.split("\n")
    assert lines[0] == 'Started profiling profile named "wave".'
    assert lines[1] == f"   Number of dynamic particles: {p.scene.n}"
    assert lines[2] == f"   Number of static particles:  {p.scene.n_static}"
    assert lines[3] == "   Simulation duration:         20s"
    assert lines[4] == "   Simulation step time:        0.01s"
    assert lines[5] == "   Steps per update:            100"
    assert lines[6] == ""
    assert lines[7] == "           HIP "
    assert lines[8] == "       0.123ms "
    assert lines[9] == "      12.500ms "


@pytest.mark.parametrize("mutate,message", [
    (lambda sp: sp.pop("particleBehaviourParameters"), "particleBehaviourParameters"),
    (lambda sp: sp["dynamicParticleGenerationParameters"].update(generatorType="Sphere"), "generatorType"),
    (lambda sp: sp["particleBoundParameters"].update(bounded=True), "bounded"),
    (lambda sp: sp["dynamicParticleGenerationParameters"].pop("spawnVolumeSize"), "spawnVolumeSize"),
])
def test_invalid_profiles_are_rejected(mutate, message):
    with open(os.path.join(PROFILES, "systemProfilingProfiles.json")) as fh:
        doc = json.load(fh)
    mutate(doc["profiles"][0]["systemParameters"])
    with pytest.raises(ProfileError, match=message):
        p = parse_profiles(doc)[0]
        p.scene  # noqa: B018 (generation is lazy)
    with pytest.raises(ProfileError):
        parse_profiles("{not json")


@pytest.mark.parametrize("key,value", [("simulationStepTime", 0.0), ("stepsPerUpdate", 0),
                                       ("simulationDuration", -1.0)])
def test_invalid_stepping_is_rejected(key, value):
    with open(os.path.join(PROFILES, "systemProfilingProfiles.json")) as fh:
        doc = json.load(fh)
    doc["profiles"][0][key] = value
    with pytest.raises(ProfileError, match="stepsPerUpdate"):
        parse_profiles(doc)


@pytest.mark.gpu
def test_wave_profile_runs_on_the_hip_engine(tmp_path):
    print("[gpu-test] systemProfilingProfiles.json on the HIP engine, 2 updates", flush=True)
    written = run_profiles(os.path.join(PROFILES, "systemProfilingProfiles.json"), out_dir=str(tmp_path),
                           max_updates=2, log=lambda m: print(m, flush=True))
    assert written == [os.path.join(str(tmp_path), "outputs/wave-ACER.txt")]
    with open(written[0]) as fh:
        lines = fh.read().splitlines()
    assert lines[0] == 'Started profiling profile named "wave".'
    assert lines[7].split() == ["HIP"]
    rows = [float(line.strip()[:-2]) for line in lines[8:]]
    assert len(rows) == 2 and all(r > 0.0 for r in rows)

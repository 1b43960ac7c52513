"""The reference application's simulation profiles, run on the HIP engine.

The reference keeps its profiling experiments as JSON files
(FluidSimulationProject/assets/simulationProfiles/*.json): how large a scene to
generate, how to step it and, per engine, how often to reorder the particles
(the reorder-interval experiments of increasingReorderTimeIntervalProfiles.json
among them). A file holds a list of, or one, object {"profiles": [...]}, or a
single profile; each profile is

    {profileName, outputFilePath, simulationDuration, simulationStepTime,
     stepsPerUpdate, systemParameters}

and systemParameters (alone, it is the whole of a visualisation file such as
waveProfile.json) holds the two generators, particleBehaviourParameters,
particleBoundParameters, bufferCount, hashesPerDynamicParticle,
hashesPerStaticParticle and implementationSpecifics {CPU: {...}, OpenCL: {...}}.

The generators are the blueprints' (FilledBoxParticleGenerator:
FilledBoxParticleSetBlueprint.cpp:23-66; BoxShellParticleGenerator:
BoxShellParticleSetBlueprint.cpp:14-116) with their properties under the same
names; the profiles give randomOffsetIntensity as a real number, which is kept.
Random offsets come from numpy's PCG64 (BlazeEngine's Random is not available),
seeded with the generator's "seed" (default 0).

ProfilingScene (FluidSimulationProject/source/Scenes/ProfilingScene/
ProfilingScene.cpp:23-36, 107-155) steps each engine with
Update(simulationStepTime, stepsPerUpdate) until simulationDuration has passed
and then writes the report of ProfileFinished (:179-251): a header, the engines'
names, and one row per update of the time per step in ms. run_profile does the
same for the HIP engine (SystemImplementationName "HIP"); the reference's CPU
and OpenCL engines are not part of this package.

Reordering: the engines read one reorder time interval today,
"reorderTimeInterval" (SimulationEngineCPU.cpp:401, SimulationEngineGPU.cpp:142).
The profile files come from an earlier version of the engines' parameters and
carry two: "reorderTimeInterval", the same 0.1 s in every profile, and
"reorderInterval", the value the reorder-interval experiments step through (their
profile names are its values, 0.000000 to 2.000000). The experiment's value is
taken: "reorderInterval" of implementationSpecifics "HIP", else "OpenCL", else
"CPU", else their "reorderTimeInterval". The engine then runs in
SPH_FLAG_PARTICLE_MAP mode with that interval and the CPU engine's rule
(SimulationEngineCPU.cpp:280-316, 442-457): an update reorders when the elapsed
time exceeds the interval at its start, and the elapsed time restarts at 0 as
soon as it exceeds the interval, so, as in that engine, an interval of 0 or more
never reorders and a negative one reorders every update. A profile without an
interval runs the engine's default mode, which reorders every step.

Hash tables: the engines size the dynamic table at 2 N buckets and the static one
at one bucket per wall (SimulationEngine.h:79-114), as the HIP engine does;
"hashesPerDynamicParticle" and "hashesPerStaticParticle" (earlier parameters)
are read and ignored, with a warning when the dynamic one is not 2.

Command line:
    python -m sphfluidsimulation_amd.profiles FILE [--out DIR] [--max-updates N] [--device D] [--every-step]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
import warnings
from dataclasses import dataclass, field

import numpy as np

from .scenes import Params, Scene, box_shell, filled_box

__all__ = ["ProfileError", "Profile", "load_profiles", "parse_profiles", "update_count", "profile_report",
           "run_profile", "run_profiles"]


class ProfileError(ValueError):
    """The profile file is not valid JSON or misses a required entry."""


def _get(obj, key, kind=None, where=""):
    if not isinstance(obj, dict) or key not in obj:
        raise ProfileError(f"{where}missing entry '{key}'")
    v = obj[key]
    if kind == "vec3":
        if not (isinstance(v, list) and len(v) == 3 and all(isinstance(c, (int, float)) for c in v)):
            raise ProfileError(f"{where}entry '{key}' is not a 3-vector")
        return [float(c) for c in v]
    if kind == "float":
        if isinstance(v, bool) or not isinstance(v, (int, float)):
            raise ProfileError(f"{where}entry '{key}' is not a number")
        return float(v)
    if kind == "uint":
        if isinstance(v, bool) or not isinstance(v, (int, float)) or v < 0 or int(v) != v:
            raise ProfileError(f"{where}entry '{key}' is not an unsigned integer")
        return int(v)
    return v


def _params(j) -> Params:
    w = "particleBehaviourParameters: "
    g = _get(j, "gravity", "vec3", w)
    return Params(particleMass=_get(j, "particleMass", "float", w), gasConstant=_get(j, "gasConstant", "float", w),
                  elasticity=_get(j, "elasticity", "float", w), viscosity=_get(j, "viscosity", "float", w),
                  gravityX=g[0], gravityY=g[1], gravityZ=g[2], restDensity=_get(j, "restDensity", "float", w),
                  maxInteractionDistance=_get(j, "maxInteractionDistance", "float", w))


def _generate(gen, where) -> np.ndarray:
    kind = _get(gen, "generatorType", None, where)
    offset = _get(gen, "spawnVolumeOffset", "vec3", where)
    size = _get(gen, "spawnVolumeSize", "vec3", where)
    intensity = _get(gen, "randomOffsetIntensity", "float", where)
    rng = np.random.default_rng(_get(gen, "seed", "uint", where) if "seed" in gen else 0)
    if kind == "FilledBoxParticleGenerator":
        return filled_box(size, _get(gen, "particlesPerUnit", "uint", where), offset=offset,
                          random_offset_intensity=intensity, rng=rng)
    if kind == "BoxShellParticleGenerator":
        return box_shell(size, _get(gen, "particleDistance", "float", where), offset=offset, jitter=intensity,
                         rng=rng)
    raise ProfileError(f"{where}unknown generatorType '{kind}'")


@dataclass
class Profile:
    name: str
    output_file_path: str
    simulation_duration: float
    step_time: float
    steps_per_update: int
    system_parameters: dict
    reorder_time_interval: float | None = None  # the GPU engine's; None: the default mode
    buffer_count: int = 3
    _scene: Scene | None = field(default=None, repr=False)

    @property
    def scene(self) -> Scene:
        """The profile's particles, generated on first use (a profile may hold
        millions of walls)."""
        if self._scene is None:
            self._scene = _scene(self.system_parameters, self.name)
        return self._scene


def _check_system(sp, name) -> float | None:
    """Validates a systemParameters object; returns the reorder time interval
    (None when it gives none)."""
    w = f"profile '{name}': "
    if not isinstance(sp, dict):
        raise ProfileError(f"{w}systemParameters is not a JSON object")
    hashes = _get(sp, "hashesPerDynamicParticle", "uint", w) if "hashesPerDynamicParticle" in sp else 2
    if hashes != 2:
        warnings.warn(f"{w}hashesPerDynamicParticle {hashes} ignored: the dynamic table has 2 N buckets")
    bound = sp.get("particleBoundParameters", {})
    if isinstance(bound, dict) and bound.get("bounded", False):
        raise ProfileError(f"{w}bounded particles (particleBoundParameters.bounded) are not supported")
    _params(_get(sp, "particleBehaviourParameters", None, w))
    for key in ("dynamicParticleGenerationParameters", "staticParticleGenerationParameters"):
        if key in sp or key.startswith("dynamic"):
            g = _get(sp, key, None, w)
            kind = _get(g, "generatorType", None, f"{w}{key}: ")
            if kind not in ("FilledBoxParticleGenerator", "BoxShellParticleGenerator"):
                raise ProfileError(f"{w}{key}: unknown generatorType '{kind}'")
    spec = sp.get("implementationSpecifics", {}) or {}
    for key in ("reorderInterval", "reorderTimeInterval"):
        for system in ("HIP", "OpenCL", "CPU"):
            entry = spec.get(system)
            if isinstance(entry, dict) and key in entry:
                return _get(entry, key, "float", f"{w}implementationSpecifics.{system}: ")
    return None


def _scene(sp, name) -> Scene:
    w = f"profile '{name}': "
    params = _params(_get(sp, "particleBehaviourParameters", None, w))
    dyn = _generate(_get(sp, "dynamicParticleGenerationParameters", None, w), w + "dynamic generator: ")
    st = (_generate(sp["staticParticleGenerationParameters"], w + "static generator: ")
          if "staticParticleGenerationParameters" in sp else np.zeros((0, 3), np.float32))
    return Scene(params, np.ascontiguousarray(dyn, dtype=np.float32), np.zeros_like(dyn, dtype=np.float32),
                 np.ascontiguousarray(st, dtype=np.float32), name=name)


def parse_profiles(doc, default_name: str = "profile") -> list[Profile]:
    """The profiles of a parsed profile file (or its JSON text)."""
    if isinstance(doc, (str, bytes)):
        try:
            doc = json.loads(doc)
        except json.JSONDecodeError as ex:
            raise ProfileError(f"profile file is not valid JSON: {ex}") from ex
    groups = doc if isinstance(doc, list) else [doc]
    out: list[Profile] = []
    for g in groups:
        if not isinstance(g, dict):
            raise ProfileError("a profile group is not a JSON object")
        if "profiles" not in g and "systemParameters" not in g:
            # A bare systemParameters object (the visualisation scene's file): one
            # profile with the application's frame step, Update(0.01, 1)
            # (SimulationVisualisationScene.cpp:154-166), for one second.
            interval = _check_system(g, default_name)
            out.append(Profile(default_name, f"{default_name}.txt", 1.0, 0.01, 1, g, interval,
                               int(g.get("bufferCount", 3))))
            continue
        for p in (_get(g, "profiles") if "profiles" in g else [g]):  # a list of profiles, or one
            name = str(_get(p, "profileName"))
            w = f"profile '{name}': "
            sp = _get(p, "systemParameters", None, w)
            interval = _check_system(sp, name)
            duration = _get(p, "simulationDuration", "float", w)
            dt = _get(p, "simulationStepTime", "float", w)
            spu = _get(p, "stepsPerUpdate", "uint", w)
            if not (dt > 0.0 and duration >= 0.0 and spu >= 1 and math.isfinite(duration)):
                raise ProfileError(f"{w}needs simulationStepTime > 0, a finite simulationDuration >= 0 and "
                                   "stepsPerUpdate >= 1")
            out.append(Profile(name, str(_get(p, "outputFilePath", None, w)), duration, dt, spu, sp, interval,
                               int(sp.get("bufferCount", 3))))
    return out


def load_profiles(path: str) -> list[Profile]:
    with open(path, "r", encoding="utf-8") as fh:
        text = fh.read()
    return parse_profiles(text, default_name=os.path.splitext(os.path.basename(path))[0])


def update_count(p: Profile) -> int:
    """Updates ProfilingScene runs (ProfilingScene.cpp:118): it stops once the
    update count reaches simulationDuration / simulationStepTime /
    stepsPerUpdate, evaluated in single precision."""
    q = np.float32(p.simulation_duration) / np.float32(p.step_time) / np.float32(max(p.steps_per_update, 1))
    return max(int(math.ceil(float(q))), 0)


def _num(x: float) -> str:
    return f"{x:g}"


def profile_report(p: Profile, systems: list[str], rows: list[list[float]]) -> str:
    """The report ProfileFinished writes (ProfilingScene.cpp:179-251): the header,
    the engines' names, then per update each engine's time per step in ms."""
    out = (f'Started profiling profile named "{p.name}".\n'
           f"   Number of dynamic particles: {p.scene.n}\n"
           f"   Number of static particles:  {p.scene.n_static}\n"
           f"   Simulation duration:         {_num(p.simulation_duration)}s\n"
           f"   Simulation step time:        {_num(p.step_time)}s\n"
           f"   Steps per update:            {p.steps_per_update}\n"
           "\n")
    out += "".join(f"{s:>14} " for s in systems) + "\n"
    for r in rows:
        out += "".join(f"{ms:12.3f}ms " for ms in r) + "\n"
    return out


def run_profile(p: Profile, device: int = 0, max_updates: int | None = None, log=None,
                every_step: bool = False) -> list[float]:
    """Runs a profile on the HIP engine; returns the time per step (ms) of each
    update, each update synchronised before the clock stops. every_step: run the
    engine's default mode (reordering every step) whatever the profile's interval."""
    from ._lib import SPH_FLAG_PARTICLE_MAP, SPH_V_NONFINITE
    from .engine import SimulationEngineHIP

    n_upd = update_count(p)
    if max_updates is not None:
        n_upd = min(n_upd, int(max_updates))
    e = SimulationEngineHIP(device)
    try:
        if p.reorder_time_interval is not None and not every_step:
            e.Initialize(p.scene, flags=SPH_FLAG_PARTICLE_MAP)
            e.set_reorder_interval(p.reorder_time_interval)
        else:
            e.Initialize(p.scene)
        e.synchronize()
        rows = []
        for u in range(n_upd):
            t0 = time.perf_counter()
            e.Update(p.step_time, p.steps_per_update)
            e.synchronize()
            rows.append((time.perf_counter() - t0) * 1e3 / max(p.steps_per_update, 1))
            if log:
                log(f"[profile] {p.name}: update {u + 1}/{n_upd}, {rows[-1]:.3f} ms/step")
        if e.validate() & ~SPH_V_NONFINITE:  # non-finite state aside: a profile may blow up
            raise RuntimeError(f"profile '{p.name}': the engine's tables failed validation")
    finally:
        e.close()
    return rows


def run_profiles(path: str, out_dir: str = ".", device: int = 0, max_updates: int | None = None,
                 log=print, every_step: bool = False) -> list[str]:
    """Runs every profile of a file on the HIP engine and writes each report to
    its outputFilePath under out_dir; returns the paths written."""
    written = []
    for p in load_profiles(path):
        if log:
            log(f'[profile] "{p.name}": {p.scene.n} dynamic, {p.scene.n_static} static particles, '
                f"{update_count(p)} updates of {p.steps_per_update} steps at {p.step_time} s, reorder interval "
                f"{'every step' if every_step or p.reorder_time_interval is None else p.reorder_time_interval}")
        rows = run_profile(p, device=device, max_updates=max_updates, log=log, every_step=every_step)
        dst = os.path.join(out_dir, p.output_file_path)
        os.makedirs(os.path.dirname(os.path.abspath(dst)), exist_ok=True)
        with open(dst, "w", encoding="utf-8") as fh:
            fh.write(profile_report(p, ["HIP"], [[r] for r in rows]))
        written.append(dst)
    return written


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Run the reference's simulation profiles on the HIP engine.")
    ap.add_argument("file", help="a simulationProfiles JSON file")
    ap.add_argument("--out", default=".", help="directory the profiles' outputFilePath are taken relative to")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--max-updates", type=int, default=None, help="stop each profile after this many updates")
    ap.add_argument("--every-step", action="store_true",
                    help="the engine's default mode (reorder every step) instead of the profile's interval")
    args = ap.parse_args(argv)
    for path in run_profiles(args.file, out_dir=args.out, device=args.device, max_updates=args.max_updates,
                             log=lambda m: print(m, flush=True), every_step=args.every_step):
        print(f"[profile] wrote {path}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

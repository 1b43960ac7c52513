"""Scene files in the reference's JSON format, for SimulationEngineHIP.Initialize.

A scene file (SceneBlueprint::LoadScene, SPHFluidSimulation/source/SPH/Core/
SceneBlueprint.cpp:134-183) holds:

* ``particleBehaviourParameters`` (required): particleMass, gasConstant,
  elasticity, viscosity, gravity [x, y, z], restDensity, maxInteractionDistance
  (SceneBlueprint.cpp:15-33);
* ``particleSetBlueprints``: a list of {layerName, particleSetBlueprintType,
  properties}. Types "FilledBox" and "BoxShell" are known; an unknown type is
  skipped with a warning, as the reference does (SceneBlueprint.cpp:43-48). The
  engine reads the layers "static" and "dynamic"; each layer is its blueprints'
  particles in file order (SceneBlueprint.h:23-48), which fixes the original
  particle ids;
* ``boundaryMesh`` (optional): {path, objectName?}, a Wavefront OBJ whose
  triangles become the boundary mesh (SceneBlueprint.cpp:63-92);
* ``otherParameters`` (optional): string key/value pairs (reorderTimeInterval,
  openCLChoosesGroupSize, ...), kept as they are.

Generators follow FilledBoxParticleSetBlueprint.cpp:23-66 and
BoxShellParticleSetBlueprint.cpp:14-116 (see scenes.filled_box and
scenes.box_shell). The reference's random offsets come from BlazeEngine's
Random, which is not available; this loader draws them from numpy's PCG64
seeded with the blueprint's ``seed`` (default 0), so jittered layers have the
reference's distribution but not its values.
"""
from __future__ import annotations

import json
import os
import warnings

import numpy as np

from .scenes import Params, Scene, box_shell, filled_box

__all__ = ["SceneFileError", "load_scene", "parse_scene", "load_obj_triangles"]


class SceneFileError(ValueError):
    """The scene file is not valid JSON or misses a required entry (the
    reference marks such a scene invalid, SceneBlueprint.cpp:166-178)."""


def _expect(obj, key, kind=None):
    if not isinstance(obj, dict) or key not in obj:
        raise SceneFileError(f"missing entry '{key}'")
    v = obj[key]
    if kind == "vec3":
        if not (isinstance(v, list) and len(v) == 3 and all(isinstance(c, (int, float)) for c in v)):
            raise SceneFileError(f"entry '{key}' is not a 3-vector")
        return [float(c) for c in v]
    if kind == "float":
        if isinstance(v, bool) or not isinstance(v, (int, float)):
            raise SceneFileError(f"entry '{key}' is not a number")
        return float(v)
    if kind == "uint":
        if isinstance(v, bool) or not isinstance(v, (int, float)) or v < 0 or int(v) != v:
            raise SceneFileError(f"entry '{key}' is not an unsigned integer")
        return int(v)
    return v


def _params(j) -> Params:
    g = _expect(j, "gravity", "vec3")
    return Params(
        particleMass=_expect(j, "particleMass", "float"),
        gasConstant=_expect(j, "gasConstant", "float"),
        elasticity=_expect(j, "elasticity", "float"),
        viscosity=_expect(j, "viscosity", "float"),
        gravityX=g[0],
        gravityY=g[1],
        gravityZ=g[2],
        restDensity=_expect(j, "restDensity", "float"),
        maxInteractionDistance=_expect(j, "maxInteractionDistance", "float"),
    )


def _filled_box(props) -> np.ndarray:
    # FilledBoxParticleSetBlueprint::Load (FilledBoxParticleSetBlueprint.cpp:23-36):
    # particlesPerUnit and randomOffsetIntensity are unsigned integers there.
    ppu = _expect(props, "particlesPerUnit", "uint")
    intensity = _expect(props, "randomOffsetIntensity", "uint")
    offset = _expect(props, "spawnVolumeOffset", "vec3")
    size = _expect(props, "spawnVolumeSize", "vec3")
    seed = _expect(props, "seed", "uint") if "seed" in props else 0
    return filled_box(size, ppu, offset=offset, random_offset_intensity=float(intensity),
                      rng=np.random.default_rng(seed))


def _box_shell(props) -> np.ndarray:
    # BoxShellParticleSetBlueprint::Load (BoxShellParticleSetBlueprint.cpp:14-29).
    spacing = _expect(props, "particleDistance", "float")
    intensity = _expect(props, "randomOffsetIntensity", "float")
    offset = _expect(props, "spawnVolumeOffset", "vec3")
    size = _expect(props, "spawnVolumeSize", "vec3")
    seed = _expect(props, "seed", "uint") if "seed" in props else 0
    return box_shell(size, spacing, offset=offset, jitter=intensity, rng=np.random.default_rng(seed))


_GENERATORS = {"FilledBox": _filled_box, "BoxShell": _box_shell}


def load_obj_triangles(path: str, object_name: str | None = None) -> np.ndarray:
    """Triangles (T, 9) float32 of a Wavefront OBJ: the faces of object
    `object_name` (an ``o`` statement; all faces if None), in file order,
    polygons fanned from their first vertex. Indices are 1-based; negative
    indices count back from the last vertex read."""
    verts = []
    tris = []
    current = None
    with open(path, "r", encoding="utf-8", errors="replace") as fh:
        for line in fh:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "v" and len(parts) >= 4:
                verts.append((float(parts[1]), float(parts[2]), float(parts[3])))
            elif parts[0] == "o":
                current = " ".join(parts[1:])
            elif parts[0] == "f" and len(parts) >= 4:
                if object_name is not None and current != object_name:
                    continue
                idx = []
                for tok in parts[1:]:
                    k = int(tok.split("/")[0])
                    idx.append(k - 1 if k > 0 else len(verts) + k)
                for t in range(1, len(idx) - 1):
                    tris.append(verts[idx[0]] + verts[idx[t]] + verts[idx[t + 1]])
    return np.asarray(tris, dtype=np.float32).reshape(-1, 9)


def parse_scene(doc, name: str = "", asset_root: str | None = None) -> Scene:
    """A Scene from a parsed scene document (a dict) or its JSON text."""
    if isinstance(doc, (str, bytes)):
        try:
            doc = json.loads(doc)
        except json.JSONDecodeError as ex:
            raise SceneFileError(f"scene is not valid JSON: {ex}") from ex
    if not isinstance(doc, dict):
        raise SceneFileError("scene is not a JSON object")
    params = _params(_expect(doc, "particleBehaviourParameters"))
    layers: dict[str, list[np.ndarray]] = {}
    for bp in doc.get("particleSetBlueprints", []):
        layer = _expect(bp, "layerName")
        kind = _expect(bp, "particleSetBlueprintType")
        gen = _GENERATORS.get(kind)
        if gen is None:
            warnings.warn(f'No particle group registered as "{kind}" found')
            continue
        layers.setdefault(layer, []).append(gen(_expect(bp, "properties")))

    def layer(n):
        parts = layers.get(n, [])
        return np.concatenate(parts).astype(np.float32) if parts else np.zeros((0, 3), np.float32)

    pos = layer("dynamic")
    triangles = None
    if "boundaryMesh" in doc:
        mesh = doc["boundaryMesh"]
        rel = _expect(mesh, "path")
        path = rel if os.path.isabs(rel) or asset_root is None else os.path.join(asset_root, rel)
        if not os.path.isfile(path):
            # SceneBlueprint.cpp:69-79: an invalid path logs an error; the scene stays valid without a mesh.
            warnings.warn(f'Path to boundary mesh is not a valid path: "{path}"')
        else:
            triangles = load_obj_triangles(path, mesh.get("objectName"))
    other = {str(k): str(v) for k, v in doc.get("otherParameters", {}).items()}
    return Scene(params, np.ascontiguousarray(pos), np.zeros_like(pos), np.ascontiguousarray(layer("static")),
                 name=name, triangles=triangles, other=other)


def load_scene(path: str, asset_root: str | None = None) -> Scene:
    """Load a scene file. A boundary mesh path is taken relative to `asset_root`;
    by default that is the directory above the file's ``assets`` directory when
    the file lies under one (the reference app runs from its project directory,
    whose scenes live in assets/simulationScenes), else the file's directory."""
    with open(path, "r", encoding="utf-8") as fh:
        text = fh.read()
    if asset_root is None:
        d = os.path.dirname(os.path.abspath(path))
        probe = d
        while probe and os.path.basename(probe) != "assets" and os.path.dirname(probe) != probe:
            probe = os.path.dirname(probe)
        asset_root = os.path.dirname(probe) if os.path.basename(probe) == "assets" else d
    name = os.path.splitext(os.path.basename(path))[0]
    return parse_scene(text, name=name, asset_root=asset_root)

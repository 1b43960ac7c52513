"""Seeded scene generators for the SPH step engine.

The reference builds scenes from JSON blueprints (SceneBlueprint.cpp:134-183)
with two particle-set generators, FilledBox (FilledBoxParticleSetBlueprint.cpp:37-66)
and BoxShell (BoxShellParticleSetBlueprint.cpp:30-116). Their jitter comes from
BlazeEngine's Random, which is not vendored, so every generator here draws from
numpy's PCG64 with an explicit seed instead; the lattice layout, particle counts
and particle order follow the reference generators.

A scene is the data SimulationEngine::Initialize consumes: dynamic particle
positions and velocities, static (wall) positions and the
ParticleBehaviourParameters of SPHFunctions.cpp:23-41.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# ParticleBehaviourParameters field order (Particles.h:23-41).
PARAM_FIELDS = (
    "particleMass",
    "gasConstant",
    "elasticity",
    "viscosity",
    "gravityX",
    "gravityY",
    "gravityZ",
    "restDensity",
    "maxInteractionDistance",
    "selfDensity",
    "smoothingKernelConstant",
)


@dataclasses.dataclass
class Params:
    """SPH::ParticleBehaviourParameters (Particles.h:23-41)."""

    particleMass: float
    gasConstant: float
    elasticity: float
    viscosity: float
    gravityX: float
    gravityY: float
    gravityZ: float
    restDensity: float
    maxInteractionDistance: float
    selfDensity: float = 0.0
    smoothingKernelConstant: float = 0.0

    def as_array(self) -> np.ndarray:
        return np.array([getattr(self, f) for f in PARAM_FIELDS], dtype=np.float32)


@dataclasses.dataclass
class Scene:
    params: Params
    pos: np.ndarray  # (N, 3) float32, dynamic particles in original order
    vel: np.ndarray  # (N, 3) float32
    static_pos: np.ndarray  # (N_s, 3) float32, wall particles in original order
    name: str = ""
    triangles: np.ndarray | None = None  # (T, 9) float32 boundary mesh (p1, p2, p3), or None
    other: dict = dataclasses.field(default_factory=dict)  # otherParameters of a scene file

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    @property
    def n_static(self) -> int:
        return int(self.static_pos.shape[0])


def kernel_constant(h: float) -> float:
    """C = 15 / (3.1415 h^6) evaluated in float32 in the reference's order (SPHFunctions.cpp:115-118)."""
    h32 = np.float32(h)
    d = np.float32(3.1415)
    for _ in range(6):
        d = np.float32(d * h32)
    return float(np.float32(np.float32(15.0) / d))


def lattice_kernel_sum(h: float, spacing: float) -> float:
    """Sum of (h - r)^3 over an infinite cubic lattice at `spacing`, self term included."""
    n = int(math.ceil(h / spacing)) + 1
    total = 0.0
    for i in range(-n, n + 1):
        for j in range(-n, n + 1):
            for k in range(-n, n + 1):
                r = spacing * math.sqrt(i * i + j * j + k * k)
                if r < h:
                    total += (h - r) ** 3
    return total


def calibrated_mass(h: float, spacing: float, rest_density: float) -> float:
    """Particle mass that puts an interior lattice particle exactly at rest density.

    With m = rho0 * s^3 the reference kernel gives a lattice density of 1236 for
    rho0 = 1000, h = 0.1, s = 0.05; this mass instead satisfies
    rho0 = m * C * sum_j (h - r_j)^3 on the lattice (about 0.10114).
    """
    return rest_density / (kernel_constant(h) * lattice_kernel_sum(h, spacing))


def dam_break_params(h: float = 0.1, spacing: float = 0.05) -> Params:
    """Parameters of the dam-break workloads.

    gasConstant 200 is what every reference scene uses (scene.json:56,
    triangleTestScene.json:16); viscosity 0.1 is that of triangleTestScene.json:18,
    the scene the application loads; rest density 1000 and h = 0.1 as the
    workloads state; mass calibrated to the lattice (calibrated_mass).
    """
    return Params(
        particleMass=float(np.float32(calibrated_mass(h, spacing, 1000.0))),
        gasConstant=200.0,
        elasticity=0.5,
        viscosity=0.1,
        gravityX=0.0,
        gravityY=-9.81,
        gravityZ=0.0,
        restDensity=1000.0,
        maxInteractionDistance=h,
    )


def box_shell(size, spacing: float, offset=(0.0, 0.0, 0.0), jitter: float = 0.0, rng=None) -> np.ndarray:
    """Hollow box lattice in the reference's BoxShell order (BoxShellParticleSetBlueprint.cpp:30-116).

    Grid size per axis is trunc(size / spacing) + 1 computed in float32; the
    count is 2 (gx*gy + gx*gz + gy*gz - 2gx - 2gy - 2gz + 4). Faces come first
    (xy at z = 0 and z = max, then xz, then yz, interior rows only), then the
    12 edges, then the 8 corners. `jitter` is randomOffsetIntensity: each
    particle moves by U[0, 1) * jitter * spacing per axis.
    """
    size32 = np.asarray(size, dtype=np.float32)
    sp = np.float32(spacing)
    g = (size32 / sp).astype(np.float64).astype(np.int64) + 1
    gx, gy, gz = (int(v) for v in g)
    pts = []

    def add(x, y, z):
        pts.append((x, y, z))

    for x in range(1, gx - 1):
        for y in range(1, gy - 1):
            add(x, y, 0)
            add(x, y, gz - 1)
    for x in range(1, gx - 1):
        for z in range(1, gz - 1):
            add(x, 0, z)
            add(x, gy - 1, z)
    for y in range(1, gy - 1):
        for z in range(1, gz - 1):
            add(0, y, z)
            add(gx - 1, y, z)
    for x in range(1, gx - 1):
        add(x, 0, 0)
        add(x, gy - 1, 0)
        add(x, 0, gz - 1)
        add(x, gy - 1, gz - 1)
    for y in range(1, gy - 1):
        add(0, y, 0)
        add(gx - 1, y, 0)
        add(0, y, gz - 1)
        add(gx - 1, y, gz - 1)
    for z in range(1, gz - 1):
        add(0, 0, z)
        add(gx - 1, 0, z)
        add(0, gy - 1, z)
        add(gx - 1, gy - 1, z)
    for c in [(0, 0, 0), (gx - 1, 0, 0), (0, gy - 1, 0), (gx - 1, gy - 1, 0),
              (0, 0, gz - 1), (gx - 1, 0, gz - 1), (0, gy - 1, gz - 1), (gx - 1, gy - 1, gz - 1)]:
        add(*c)
    idx = np.asarray(pts, dtype=np.float32)
    expected = (gx * gy + gx * gz + gy * gz - 2 * gx - 2 * gy - 2 * gz + 4) * 2
    assert idx.shape[0] == expected, (idx.shape, expected)
    out = (idx * sp).astype(np.float32)
    if jitter:
        rng = rng or np.random.default_rng(0)
        out = out + (rng.random(out.shape, dtype=np.float64).astype(np.float32) * np.float32(jitter) * sp)
    return (out + np.asarray(offset, dtype=np.float32)).astype(np.float32)


def box_shell_count(size, spacing: float) -> int:
    """BoxShellParticleSetBlueprint::GetParticleCount (BoxShellParticleSetBlueprint.cpp:30-41)."""
    g = (np.asarray(size, dtype=np.float32) / np.float32(spacing)).astype(np.float64).astype(np.int64) + 1
    gx, gy, gz = (int(v) for v in g)
    return (gx * gy + gx * gz + gy * gz - 2 * gx - 2 * gy - 2 * gz + 4) * 2


def lattice(nx: int, ny: int, nz: int, spacing: float, offset=(0.0, 0.0, 0.0), jitter_frac: float = 0.0,
            rng=None) -> np.ndarray:
    """Jittered cubic lattice, x outermost and z innermost as FilledBox writes it
    (FilledBoxParticleSetBlueprint.cpp:56-66): particle (i, j, k) sits at
    ((i, j, k) + 0.5) * spacing + offset, moved by U(-1, 1) * jitter_frac * spacing per axis."""
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    idx = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
    p = (idx + 0.5) * spacing + np.asarray(offset, dtype=np.float64)
    if jitter_frac:
        rng = rng or np.random.default_rng(0)
        p = p + rng.uniform(-1.0, 1.0, size=p.shape) * (jitter_frac * spacing)
    return p.astype(np.float32)


def filled_box(size, particles_per_unit: int, offset=(0.0, 0.0, 0.0), random_offset_intensity: float = 0.0,
               rng=None) -> np.ndarray:
    """FilledBox generator (FilledBoxParticleSetBlueprint.cpp:37-66): lattice of
    cbrt(particlesPerUnit) per unit length, grid size trunc(size * density)."""
    lin = np.float32(float(particles_per_unit) ** (1.0 / 3.0))
    dist = np.float32(1.0) / lin
    g = (np.asarray(size, dtype=np.float32) * lin).astype(np.float64).astype(np.int64)
    i, j, k = np.meshgrid(np.arange(g[0]), np.arange(g[1]), np.arange(g[2]), indexing="ij")
    idx = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float32)
    p = idx * dist + (np.float32(dist / 2) + np.asarray(offset, dtype=np.float32))
    if random_offset_intensity:
        rng = rng or np.random.default_rng(0)
        amp = np.float32(dist / 2.0 * random_offset_intensity)
        p = p + rng.uniform(-1.0, 1.0, size=p.shape).astype(np.float32) * amp
    return p.astype(np.float32)


def dam_break(nx: int, ny: int, nz: int, box, spacing: float = 0.05, h: float = 0.1, jitter_frac: float = 0.05,
              seed: int = 0, shuffle: bool = False, vel_sigma: float = 0.0, name: str = "") -> Scene:
    """A jittered fluid block in the corner of a box-shell-walled box."""
    rng = np.random.default_rng(seed)
    pos = lattice(nx, ny, nz, spacing, jitter_frac=jitter_frac, rng=rng)
    vel = np.zeros_like(pos)
    if vel_sigma:
        vel = rng.normal(0.0, vel_sigma, size=pos.shape).astype(np.float32)
    if shuffle:
        perm = rng.permutation(pos.shape[0])
        pos = pos[perm]
        vel = vel[perm]
    walls = box_shell(box, spacing)
    return Scene(dam_break_params(h, spacing), np.ascontiguousarray(pos), np.ascontiguousarray(vel), walls,
                 name=name)


# ---- the workloads ---------------------------------------------------------

def cpu_16k(seed: int = 1) -> Scene:
    """16,000 particles: 40x20x20 lattice at 0.5h in the corner of a 4x2x2 box (16,002 walls)."""
    return dam_break(40, 20, 20, (4.0, 2.0, 2.0), seed=seed, name="dam-break-16k")


def gpu_1m(seed: int = 2) -> Scene:
    """1,048,576 particles: 128x64x128 lattice in a 12.8x6.4x12.8 box, memory order shuffled."""
    return dam_break(128, 64, 128, (12.8, 6.4, 12.8), seed=seed, shuffle=True, name="dam-break-1m")


def gpu_8m(seed: int = 3) -> Scene:
    """8,388,608 particles: 256x128x256 lattice in a 25.6x12.8x25.6 box, Gaussian
    velocities (sigma 0.1), memory order permuted."""
    return dam_break(256, 128, 256, (25.6, 12.8, 25.6), seed=seed, shuffle=True, vel_sigma=0.1,
                     name="dam-break-8m")


def batch_scenes(n_scenes: int, lo: int, hi: int, seed: int, spacing: float = 0.05, h: float = 0.1):
    """Independent dam breaks with particle counts log-uniform in [lo, hi] and a
    block aspect ratio uniform in [0.5, 2] per axis, each in a box twice the
    block's extent."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n_scenes):
        target = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        aspect = rng.uniform(0.5, 2.0, size=3)
        t = (target / float(np.prod(aspect))) ** (1.0 / 3.0)
        dims = [max(1, int(round(a * t))) for a in aspect]
        box = tuple(2.0 * d * spacing for d in dims)
        out.append(dam_break(dims[0], dims[1], dims[2], box, spacing=spacing, h=h, seed=seed * 100003 + s,
                             shuffle=True, name=f"batch-{s}"))
    return out


def batch_256(seed: int = 4):
    return batch_scenes(256, 2048, 200000, seed)


def batch_64_mixed(seed: int = 5):
    return batch_scenes(64, 65536, 1048576, seed)


def debug_smallest() -> Scene:
    """debugSmallestScene.json translated: a 1x1x1 FilledBox at 8 particles per
    unit (8 particles) inside a 2x2x2 BoxShell at spacing 0.2, with that scene's
    parameters. The shell's jitter (randomOffsetIntensity 1) uses this module's RNG."""
    pos = filled_box((1, 1, 1), 8, offset=(-0.5, -0.5, -0.5))
    walls = box_shell((2, 2, 2), 0.2, offset=(-1, -1, -1), jitter=1.0, rng=np.random.default_rng(0))
    params = Params(1.0, 200.0, 0.5, 0.5, 0.0, -9.81, 0.0, 7.0, 1.0)
    return Scene(params, pos, np.zeros_like(pos), walls, name="debugSmallestScene")


def sphere_triangles(center, radius: float, n_lat: int = 8, n_lon: int = 12) -> np.ndarray:
    """A closed UV-sphere mesh (T, 9) float32: n_lon triangles at each pole cap
    and two per quad between, outward-facing (a boundary obstacle for tests)."""
    c = np.asarray(center, dtype=np.float64)
    th = np.linspace(0.0, np.pi, n_lat + 1)
    ph = np.linspace(0.0, 2.0 * np.pi, n_lon + 1)

    def p(i, j):
        return c + radius * np.array([np.sin(th[i]) * np.cos(ph[j]), np.cos(th[i]), np.sin(th[i]) * np.sin(ph[j])])

    tris = []
    for i in range(n_lat):
        for j in range(n_lon):
            a, b, cc, d = p(i, j), p(i, j + 1), p(i + 1, j), p(i + 1, j + 1)
            if i > 0:
                tris.append(np.concatenate([a, b, cc]))
            if i < n_lat - 1:
                tris.append(np.concatenate([b, d, cc]))
    return np.asarray(tris, dtype=np.float32)


def concat(scenes):
    """Concatenate scenes for a batched initialize: returns (params list,
    dyn offsets, pos, vel, static offsets, static pos)."""
    dyn_off = np.zeros(len(scenes) + 1, dtype=np.uint64)
    st_off = np.zeros(len(scenes) + 1, dtype=np.uint64)
    for s, sc in enumerate(scenes):
        dyn_off[s + 1] = dyn_off[s] + sc.n
        st_off[s + 1] = st_off[s] + sc.n_static
    pos = np.concatenate([sc.pos for sc in scenes]).astype(np.float32)
    vel = np.concatenate([sc.vel for sc in scenes]).astype(np.float32)
    spos = (np.concatenate([sc.static_pos for sc in scenes]).astype(np.float32)
            if any(sc.n_static for sc in scenes) else np.zeros((0, 3), np.float32))
    return [sc.params for sc in scenes], dyn_off, pos, vel, st_off, spos

"""Seeded scenario generator for the benchmark configurations.

Plays the role of the reference's tools/gen_grid.py (a cubic lattice at
spacing h/2 written as a .in/.par pair, gen_grid.py:4,27-48) for the five
benchmark configurations:

  c0  32^3 dam break, box 4 x 2 x 2, 100 steps (the CPU-runnable case)
  c1  100^3 dam break, box 12 x 6 x 6, 200 steps (1,000,000 particles)
  c2  two 128 x 256 x 128 columns colliding, box 38.4 x 19.2 x 19.2
  c3  256 independent scenes, N_s log-uniform in [2048, 65536]
  c4  256 x 128 x 128 block + seeded sphere obstacle, box 24 x 8 x 8

Common: h = 0.1, lattice spacing 0.5 h, first site at min + 1.1 h, jitter
uniform in +-0.01 h from splitmix64 (seed 1000 + config index), v = 0,
m = 1, rho0 = 6378, dt = 1/120, g = -9.81.  The boxes are the stated ones
scaled by the smallest integer that holds the lattice (2, 3, 3, 2 for c0,
c1, c2, c4); c3 derives each box from its block and fill fraction.

Random numbers: splitmix64 (Steele, Lea & Flood 2014) as a counter-based
stream, value k = mix(seed + (k + 1) * 0x9E3779B97F4A7C15); uniforms use the
top 53 bits.  Jitter is computed in double and rounded to float once.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

import pbf

H = 0.1
SPACING = 0.5 * H
MARGIN = 1.1 * H
JITTER = 0.01 * H
REST_DENSITY = 6378.0
DT = 1.0 / 120.0

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Values start .. start+count-1 of the splitmix64 stream of ``seed``."""
    with np.errstate(over="ignore"):
        k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Uniform doubles in [0, 1) from the top 53 bits."""
    return (splitmix64(seed, count, start) >> np.uint64(11)).astype(np.float64) * (
        2.0 ** -53)


class _Stream:
    """Sequential draws from one splitmix64 stream."""

    def __init__(self, seed: int):
        self.seed = seed
        self.pos = 0

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        u = float(uniform(self.seed, 1, self.pos)[0])
        self.pos += 1
        return lo + (hi - lo) * u


def lattice(dims: Sequence[int], first: Sequence[float], seed: int,
            spacing: float = SPACING, jitter: float = JITTER,
            count: Optional[int] = None, stream_offset: int = 0) -> np.ndarray:
    """Jittered cubic lattice, x slowest / z fastest as gen_grid.py:27-31
    writes it.  Returns float32 (n, 3); ``count`` keeps the first n sites."""
    nx, ny, nz = (int(d) for d in dims)
    n = nx * ny * nz if count is None else int(count)
    idx = np.arange(n, dtype=np.int64)
    ix = idx // (ny * nz)
    iy = (idx // nz) % ny
    iz = idx % nz
    u = uniform(seed, 3 * n, 3 * stream_offset).reshape(n, 3)
    j = (2.0 * u - 1.0) * jitter
    pos = np.empty((n, 3), np.float64)
    pos[:, 0] = first[0] + ix * spacing + j[:, 0]
    pos[:, 1] = first[1] + iy * spacing + j[:, 1]
    pos[:, 2] = first[2] + iz * spacing + j[:, 2]
    return pos.astype(np.float32)


def particles_from_positions(pos: np.ndarray, mass: float = 1.0) -> np.ndarray:
    """(n, 3) positions -> (n, 10) particle records, v = 0."""
    out = np.zeros((pos.shape[0], 10), np.float32)
    out[:, 0] = mass
    out[:, 1:4] = pos
    return out


def make_params(box_max: Sequence[float], physics: int = pbf.PHYSICS_PAPER,
                box_min: Sequence[float] = (0.0, 0.0, 0.0), h: float = H,
                dt: float = DT, rest_density: float = REST_DENSITY,
                time_end: float = 0.0, **overrides) -> pbf.Params:
    p = pbf.default_params(physics)
    p.time_step_length = dt
    p.time_end = time_end
    p.x_min, p.y_min, p.z_min = box_min
    p.x_max, p.y_max, p.z_max = box_max
    for a, name in enumerate(("x_n", "y_n", "z_n")):
        n = round((box_max[a] - box_min[a]) / h)
        setattr(p, name, float(n))
    p.rest_density = rest_density
    for k, v in overrides.items():
        if k == "sphere":
            for a in range(4):
                p.sphere[a] = v[a]
        else:
            setattr(p, k, v)
    return p


@dataclass
class Scenario:
    name: str
    params: List[pbf.Params]
    particles: List[np.ndarray]
    steps: int
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(sum(p.shape[0] for p in self.particles))


def _dam(index: int, dims, box, physics, steps, name, **kw) -> Scenario:
    first = (MARGIN, MARGIN, MARGIN)
    pos = lattice(dims, first, seed=1000 + index)
    params = make_params(box, physics, time_end=steps * DT, **kw)
    return Scenario(name, [params], [particles_from_positions(pos)], steps,
                    {"dims": dims, "box": box})


def config(index: int, physics: int = pbf.PHYSICS_PAPER, **kw) -> Scenario:
    """Benchmark configuration ``index`` (0..4)."""
    if index == 0:
        return _dam(0, (32, 32, 32), (4.0, 2.0, 2.0), physics, 100, "c0", **kw)
    if index == 1:
        return _dam(1, (100, 100, 100), (12.0, 6.0, 6.0), physics, 200, "c1", **kw)
    if index == 2:
        dims = (128, 256, 128)
        box = (38.4, 19.2, 19.2)
        a = lattice(dims, (MARGIN, MARGIN, MARGIN), seed=1002)
        x2 = box[0] - MARGIN - (dims[0] - 1) * SPACING
        b = lattice(dims, (x2, MARGIN, MARGIN), seed=1002,
                    stream_offset=a.shape[0])
        pos = np.concatenate([a, b])
        params = make_params(box, physics, time_end=100 * DT, **kw)
        return Scenario("c2", [params], [particles_from_positions(pos)], 100,
                        {"dims": dims, "box": box})
    if index == 3:
        return batch_config(physics=physics, **kw)
    if index == 4:
        return sphere_dam((256, 128, 128), (24.0, 8.0, 8.0), seed=1004,
                          steps=50, name="c4", physics=physics, **kw)
    raise ValueError(f"unknown config {index}")


def sphere_dam(dims, box, seed: int, steps: int = 50, name: str = "sphere",
               physics: int = pbf.PHYSICS_PAPER, **kw) -> Scenario:
    """A jittered block in a box with a seeded obstacle sphere (radius
    U[0.1, 0.3] of the box height) placed outside the block grown by r + h,
    so the fluid falls onto or flows around it.  c4 is (256, 128, 128) in a
    24 x 8 x 8 box (room for the obstacle beside the 12.8 x 6.4 x 6.4
    block)."""
    pos = lattice(dims, (MARGIN, MARGIN, MARGIN), seed=seed)
    rng = _Stream(seed + 1000)
    r = rng.uniform(0.1, 0.3) * box[1]
    lo = np.full(3, MARGIN) - (r + H)
    hi = np.array([MARGIN + (d - 1) * SPACING for d in dims]) + (r + H)
    centre = None
    for _ in range(10000):
        c = np.array([rng.uniform(r + H, box[a] - r - H) for a in range(3)])
        # outside the block grown by r + h (distance to the box > r + h)
        d = np.maximum(np.maximum(lo + (r + H) - c, c - hi + (r + H)), 0.0)
        outside = np.any((c < lo) | (c > hi))
        if outside and np.linalg.norm(d) > r + H:
            centre = c
            break
    if centre is None:
        raise RuntimeError("no sphere placement found")
    sphere = (float(centre[0]), float(centre[1]), float(centre[2]), float(r))
    params = make_params(box, physics, time_end=steps * DT, sphere=sphere, **kw)
    return Scenario(name, [params], [particles_from_positions(pos)], steps,
                    {"dims": dims, "box": box, "sphere": sphere})


def batch_config(nscenes: int = 256, physics: int = pbf.PHYSICS_PAPER,
                 nmin: int = 2048, nmax: int = 65536, seed: int = 1003,
                 **kw) -> Scenario:
    """c3: independent scenes; N_s = floor(exp(U[ln nmin, ln nmax])), box
    aspect ratios U[1, 4]^3, fill fraction U[0.2, 0.5]; the block has the
    box's aspect ratio and sits at a seeded position >= 1.1 h from walls."""
    rng = _Stream(seed)
    params, parts, counts = [], [], []
    for s in range(nscenes):
        ns = int(math.floor(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        ns = min(max(ns, nmin), nmax)
        aspect = [rng.uniform(1.0, 4.0) for _ in range(3)]
        fill = rng.uniform(0.2, 0.5)
        t = (ns / (aspect[0] * aspect[1] * aspect[2])) ** (1.0 / 3.0)
        dims = [max(1, int(math.ceil(a * t))) for a in aspect]
        while dims[0] * dims[1] * dims[2] < ns:
            dims[int(np.argmin(dims))] += 1
        ext = [(d - 1) * SPACING for d in dims]
        scale = fill ** (-1.0 / 3.0)
        box = []
        for a in range(3):
            L = max((ext[a] + SPACING) * scale, ext[a] + 2 * MARGIN + H)
            box.append(math.ceil(L / H - 1e-9) * H)
        first = [rng.uniform(MARGIN, box[a] - MARGIN - ext[a]) for a in range(3)]
        pos = lattice(dims, first, seed=seed * 1000 + s, count=ns)
        params.append(make_params(tuple(box), physics, time_end=100 * DT, **kw))
        parts.append(particles_from_positions(pos))
        counts.append(ns)
    return Scenario("c3", params, parts, 100, {"counts": counts})


def small_dam(dims=(8, 8, 8), box=(1.0, 1.0, 1.0), physics=pbf.PHYSICS_REF,
              seed: int = 7, origin=(MARGIN, MARGIN, MARGIN), **kw) -> Scenario:
    """A small jittered dam block for tests (first site at ``origin``)."""
    pos = lattice(dims, origin, seed=seed)
    params = make_params(box, physics, **kw)
    return Scenario("small", [params], [particles_from_positions(pos)], 10,
                    {"dims": dims, "box": box})


def write_in(path: str, particles: np.ndarray) -> None:
    """Particle file in the reference format (PartReader.cpp:35-89)."""
    with open(path, "w") as f:
        f.write(f"{particles.shape[0]}\n")
        for q in particles:
            f.write("{} {} {} {} {} {} {}\n".format(*(repr(float(v)) for v in q[:7])))


def write_par(path: str, params: pbf.Params, part_file: str) -> None:
    """Scenario file in the reference format (ConfigReader.cpp:45-84)."""
    p = params
    lines = [
        ("part_input_file", part_file),
        ("timestep_length", repr(p.time_step_length)),
        ("time_end", repr(p.time_end)),
        ("x_min", repr(p.x_min)), ("x_max", repr(p.x_max)),
        ("y_min", repr(p.y_min)), ("y_max", repr(p.y_max)),
        ("z_min", repr(p.z_min)), ("z_max", repr(p.z_max)),
        ("x_n", repr(p.x_n)), ("y_n", repr(p.y_n)), ("z_n", repr(p.z_n)),
        ("restdensity", repr(p.rest_density)),
    ]
    with open(path, "w") as f:
        for k, v in lines:
            f.write(f"{k:<22}{v}\n")


def load(par_path: str):
    """Read a .par scenario and its particle file (main.cpp:44-53)."""
    params, part = pbf.read_par(par_path)
    parts = pbf.read_particles(os.path.join(os.path.dirname(par_path), part))
    return params, parts

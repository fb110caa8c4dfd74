"""Seeded shape generators for the benchmark / parity workloads.

Every generator is deterministic: a splitmix64 stream seeded by the caller
draws float32 parameters, the shape is built with :class:`frep.tape.Builder`
and compiled to a tape.  Workloads:

* ``spheres_union``  -- union (min) of spheres, SDF form
  ``sqrt((x-cx)^2 + (y-cy)^2 + (z-cz)^2) - r``; 8 spheres, centres U[-1,1]^3,
  radii U[0.1,0.4], seed 1, 3-level tree 8/8/4 (256^3).
* ``random_csg``     -- random binary CSG tree of primitives (sphere SDF, axis
  box ``max(max(|x-cx|-hx, |y-cy|-hy), |z-cz|-hz)`` with ``|u| = sqrt(u*u)``
  (the reference's Abs, vulkan/src/csg/lib.cpp:68-71), half-plane
  ``n.p - d``), internal nodes union = min, intersection = max,
  difference = max(a, -b) (vulkan/src/csg/lib.cpp:117-135) chosen uniformly;
  24 primitives, seed 2, tree 16/8/8 (1024^3).
* ``batch_mix``      -- many sphere unions with log-uniform primitive counts
  (4..400) at resolutions 128^3..1024^3 (branching per BATCH_RESOLUTIONS).
* ``gyroid_blend``   -- gyroid lattice with sin/cos, thickened, clipped by a
  box and blended with spheres by a polynomial smooth minimum.
"""
from __future__ import annotations

import math

import numpy as np

from .tape import Builder, Tape, compile_tape, f32

MASK64 = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014); uniform floats from the top 24 bits."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float, hi: float) -> float:
        u = (self.next_u64() >> 40) * (1.0 / (1 << 24))
        return f32(lo + (hi - lo) * u)

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi] inclusive."""
        return lo + self.next_u64() % (hi - lo + 1)


def sphere_sdf(b: Builder, c, r) -> int:
    dx = b.sub(b.x(), b.const(c[0]))
    dy = b.sub(b.y(), b.const(c[1]))
    dz = b.sub(b.z(), b.const(c[2]))
    d2 = b.add(b.add(b.square(dx), b.square(dy)), b.square(dz))
    return b.sub(b.sqrt(d2), b.const(r))


def box_sdf(b: Builder, c, h) -> int:
    ax = b.sub(b.abs(b.sub(b.x(), b.const(c[0]))), b.const(h[0]))
    ay = b.sub(b.abs(b.sub(b.y(), b.const(c[1]))), b.const(h[1]))
    az = b.sub(b.abs(b.sub(b.z(), b.const(c[2]))), b.const(h[2]))
    return b.max(b.max(ax, ay), az)


def halfplane(b: Builder, n, d) -> int:
    dot = b.add(b.add(b.mul(b.const(n[0]), b.x()), b.mul(b.const(n[1]), b.y())),
                b.mul(b.const(n[2]), b.z()))
    return b.sub(dot, b.const(d))


def spheres_union(n_spheres: int = 8, seed: int = 1, rmin: float = 0.1,
                  rmax: float = 0.4) -> Tape:
    rng = SplitMix64(seed)
    b = Builder()
    acc = None
    params = []
    for _ in range(n_spheres):
        c = (rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1))
        r = rng.uniform(rmin, rmax)
        params.append((c, r))
        s = sphere_sdf(b, c, r)
        acc = s if acc is None else b.min(acc, s)
    tape = compile_tape(b, acc)
    tape.meta = {"kind": "spheres_union", "seed": seed, "n": n_spheres, "spheres": params,
                 "builder": b, "root": acc}
    return tape


def _random_primitive(b: Builder, rng: SplitMix64):
    kind = rng.randint(0, 2)
    if kind == 0:
        c = (rng.uniform(-0.8, 0.8), rng.uniform(-0.8, 0.8), rng.uniform(-0.8, 0.8))
        return sphere_sdf(b, c, rng.uniform(0.1, 0.5))
    if kind == 1:
        c = (rng.uniform(-0.8, 0.8), rng.uniform(-0.8, 0.8), rng.uniform(-0.8, 0.8))
        h = (rng.uniform(0.1, 0.5), rng.uniform(0.1, 0.5), rng.uniform(0.1, 0.5))
        return box_sdf(b, c, h)
    v = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1)],
                 dtype=np.float32)
    nrm = float(np.sqrt(np.float32(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])))
    if nrm < 1e-3:
        v, nrm = np.array([0, 0, 1], dtype=np.float32), 1.0
    n = [f32(v[k] / np.float32(nrm)) for k in range(3)]
    return halfplane(b, n, rng.uniform(-0.5, 0.5))


def random_csg(n_prims: int = 24, seed: int = 2) -> Tape:
    rng = SplitMix64(seed)
    b = Builder()

    def build(n):
        if n == 1:
            return _random_primitive(b, rng)
        k = rng.randint(1, n - 1)
        op = rng.randint(0, 2)
        left = build(k)
        right = build(n - k)
        if op == 0:
            return b.min(left, right)
        if op == 1:
            return b.max(left, right)
        return b.max(left, b.neg(right))

    root = build(n_prims)
    tape = compile_tape(b, root)
    tape.meta = {"kind": "random_csg", "seed": seed, "n": n_prims, "builder": b, "root": root}
    return tape


def gyroid_blend(k: float = 20.0, n_spheres: int = 32, seed: int = 5,
                 thickness: float = 0.1, blend: float = 0.05, balanced: bool = False,
                 rmin: float = 0.05, rmax: float = 0.25) -> Tape:
    """|gyroid(k p)| - thickness, intersected with a box, smooth-unioned with
    spheres by the polynomial smooth minimum
    ``smin(a, b) = min(a, b) - h*h/(4r)`` with ``h = max(r - |a - b|, 0)``.
    By default the spheres are blended in one after another; with `balanced`
    they are first smooth-unioned in a balanced tree that splits them at the
    median of their widest axis (so interval pruning can drop whole far-away
    subtrees), which is then blended with the lattice."""
    rng = SplitMix64(seed)
    b = Builder()
    kx, ky, kz = (b.mul(b.const(k), b.x()), b.mul(b.const(k), b.y()),
                  b.mul(b.const(k), b.z()))
    g = b.add(b.add(b.mul(b.sin(kx), b.cos(ky)), b.mul(b.sin(ky), b.cos(kz))),
              b.mul(b.sin(kz), b.cos(kx)))
    # scale by 1/k so the field is roughly a distance
    g = b.sub(b.abs(b.div(g, b.const(k))), b.const(thickness / k))
    shape = b.max(g, box_sdf(b, (0.0, 0.0, 0.0), (0.8, 0.8, 0.8)))
    inv4r = b.const(1.0 / (4.0 * blend))
    rr = b.const(blend)

    def smin(p, q):
        h = b.max(b.sub(rr, b.abs(b.sub(p, q))), b.const(0.0))
        return b.sub(b.min(p, q), b.mul(b.mul(h, h), inv4r))

    params = []
    for _ in range(n_spheres):
        c = (rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1))
        params.append((c, rng.uniform(rmin, rmax)))

    def kd(items):
        # smooth union of a set of spheres, split at the median of the axis
        # of widest spread, so that each subtree covers one region of space
        if len(items) == 1:
            c, r = items[0]
            return sphere_sdf(b, c, r)
        ax = max(range(3), key=lambda a: max(p[0][a] for p in items) - min(p[0][a] for p in items))
        items = sorted(items, key=lambda p: p[0][ax])
        h = len(items) // 2
        return smin(kd(items[:h]), kd(items[h:]))

    if balanced and params:
        shape = smin(shape, kd(params))
    else:
        for c, r in params:
            shape = smin(shape, sphere_sdf(b, c, r))
    tape = compile_tape(b, shape)
    tape.meta = {"kind": "gyroid_blend", "seed": seed, "builder": b, "root": shape}
    return tape


def make_grid(dims, low=(-1.0, -1.0, -1.0), world_size=2.0, tape_time=0.0) -> dict:
    return {"dims": list(dims), "low": [f32(v) for v in low], "world_size": f32(world_size),
            "tape_time": f32(tape_time)}


# Resolutions of the batched mix: effective resolution -> branching per level.
BATCH_RESOLUTIONS = {128: [8, 4, 4], 256: [8, 8, 4], 512: [8, 8, 8], 1024: [16, 8, 8]}


def batch_mix(n_shapes: int = 64, seed_base: int = 1000, min_prims: int = 4,
              max_prims: int = 400, resolutions=(128, 256, 512, 1024)):
    """The batched workload: shape k is a union of spheres (as spheres_union,
    seed seed_base + k) with round(exp(U[ln min_prims, ln max_prims]))
    spheres, voxelized at a resolution drawn uniformly from `resolutions`.
    Returns a list of (tape, grid, resolution)."""
    out = []
    for k in range(n_shapes):
        rng = SplitMix64((seed_base + k) ^ 0x5EED)
        lo, hi = math.log(min_prims), math.log(max_prims)
        count = int(round(math.exp(lo + (hi - lo) * rng.uniform(0.0, 1.0))))
        count = max(min_prims, min(max_prims, count))
        res = resolutions[rng.randint(0, len(resolutions) - 1)]
        tape = spheres_union(count, seed_base + k)
        out.append((tape, make_grid(BATCH_RESOLUTIONS[res]), res))
    return out


# Named workloads: shape factory, grid, and the frame rendered of it.
WORKLOADS = {
    # 8 spheres, 256^3 (8/8/4), one 512x512 orthographic frame looking down -z
    "spheres8_256": (lambda: spheres_union(8, 1), make_grid([8, 8, 4]),
                     dict(pos=(0.0, 0.0, 3.0), fwd=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0),
                          width=512, height=512, ortho=True, extent=(2.0, 2.0), spp=1,
                          jitter=False)),
    # 24-primitive random CSG, 1024^3 (16/8/8), 1920x1080 perspective
    "csg24_1024": (lambda: random_csg(24, 2), make_grid([16, 8, 8]),
                   dict(pos=(1.8, 1.2, 2.4), fwd=(-0.55, -0.35, -0.75), up=(0.0, 1.0, 0.0),
                        width=1920, height=1080, ortho=False, spp=1, jitter=False)),
    # gyroid lattice blended with 416 spheres in a balanced smooth union
    # (9607-op tape), 4096^3 (16/16/16), 1920x1080 at 4 jittered samples per
    # pixel
    "gyroid_4096": (lambda: gyroid_blend(n_spheres=416, balanced=True), make_grid([16, 16, 16]),
                    dict(pos=(1.8, 1.2, 2.4), fwd=(-0.55, -0.35, -0.75), up=(0.0, 1.0, 0.0),
                         width=1920, height=1080, ortho=False, spp=4, jitter=True)),
}


def frame(name: str) -> dict:
    """The frame rendered of a named workload (camera and sampling)."""
    return dict(WORKLOADS[name][2])


def workload(name: str):
    fn, grid, _ = WORKLOADS[name]
    return fn(), dict(grid)


def camera_basis(pos, forward, initial_up, fov_deg, width, height, ortho=False,
                 extent=(2.0, 2.0)):
    """Camera fields as the reference computes them (camera.cpp:16-27,
    raytracer.cpp:177-181), in float32 with the same operation order as the
    oracle's and the library's helpers."""
    f = [f32(v) for v in forward]
    u0 = [f32(v) for v in initial_up]
    F = np.float32
    cr = [F(f[1]) * F(u0[2]) - F(f[2]) * F(u0[1]), F(f[2]) * F(u0[0]) - F(f[0]) * F(u0[2]),
          F(f[0]) * F(u0[1]) - F(f[1]) * F(u0[0])]
    ln = np.sqrt(F(cr[0] * cr[0] + cr[1] * cr[1]) + F(cr[2] * cr[2]))
    r = [F(c / ln) for c in cr]
    up = [r[1] * F(f[2]) - r[2] * F(f[1]), r[2] * F(f[0]) - r[0] * F(f[2]),
          r[0] * F(f[1]) - r[1] * F(f[0])]
    if ortho:
        sw = (f32(extent[0]), f32(extent[1]))
    else:
        rad = F(fov_deg) * F(math.pi / 180.0)
        swx = F(2.0) * F(math.tan(float(rad / F(2.0))))
        sw = (float(swx), float(swx * F(F(height) / F(width))))
    return {"pos": [f32(v) for v in pos], "forward": f, "right": [float(v) for v in r],
            "up": [float(v) for v in up], "screen_world": list(sw)}


def orbit_basis(azimuth: float, elevation: float, radius: float = 3.0, width: int = 960,
                height: int = 540, fov_deg: float = 45.0):
    """Camera on a sphere of `radius` around the origin looking at it
    (angles in radians, y up)."""
    ce, se = math.cos(elevation), math.sin(elevation)
    pos = (radius * ce * math.cos(azimuth), radius * se, radius * ce * math.sin(azimuth))
    fwd = tuple(-c / radius for c in pos)
    return camera_basis(pos, fwd, (0.0, 1.0, 0.0), fov_deg, width, height)


def served_requests(n: int, seed: int = 7, rate: float = 100.0, p_rerender: float = 0.5,
                    width: int = 960, height: int = 540):
    """The served workload: Poisson arrivals at `rate` req/s (exponential
    gaps).  Each request renders a random orbit view (radius 3, azimuth
    U[0, 2pi), elevation U[-60, 60] degrees) of a scene; with probability
    p_rerender the scene is one requested before (a resident tree), else a new
    24-primitive random CSG (as random_csg, seed 10000 + index) at 1024^3.
    Returns a list of (arrival_s, scene_index, azimuth, elevation, seed);
    scene_index k names shape random_csg(24, 10000 + k)."""
    rng = SplitMix64(seed)
    out, t, n_scenes = [], 0.0, 0
    for i in range(n):
        u = max(rng.uniform(0.0, 1.0), 1e-7)
        t += -math.log(u) / rate
        if n_scenes and rng.uniform(0.0, 1.0) < p_rerender:
            k = rng.randint(0, n_scenes - 1)
        else:
            k = n_scenes
            n_scenes += 1
        az = rng.uniform(0.0, 2.0 * math.pi)
        el = rng.uniform(-math.pi / 3.0, math.pi / 3.0)
        out.append((t, k, az, el, 1000 + i))
    return out

"""Synthetic workloads: the plain small scene, the large single scene, the string-mod
accumulation, and the batched uneven scene set. All randomness is SplitMix64 with the draw
order written out below, so the same seed gives the same scene everywhere.
"""
import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from . import abi
from . import geo as G
from .tracer import DirectionalLight, Material, Object, PointLight, SpotLight

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): state += 0x9E3779B97F4A7C15, then two
    xor-shift-multiply rounds."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def log_uniform(self, lo: float, hi: float) -> float:
        return math.exp(self.uniform(math.log(lo), math.log(hi)))

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi] (inclusive)."""
        return lo + int(self.uniform() * (hi - lo + 1)) if hi > lo else lo


@dataclass
class Scene:
    objects: List[Object]
    lights: List[object]
    max_bounce: int = 5
    cutoff: Tuple[float, float, float, float] = (0.001, 0.001, 0.001, 0.001)
    canvas: Tuple[float, float, float, float] = (-1.0, 1.0, -1.0, 1.0)  # left right bottom top
    width: int = 1024
    height: int = 1024
    name: str = ""

    def n_rays(self) -> int:
        return sum(int(l.num_rays) for l in self.lights)

    def params(self, fp64: bool = False) -> abi.LgParams:
        p = abi.LgParams()
        p.max_bounce = int(self.max_bounce)
        p.flags = abi.TRACE_FP64 if fp64 else 0
        p.cutoff_color[:] = list(self.cutoff)
        p.canvas[:] = list(self.canvas)
        return p

    def view(self) -> abi.LgView:
        c = self.canvas
        return abi.LgView(c[0], c[1], c[2], c[3], self.width, self.height)


def arc_bezier(centre, r, theta0, theta1) -> G.CubicBezier:
    """Cubic Bezier approximation of a circular arc with the standard handle length
    k = 4/3 tan((theta1 - theta0) / 4)."""
    k = 4.0 / 3.0 * math.tan((theta1 - theta0) / 4.0)
    cx, cy = centre
    p0 = (cx + r * math.cos(theta0), cy + r * math.sin(theta0))
    p3 = (cx + r * math.cos(theta1), cy + r * math.sin(theta1))
    p1 = (p0[0] - k * r * math.sin(theta0), p0[1] + k * r * math.cos(theta0))
    p2 = (p3[0] + k * r * math.sin(theta1), p3[1] - k * r * math.cos(theta1))
    return G.CubicBezier([p0, p1, p2, p3])


def plain_small(num_rays: int = 20000) -> Scene:
    """World [-1,1]^2, white point light at the origin, circle, rect, flat mirror and the
    60-degree curved mirror (a Bezier, as the reference's CurvedMirror is); 5 bounces;
    1024x1024."""
    circle = Object.new_circle((0.5, 0.0), 0.2)
    circle.material_opt = Material(1.5)
    rect = Object.new_rect((-0.5, 0.3), 0.3, 0.2)
    rect.material_opt = Material(1.33)
    mirror = Object.new_mirror((-0.6, -0.6), (0.6, -0.5))
    curved = Object.new_curved_mirror(arc_bezier((0.0, 0.2), 0.8, math.radians(-30.0), math.radians(30.0)))
    light = PointLight((0.0, 0.0), num_rays, [0.01, 0.01, 0.01, 0.01])
    return Scene([circle, rect, mirror, curved], [light], max_bounce=5, width=1024, height=1024,
                 name="plain_small")


def random_objects(rng: SplitMix64, n: int, first_type: int = 0) -> List[Object]:
    """n objects cycling circle, rect, lens, mirror; centres uniform in [-0.9, 0.9]^2."""
    objs = []
    for i in range(n):
        t = (first_type + i) % 4
        cx, cy = rng.uniform(-0.9, 0.9), rng.uniform(-0.9, 0.9)
        if t == 0:
            o = Object.new_circle((cx, cy), rng.uniform(0.02, 0.15))
        elif t == 1:
            w, h = rng.uniform(0.02, 0.2), rng.uniform(0.02, 0.2)
            o = Object(G.Rect((cx, cy), G.rotation(rng.uniform(0.0, math.pi)), w, h), Material())
        elif t == 2:
            R = rng.uniform(0.03, 0.15)
            dist = rng.uniform(0.5 * R, 1.6 * R)
            lens = G.lens((cx, cy), R, dist)
            lens.rotation = G.rotation(rng.uniform(0.0, math.pi))
            o = Object(lens, Material())
        else:
            length, ang = rng.uniform(0.05, 0.3), rng.uniform(0.0, math.pi)
            dx, dy = 0.5 * length * math.cos(ang), 0.5 * length * math.sin(ang)
            o = Object.new_mirror((cx - dx, cy - dy), (cx + dx, cy + dy))
        if o.material_opt is not None:
            o.material_opt = Material(rng.uniform(1.3, 1.8))
        objs.append(o)
    return objs


def random_lights(rng: SplitMix64, counts: List[int], kinds: List[int]) -> List[object]:
    """kinds: 0 point, 1 spot (30 degree cone), 2 directional (band length 0.4). Each light
    draws its placement, then a hue u ~ U(0,1)^3 scaled so its largest channel is 0.002
    (so no light starts below the 0.001 cutoff); alpha 0.002."""
    lights = []
    for n, k in zip(counts, kinds):
        px, py = rng.uniform(-0.9, 0.9), rng.uniform(-0.9, 0.9)
        if k == 0:
            extra = None
        elif k == 1:
            a = rng.uniform(0.0, 2.0 * math.pi)
            extra = (math.cos(a), math.sin(a))
        else:
            a = rng.uniform(0.0, math.pi)
            extra = (0.2 * math.cos(a), 0.2 * math.sin(a))
        u = [rng.uniform(), rng.uniform(), rng.uniform()]
        m = max(u) if max(u) > 0.0 else 1.0
        col = [0.002 * u[0] / m, 0.002 * u[1] / m, 0.002 * u[2] / m, 0.002]
        if k == 0:
            lights.append(PointLight((px, py), n, col))
        elif k == 1:
            lights.append(SpotLight((px, py), math.pi / 6.0, extra, n, col))
        else:
            seg = G.LineSegment((px - extra[0], py - extra[1]), (px + extra[0], py + extra[1]))
            lights.append(DirectionalLight(col, n, seg))
    return lights


def large_single(num_rays: int = 1_000_000, seed: int = 2, width: int = 2048) -> Scene:
    """1M rays from a point, a 30-degree spot and a 0.4-wide directional light with seeded
    colours; 64 objects (16 each of circle, rect, lens, mirror) in [-0.9, 0.9]^2; n in
    [1.3, 1.8]; 10 bounces; 2048x2048."""
    rng = SplitMix64(seed)
    objs = random_objects(rng, 64)
    a = num_rays // 3
    counts = [num_rays - 2 * a, a, a]
    lights = random_lights(rng, counts, [0, 1, 2])
    return Scene(objs, lights, max_bounce=10, width=width, height=width, name="large_single")


def palette(seed: int = 3, n: int = 18) -> np.ndarray:
    rng = SplitMix64(seed)
    return np.array([[rng.uniform(), rng.uniform(), rng.uniform()] for _ in range(n)], np.float64)


@dataclass
class StringModSpec:
    modulo: int = 100_000
    mults: Tuple[int, ...] = tuple(range(2, 18))
    scale: float = 0.9
    alpha: float = 0.002
    width: int = 4096
    height: int = 4096
    seed: int = 3

    def colors(self) -> List[List[float]]:
        pal = palette(self.seed)
        return [[float(pal[k][0] * 0.002), float(pal[k][1] * 0.002), float(pal[k][2] * 0.002),
                 self.alpha] for k in self.mults]


def string_mod_segments(spec: StringModSpec):
    """Segments of the string-mod accumulation: for each k, the Mul-mode StringMod with
    modulo N and num k on the circle of radius `scale` (string_mod.rs:45-54, 104-122).
    Returns pos (S, 4) float64, color (S, 4) float32 (both vertices share the colour)."""
    from .string_mod import StringMod
    pos, col = [], []
    for k, c in zip(spec.mults, spec.colors()):
        sm = StringMod(modulo=spec.modulo, num=k, color=c, scale=spec.scale)
        p, c0, _ = sm.segments()
        pos.append(p)
        col.append(c0)
    return np.concatenate(pos), np.concatenate(col)


def batched_uneven(n_scenes: int = 256, seed: int = 4, ray_scale: float = 1.0) -> List[Scene]:
    """Scenes with rays log-uniform in [1e3, 2e6] over 1-3 lights, objects log-uniform in
    [1, 256] drawn as the large scene's, max bounce uniform in {3..12}, image side in
    {512, 1024, 2048}. ray_scale shrinks every ray count (for tests)."""
    rng = SplitMix64(seed)
    scenes = []
    for i in range(n_scenes):
        n_rays = int(round(rng.log_uniform(1e3, 2e6) * ray_scale))
        n_rays = max(n_rays, 3)
        n_obj = int(round(rng.log_uniform(1.0, 256.0)))
        max_b = rng.randint(3, 12)
        side = [512, 1024, 2048][rng.randint(0, 2)]
        n_l = rng.randint(1, 3)
        objs = random_objects(rng, n_obj, first_type=i % 4)
        base = n_rays // n_l
        counts = [n_rays - base * (n_l - 1)] + [base] * (n_l - 1)
        kinds = [rng.randint(0, 2) for _ in range(n_l)]
        lights = random_lights(rng, counts, kinds)
        scenes.append(Scene(objs, lights, max_bounce=max_b, width=side, height=side,
                            name="batched_%d" % i))
    return scenes


def edit_object(obj: Object, rng: SplitMix64) -> Object:
    """A serving edit: the object moved by up to 0.05 in x and y, or resized by a factor in
    [0.8, 1.2] about its own centre (even draw: move, odd: resize)."""
    if rng.randint(0, 1) == 0:
        geo = obj.geometry.translated(rng.uniform(-0.05, 0.05), rng.uniform(-0.05, 0.05))
    else:
        geo = obj.geometry.scaled(rng.uniform(0.8, 1.2))
    return Object(geo, obj.material_opt)

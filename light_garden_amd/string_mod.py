"""StringMod: modular string art (src/light_garden/string_mod.rs:3-188), the second
producer of segments for the accumulation.

draw(ctx, mods, segs) generates every StringMod's segments on the GPU (lg_string_mod_draw:
all curves and modes, modulo colours, nested StringMods over line crossings);
crossings(ctx, sm) is line_crossings_as_points (lg_string_mod_crossings). The numpy
`segments()` below is the host form of the Circle curve, kept to build inputs for the
upload path (lg_segments_upload) and as a second restatement in tests.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import abi

MODE_ADD, MODE_MUL, MODE_POW, MODE_BASE = 0, 1, 2, 3
CURVE_CIRCLE, CURVE_COMPLEX_EXP, CURVE_HYPOTROCHOID, CURVE_LISSAJOUS = 0, 1, 2, 3


@dataclass
class ModRemColor:
    """string_mod.rs:167-172: points with ix % modulo == rem take this colour."""
    modulo: int
    rem: int
    color: Sequence[float]


def _wrapping_pow(base: np.ndarray, exp: int) -> np.ndarray:
    """u64::pow with wrap-around (release-build Rust) for an array of bases."""
    r = np.ones_like(base, dtype=np.uint64)
    b = base.astype(np.uint64)
    e = int(exp)
    with np.errstate(over="ignore"):
        while e:
            if e & 1:
                r = r * b
            b = b * b
            e >>= 1
    return r


def _wrapping_pow_exps(base: int, exps: np.ndarray) -> np.ndarray:
    """u64::pow(base, e) with wrap-around for an array of exponents."""
    out = np.empty(exps.shape, np.uint64)
    b = np.uint64(base)
    with np.errstate(over="ignore"):
        for i, e in enumerate(exps.tolist()):
            r, bb, ee = np.uint64(1), b, int(e)
            while ee:
                if ee & 1:
                    r = r * bb
                bb = bb * bb
                ee >>= 1
            out[i] = r
    return out


@dataclass
class StringMod:
    modulo: int = 5
    num: int = 1
    color: Sequence[float] = (1.0, 1.0, 1.0, 1.0)
    turns: int = 1
    mode: int = MODE_MUL
    modulo_colors: List[ModRemColor] = field(default_factory=list)
    scale: float = 1.0  # the reference's curves have radius 1; this build adds a radius
    curve: int = CURVE_CIRCLE
    curve_u: Tuple[int, int, int] = (0, 0, 0)  # Hypotrochoid r, s, d; Lissajous a, b
    curve_f: Tuple[float, float] = (0.0, 0.0)  # ComplexExp c = re + i im; Lissajous delta
    nested: Optional["StringMod"] = None       # Option<Box<StringMod>>

    @staticmethod
    def complex_exp(c: complex, **kw) -> "StringMod":
        return StringMod(curve=CURVE_COMPLEX_EXP, curve_f=(c.real, c.imag), **kw)

    @staticmethod
    def hypotrochoid(r: int, s: int, d: int, **kw) -> "StringMod":
        return StringMod(curve=CURVE_HYPOTROCHOID, curve_u=(r, s, d), **kw)

    @staticmethod
    def lissajous(a: int, b: int, delta: float, **kw) -> "StringMod":
        return StringMod(curve=CURVE_LISSAJOUS, curve_u=(a, b, 0), curve_f=(delta, 0.0), **kw)

    def to_abi(self):
        """(lg_string_mod, keep-alive list): the struct points into arrays held by the list."""
        st = abi.LgStringMod()
        keep = []
        st.modulo, st.num, st.turns = int(self.modulo), int(self.num), int(self.turns)
        st.mode, st.curve = int(self.mode), int(self.curve)
        st.curve_u[:] = [int(v) for v in self.curve_u]
        st.curve_f[:] = [float(v) for v in self.curve_f]
        st.scale = float(self.scale)
        st.color[:] = [float(c) for c in self.color]
        if self.modulo_colors:
            arr = (abi.LgModRemColor * len(self.modulo_colors))()
            for i, m in enumerate(self.modulo_colors):
                arr[i].modulo, arr[i].rem = int(m.modulo), int(m.rem)
                arr[i].color[:] = [float(c) for c in m.color]
            st.modulo_colors = C.cast(arr, C.POINTER(abi.LgModRemColor))
            keep.append(arr)
        st.n_modulo_colors = len(self.modulo_colors)
        if self.nested is not None:
            inner, inner_keep = self.nested.to_abi()
            inner.nested = None  # draw() nests one level (string_mod.rs:152-158)
            st.nested = C.pointer(inner)
            keep += [inner, inner_keep]
        return st, keep

    def init_points(self) -> np.ndarray:
        """Curve::Circle points (string_mod.rs:45-54): angle = (turns * n) * TAU / m."""
        m = self.modulo
        n = np.arange(m, dtype=np.uint64)
        ang = (self.turns * n).astype(np.float64) * (2.0 * np.pi) / float(m)
        return self.scale * np.stack([np.cos(ang), np.sin(ang)], axis=1)

    def targets(self) -> np.ndarray:
        """ix for every iix (string_mod.rs:111-117)."""
        m = np.uint64(self.modulo)
        iix = np.arange(self.modulo, dtype=np.uint64)
        with np.errstate(over="ignore"):
            if self.mode == MODE_ADD:
                return (iix + np.uint64(self.num)) % m
            if self.mode == MODE_MUL:
                return (iix * np.uint64(self.num)) % m
            # `self.num as u32` (Pow) and `iix as u32` (Base) truncate before u64::pow
            if self.mode == MODE_POW:
                return _wrapping_pow(iix, int(self.num) & 0xFFFFFFFF) % m
            return _wrapping_pow_exps(self.num, iix & np.uint64(0xFFFFFFFF)) % m

    def get_colors(self, ix: np.ndarray) -> np.ndarray:
        """StringMod::get_color (string_mod.rs:124-150) for an index array."""
        base = np.asarray(self.color, np.float32)
        if not self.modulo_colors:
            return np.broadcast_to(base, (ix.shape[0], 4)).copy()
        acc = np.zeros((ix.shape[0], 4), np.float32)
        cnt = np.zeros(ix.shape[0], np.int64)
        for mrc in self.modulo_colors:
            hit = (ix % np.uint64(mrc.modulo)) == np.uint64(mrc.rem)
            acc[hit] += np.asarray(mrc.color, np.float32)
            cnt[hit] += 1
        out = np.where(cnt[:, None] > 0, acc / np.maximum(cnt, 1)[:, None].astype(np.float32), base)
        return out.astype(np.float32)

    def segments(self):
        """draw_init_points (string_mod.rs:104-122): segment iix -> ix. Returns pos (m, 4)
        float64 and per-vertex colours color0, color1 (m, 4) float32."""
        pts = self.init_points()
        iix = np.arange(self.modulo, dtype=np.uint64)
        ix = self.targets()
        pos = np.concatenate([pts[iix.astype(np.int64)], pts[ix.astype(np.int64)]], axis=1)
        return pos, self.get_colors(iix), self.get_colors(ix)


def draw(ctx, mods: Sequence[StringMod], segs) -> int:
    """LightGarden::draw in Mode::StringMod (mod.rs:681-689) on the GPU: every StringMod's
    draw() into `segs` (a Segments created with vertex_colors=True). Returns the count."""
    arr = (abi.LgStringMod * max(1, len(mods)))()
    keep = []
    for i, m in enumerate(mods):
        st, k = m.to_abi()
        arr[i] = st
        keep.append(k)
    abi.call("lg_string_mod_draw", ctx.handle, len(mods), arr, segs.handle)
    return segs.count()


def crossings(ctx, sm: StringMod) -> np.ndarray:
    """StringMod::line_crossings_as_points (string_mod.rs:87-102) on the GPU: (P, 2) float64."""
    st, _keep = sm.to_abi()
    n = C.c_uint64(0)
    rc = abi.load().lg_string_mod_crossings(ctx.handle, C.byref(st), None, 0, C.byref(n))
    if rc not in (abi.LG_OK, abi.LG_E_CAPACITY):
        abi.check(rc, "lg_string_mod_crossings")
    out = np.empty((max(n.value, 1), 2), np.float64)
    abi.call("lg_string_mod_crossings", ctx.handle, C.byref(st), out.ctypes.data, n.value, C.byref(n))
    return out[:n.value]

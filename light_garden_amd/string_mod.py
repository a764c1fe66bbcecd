"""StringMod: modular string art (src/light_garden/string_mod.rs:3-188), the second
producer of segments for the accumulation.

Segments are generated on the host in FP64 like the reference, then uploaded as FP32
(the reference's vertex conversion, sub_render_pass.rs:188-196).
"""
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

MODE_ADD, MODE_MUL, MODE_POW, MODE_BASE = 0, 1, 2, 3


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
    scale: float = 1.0  # Curve::Circle is the unit circle; this build adds a radius

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
            if self.mode == MODE_POW:
                return _wrapping_pow(iix, self.num) % m
            return _wrapping_pow_exps(self.num, iix) % m

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

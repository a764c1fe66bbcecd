"""Pins the oracle's accumulation rule, FP16 store and string-mod producer. CPU only."""
import math
from fractions import Fraction

import numpy as np
import pytest

from light_garden_amd import abi, configs
from light_garden_amd.string_mod import MODE_ADD, MODE_BASE, MODE_MUL, MODE_POW, ModRemColor, StringMod


def view(w, h, b=(-1.0, 1.0, -1.0, 1.0)):
    return abi.LgView(b[0], b[1], b[2], b[3], w, h)


def rule_exact(seg, col0, col1, W, H, bounds):
    """The sampling rule evaluated with exact rationals on small images."""
    left, right, bottom, top = [Fraction(float(x)) for x in bounds]
    x0, y0, x1, y1 = [Fraction(float(v)) for v in seg]
    sx, sy = Fraction(W) / (right - left), Fraction(H) / (top - bottom)
    X0, Y0, X1, Y1 = (x0 - left) * sx, (top - y0) * sy, (x1 - left) * sx, (top - y1) * sy
    out = np.zeros((H, W, 4))
    dX, dY = X1 - X0, Y1 - Y0
    xm = abs(dX) >= abs(dY)
    M0, M1, m0, dM, dm = (X0, X1, Y0, dX, dY) if xm else (Y0, Y1, X0, dY, dX)
    if dM == 0:
        return out
    lim = W if xm else H
    for k in range(lim):
        c = Fraction(2 * k + 1, 2)
        inside = (M0 <= c < M1) if dM > 0 else (M1 < c <= M0)
        if not inside:
            continue
        s = (c - M0) / dM
        m = m0 + s * dm
        v = m - Fraction(1, 2)
        j = math.floor(v)
        f = v - j
        a, b = [Fraction(float(x)) for x in col0], [Fraction(float(x)) for x in col1]
        colr = [float(a[i] + s * (b[i] - a[i])) for i in range(4)]
        for jj, w in ((j, 1 - f), (j + 1, f)):
            px, py = (k, jj) if xm else (jj, k)
            if 0 <= px < W and 0 <= py < H:
                wf = float(w)
                out[py, px, :3] += [wf * colr[0], wf * colr[1], wf * colr[2]]
                out[py, px, 3] += wf * colr[3] * colr[3]
    return out


def test_rule_matches_exact_rationals(orc):
    rng = np.random.default_rng(8)
    W, H, b = 23, 17, (-1.0, 1.3, -0.6, 0.9)
    for _ in range(60):
        seg = rng.uniform(-1.6, 1.6, 4)
        if rng.random() < 0.2:  # axis-aligned and pixel-centre cases
            seg[3] = seg[1]
        c0 = rng.uniform(0, 1, 4).astype(np.float32)
        c1 = rng.uniform(0, 1, 4).astype(np.float32)
        ref = rule_exact(seg, c0, c1, W, H, b)
        got = orc.accumulate_f64(seg[None, :], c0[None, :], c1[None, :], view(W, H, b))
        np.testing.assert_allclose(got, ref, atol=1e-12)


def test_horizontal_known_answer(orc):
    # 8x8 image of [-1,1]^2: the segment y = -0.125 (pixel row centre 4.5 -> v = 4.0,
    # f = 0) from pixel centre 1.5 to 5.5: columns 1..4, all weight on row 4
    W = H = 8
    seg = np.array([[-1 + 1.5 / 4, -0.125, -1 + 5.5 / 4, -0.125]])
    col = np.array([[1.0, 0.5, 0.25, 0.5]], np.float32)
    acc = orc.accumulate_f64(seg, col, None, view(W, H))
    assert np.array_equal(np.nonzero(acc[..., 0])[1], np.array([1, 2, 3, 4]))
    assert np.all(np.nonzero(acc[..., 0])[0] == 4)
    assert acc[4, 1:5, 0] == pytest.approx([1.0] * 4)
    assert acc[4, 1:5, 3] == pytest.approx([0.25] * 4)  # alpha deposits w * a * a
    # a quarter-pixel shift splits 3/4 : 1/4 between rows 4 and 5
    seg2 = seg.copy()
    seg2[0, 1] = seg2[0, 3] = -0.125 - 0.25 / 4
    acc = orc.accumulate_f64(seg2, col, None, view(W, H))
    assert acc[4, 1:5, 0] == pytest.approx([0.75] * 4)
    assert acc[5, 1:5, 0] == pytest.approx([0.25] * 4)
    # reversed direction covers the same columns (half-open span from the start point)
    rev = seg[:, [2, 3, 0, 1]]
    acc_r = orc.accumulate_f64(rev, col, None, view(W, H))
    assert np.array_equal(np.nonzero(acc_r[..., 0])[1], np.array([2, 3, 4, 5]))


def test_zero_length_and_conservation(orc):
    W = H = 64
    acc = orc.accumulate_f64(np.array([[0.1, 0.1, 0.1, 0.1]]), np.ones((1, 4), np.float32), None, view(W, H))
    assert not acc.any()
    rng = np.random.default_rng(2)
    segs = rng.uniform(-0.9, 0.9, (500, 4))
    col = np.tile(np.array([[0.3, 0.2, 0.1, 1.0]], np.float32), (500, 1))
    acc = orc.accumulate_f64(segs, col, None, view(W, H))
    n_samples = orc.sample_count(segs, view(W, H))
    # fully inside the image: every sample deposits weight 1 in total
    assert acc[..., 0].sum() == pytest.approx(float(np.float32(0.3)) * n_samples, rel=1e-12)
    assert acc[..., 3].sum() == pytest.approx(1.0 * n_samples, rel=1e-12)


def test_thread_invariance_and_f32(orc):
    rng = np.random.default_rng(4)
    segs = rng.uniform(-1.2, 1.2, (3000, 4)).astype(np.float32).astype(np.float64)
    c0 = rng.uniform(0, 0.01, (3000, 4)).astype(np.float32)
    v = view(200, 150)
    a1 = orc.accumulate_f64(segs, c0, None, v, threads=1)
    a5 = orc.accumulate_f64(segs, c0, None, v, threads=5)
    assert np.array_equal(a1, a5)
    f1 = orc.accumulate_f32(segs.astype(np.float32), c0, None, v, threads=3)
    np.testing.assert_allclose(f1, a1, rtol=1e-4, atol=1e-9)


def test_half_conversion_matches_numpy(orc):
    rng = np.random.default_rng(9)
    vals = np.concatenate([
        rng.uniform(-70000, 70000, 2000), rng.uniform(-1, 1, 2000) * 1e-5, rng.uniform(0, 1, 2000),
        np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e9, 2.0 ** -24, 2.0 ** -25, 1.5 * 2.0 ** -25,
                  2.0 ** -14, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, np.inf, -np.inf]),
    ])
    for x in vals:
        got = orc.half_bits(x)
        want = int(np.array(x, np.float64).astype(np.float16).view(np.uint16))
        assert got == want, (x, hex(got), hex(want))
        if np.isfinite(np.float16(x)):
            assert orc.half_value(got) == float(np.float16(x))


def test_resolve(orc):
    acc = np.zeros((2, 3, 4))
    acc[0, 0] = [0.5, 1e-6, 70000.0, 0.25]
    out = orc.resolve_f64(acc)
    dec = orc.decode_half(out)
    assert dec[0, 0, 0] == 0.5 and dec[0, 0, 2] == np.inf
    assert dec[0, 0, 3] == 1.25 and dec[1, 2, 3] == 1.0   # the target's clear alpha is 1
    assert dec[0, 0, 1] == float(np.float16(1e-6))


@pytest.mark.parametrize("mode", [MODE_ADD, MODE_MUL, MODE_POW, MODE_BASE])
def test_string_mod_matches_package(orc, mode):
    mrc = [ModRemColor(3, 1, [1.0, 0.0, 0.0, 1.0]), ModRemColor(5, 2, [0.0, 1.0, 0.0, 0.5])]
    sm = StringMod(modulo=997, num=13, color=[0.2, 0.3, 0.4, 0.5], turns=2, mode=mode, modulo_colors=mrc,
                   scale=0.9)
    pos, c0, c1 = sm.segments()
    opos, oc0, oc1 = orc.string_mod(997, 13, 2, mode, 0.9, [0.2, 0.3, 0.4, 0.5],
                                    [(m.modulo, m.rem, m.color) for m in mrc])
    np.testing.assert_allclose(pos, opos, atol=2e-16)
    np.testing.assert_array_equal(c0, oc0)
    np.testing.assert_array_equal(c1, oc1)


def test_string_mod_known_answers():
    # Mul mode, modulo 4, num 2: 0->0, 1->2, 2->0, 3->2 on the unit circle
    sm = StringMod(modulo=4, num=2)
    pos, _, _ = sm.segments()
    np.testing.assert_allclose(pos, [[1, 0, 1, 0], [0, 1, -1, 0], [-1, 0, 1, 0], [0, -1, -1, 0]], atol=1e-15)
    # Pow mode wraps like u64::pow in a release build
    sm = StringMod(modulo=1000, num=17, mode=MODE_POW)
    t = sm.targets()
    assert int(t[999]) == pow(999, 17, 1 << 64) % 1000
    sm = StringMod(modulo=1000, num=3, mode=MODE_BASE)
    assert int(sm.targets()[77]) == pow(3, 77, 1 << 64) % 1000

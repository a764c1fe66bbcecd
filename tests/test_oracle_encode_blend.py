"""The oracle's screenshot conversion (renderer.rs:313-328) and blend variants (mod.rs:57-73,
gui/settings.rs:57-133), pinned by an independent evaluation and by the algebra of the blend
equation."""
import math

import numpy as np
import pytest

from light_garden_amd import abi
from light_garden_amd.tracer import blend_struct

E = float(np.float32(1.0) / np.float32(2.2))  # `1. / 2.2` in f32


def test_encode_every_half_against_double_evaluation(orc):
    """All 65536 halves: the oracle's f32 powf path against pow in double, except where the
    double value lies within 1e-3 of an integer (where f32 rounding may decide)."""
    h = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    f = h.view(np.float16).astype(np.float64)
    got = np.array([orc.encode_value(int(x)) for x in h], np.int64)
    ambiguous = 0
    for i, x in enumerate(f):
        if x == -math.inf:
            want = 255  # pow(-inf, y) = +inf for y > 0 not an odd integer (C99 F.9.4.4)
        elif math.isnan(x) or x <= 0.0:
            want = 0
        elif x >= 1.0:
            want = 255  # x^e >= 1 and 255 * x^e saturates
        else:
            v = math.pow(x, E) * 255.0
            if abs(v - round(v)) < 1e-3:
                ambiguous += 1
                continue
            want = int(v)
        assert got[i] == want, (hex(int(h[i])), x, got[i], want)
    assert ambiguous < 200


def test_encode_special_values_and_order(orc):
    bits = lambda v: int(np.array(v, np.float16).view(np.uint16))
    assert orc.encode_value(bits(0.5)) == 186  # 0.5^(1/2.2) * 255 = 186.08
    assert orc.encode_value(bits(1.0)) == 255
    assert orc.encode_value(bits(65504.0)) == 255
    assert orc.encode_value(0x7C00) == 255      # +inf
    assert orc.encode_value(0xFC00) == 255      # -inf: powf(-inf, 1/2.2) = +inf
    assert orc.encode_value(0x7E00) == 0        # NaN
    assert orc.encode_value(bits(-0.25)) == 0   # powf(negative, 1/2.2) is NaN
    assert orc.encode_value(0x8000) == 0 and orc.encode_value(0) == 0
    assert orc.encode_value(1) == 0             # the smallest subnormal: 255 * 2^-24^0.4545 < 1
    px = np.array([[[bits(0.5), bits(0.25), bits(1.0), bits(0.0)]]], np.uint16)
    r, g, b, a = (orc.encode_value(int(v)) for v in px[0, 0])
    assert orc.encode_8bit(px, 0).tolist() == [[[b, g, r, a]]]  # the reference's B, G, R, A
    assert orc.encode_8bit(px, 1).tolist() == [[[r, g, b, a]]]


def segments(seed=3, n=300, vertex=False):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.95, 0.95, (n, 4))
    c0 = rng.uniform(0.0, 0.3, (n, 4)).astype(np.float32)
    c1 = rng.uniform(0.0, 0.3, (n, 4)).astype(np.float32) if vertex else None
    return pos, c0, c1


VIEW = abi.LgView(-1, 1, -1, 1, 96, 80)
ONE, ADD, RSUB = abi.BF_ONE, abi.BO_ADD, abi.BO_REVERSE_SUBTRACT


def acc(orc, blend, vertex=False):
    pos, c0, c1 = segments(vertex=vertex)
    return orc.accumulate_blend_f64(pos, c0, c1, VIEW, blend)


@pytest.mark.parametrize("vertex", [False, True])
def test_default_blend_is_the_plain_accumulation(orc, vertex):
    pos, c0, c1 = segments(vertex=vertex)
    d = blend_struct((ONE, ONE, ADD), (abi.BF_SRC_ALPHA, ONE, ADD))
    np.testing.assert_array_equal(orc.accumulate_blend_f64(pos, c0, c1, VIEW, d),
                                  orc.accumulate_f64(pos, c0, c1, VIEW))


@pytest.mark.parametrize("vertex", [False, True])
def test_blend_algebra(orc, vertex):
    one = acc(orc, blend_struct((ONE, ONE, ADD), (ONE, ONE, ADD)), vertex)
    # ReverseSubtract is the exact negation of Add
    np.testing.assert_array_equal(acc(orc, blend_struct((ONE, ONE, RSUB), (ONE, ONE, RSUB)), vertex), -one)
    # Zero adds nothing
    assert not acc(orc, blend_struct((abi.BF_ZERO, ONE, ADD), (abi.BF_ZERO, ONE, ADD)), vertex).any()
    # Constant k scales the One sum by k channel by channel; OneMinusConstant by 1 - k
    k = (0.5, 0.25, 2.0, 0.125)
    np.testing.assert_allclose(acc(orc, blend_struct((abi.BF_CONSTANT, ONE, ADD), (abi.BF_CONSTANT, ONE, ADD), k),
                                   vertex), one * np.array(k), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(acc(orc, blend_struct((abi.BF_ONE_MINUS_CONSTANT, ONE, ADD),
                                                     (abi.BF_ONE_MINUS_CONSTANT, ONE, ADD), k), vertex),
                               one * (1 - np.array(k)), rtol=1e-12, atol=1e-15)
    # OneMinusSrc = One - Src, OneMinusSrcAlpha = One - SrcAlpha (the sums are linear)
    src = acc(orc, blend_struct((abi.BF_SRC, ONE, ADD), (abi.BF_SRC, ONE, ADD)), vertex)
    np.testing.assert_allclose(acc(orc, blend_struct((abi.BF_ONE_MINUS_SRC, ONE, ADD),
                                                     (abi.BF_ONE_MINUS_SRC, ONE, ADD)), vertex),
                               one - src, rtol=1e-9, atol=1e-14)
    sa = acc(orc, blend_struct((abi.BF_SRC_ALPHA, ONE, ADD), (abi.BF_SRC_ALPHA, ONE, ADD)), vertex)
    np.testing.assert_allclose(acc(orc, blend_struct((abi.BF_ONE_MINUS_SRC_ALPHA, ONE, ADD),
                                                     (abi.BF_ONE_MINUS_SRC_ALPHA, ONE, ADD)), vertex),
                               one - sa, rtol=1e-9, atol=1e-14)
    # the alpha channel's Src is its own alpha: equal to SrcAlpha there
    np.testing.assert_array_equal(src[..., 3], sa[..., 3])


def test_blend_known_answer_one_segment(orc):
    """A horizontal segment through pixel centres: each column's pixel gets weight 1; with Src
    on rgb every sample adds c^2, with SrcAlpha on rgb c * a."""
    view = abi.LgView(0, 8, 0, 4, 8, 4)  # one world unit per pixel
    pos = np.array([[0.0, 2.5, 8.0, 2.5]])  # row 1 (from the top), 8 columns
    col = np.array([[0.5, 0.25, 2.0, 0.5]], np.float32)
    a = orc.accumulate_blend_f64(pos, col, None, view, blend_struct((abi.BF_SRC, ONE, ADD), (ONE, ONE, ADD)))
    np.testing.assert_array_equal(a[1, :, :3], np.tile([0.25, 0.0625, 4.0], (8, 1)))
    np.testing.assert_array_equal(a[1, :, 3], np.full(8, 0.5))
    assert not a[[0, 2, 3]].any()
    b = orc.accumulate_blend_f64(pos, col, None, view, blend_struct((abi.BF_SRC_ALPHA, ONE, RSUB), (ONE, ONE, ADD)))
    np.testing.assert_array_equal(b[1, :, :3], np.tile([-0.25, -0.125, -1.0], (8, 1)))


@pytest.mark.parametrize("color,alpha", [
    ((abi.BF_ONE, abi.BF_ONE_MINUS_SRC_ALPHA, ADD), (ONE, ONE, ADD)),  # "over": order dependent
    ((ONE, ONE, abi.BO_SUBTRACT), (ONE, ONE, ADD)),
    ((ONE, ONE, abi.BO_MAX), (ONE, ONE, ADD)),
    ((abi.BF_DST, ONE, ADD), (ONE, ONE, ADD)),
    ((ONE, ONE, ADD), (abi.BF_SRC_ALPHA_SATURATED, ONE, ADD)),
])
def test_order_dependent_blends_are_not_restated(orc, color, alpha):
    pos, c0, c1 = segments()
    assert orc.accumulate_blend_f64(pos, c0, c1, VIEW, blend_struct(color, alpha)) is None

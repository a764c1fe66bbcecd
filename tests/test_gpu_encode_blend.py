"""GPU 8-bit encode (lg_image_encode_8bit, lg_encode_table) and blend variants
(lg_image_set_blend) against the oracle."""
import ctypes as C

import numpy as np
import pytest

import light_garden_amd as lg
from light_garden_amd import abi, configs
from light_garden_amd import string_mod as S
from light_garden_amd.tracer import blend_struct, encode_table

pytestmark = pytest.mark.gpu

ONE, ADD, RSUB = abi.BF_ONE, abi.BO_ADD, abi.BO_REVERSE_SUBTRACT


def test_encode_table_equals_oracle_for_every_half(ctx, orc):
    """The whole conversion domain: 65536 halves, device table against libm's powf."""
    table = encode_table(ctx)
    want = np.array([orc.encode_value(h) for h in range(65536)], np.uint8)
    bad = np.nonzero(table != want)[0]
    assert bad.size == 0, [(hex(int(h)), int(table[h]), int(want[h])) for h in bad[:10]]


def render(ctx, scene, blend=None):
    tr = lg.Tracer(ctx, scene.canvas)
    for o in scene.objects:
        tr.push_object(o)
    for l in scene.lights:
        tr.push_light(l)
    tr.max_bounce = scene.max_bounce
    img = lg.Image(ctx, scene.width, scene.height, scene.canvas)
    if blend is not None:
        img.set_blend(*blend)
    segs = lg.Segments(ctx)
    tr.render(segs, img)
    pos, col, _, _ = segs.download()
    out = img.rgba16f(), img.accum(), img.encode_8bit(abi.ENCODE_BGRA), img.encode_8bit(abi.ENCODE_RGBA), pos, col
    segs.close()
    img.close()
    tr.close()
    return out


def test_encode_rendered_image(ctx, orc):
    scene = configs.plain_small(num_rays=5000)
    scene.width = scene.height = 256
    rgba16f, _, bgra, rgba, _, _ = render(ctx, scene)
    np.testing.assert_array_equal(bgra, orc.encode_8bit(rgba16f, 0))
    np.testing.assert_array_equal(rgba, orc.encode_8bit(rgba16f, 1))
    np.testing.assert_array_equal(bgra[..., [2, 1, 0, 3]], rgba)
    assert rgba[..., 3].min() == 255  # the clear alpha of 1


def test_default_blend_set_explicitly_is_bit_identical(ctx):
    scene = configs.plain_small(num_rays=5000)
    scene.width = scene.height = 256
    a = render(ctx, scene)[0]
    b = render(ctx, scene, blend=((ONE, ONE, ADD), (abi.BF_SRC_ALPHA, ONE, ADD)))[0]
    np.testing.assert_array_equal(a, b)
    img = lg.Image(ctx, 8, 8)
    assert img.blend() == ((ONE, ONE, ADD), (abi.BF_SRC_ALPHA, ONE, ADD), (0.0, 0.0, 0.0, 0.0))
    img.close()


BLENDS = {
    "one_one": ((ONE, ONE, ADD), (ONE, ONE, ADD), (0, 0, 0, 0)),
    "src": ((abi.BF_SRC, ONE, ADD), (abi.BF_SRC, ONE, ADD), (0, 0, 0, 0)),
    "one_minus_src_alpha": ((abi.BF_ONE_MINUS_SRC_ALPHA, ONE, ADD), (abi.BF_ONE_MINUS_SRC, ONE, ADD), (0, 0, 0, 0)),
    "constant": ((abi.BF_CONSTANT, ONE, ADD), (abi.BF_ONE_MINUS_CONSTANT, ONE, ADD), (0.5, 2.0, 0.25, 0.75)),
    "reverse_subtract": ((abi.BF_SRC_ALPHA, ONE, RSUB), (ONE, ONE, RSUB), (0, 0, 0, 0)),
}


def rel_l1(got, ref):
    return np.abs(got - ref).sum() / max(np.abs(ref).sum(), 1e-30)


@pytest.mark.parametrize("name", sorted(BLENDS))
def test_blend_trace_scene_matches_oracle(ctx, orc, name):
    scene = configs.plain_small(num_rays=20000)
    scene.width = scene.height = 384
    blend = BLENDS[name]
    _, accum, _, _, pos, col = render(ctx, scene, blend=blend)
    ref = orc.accumulate_blend_f64(pos.astype(np.float64), col, None, scene.view(), blend_struct(*blend))
    err = rel_l1(accum.astype(np.float64), ref)
    print(name, err)
    assert err < 1e-3


@pytest.mark.parametrize("name", ["src", "one_minus_src_alpha", "reverse_subtract"])
def test_blend_vertex_colours_match_oracle(ctx, orc, name):
    """Per-vertex colours (string mod with modulo colours): the factor is applied per sample
    to the interpolated colour."""
    mods = [S.StringMod(modulo=3000, num=7, color=[0.01, 0.02, 0.03, 0.5],
                        modulo_colors=[S.ModRemColor(2, 1, [0.05, 0.0, 0.01, 0.9])], scale=0.9)]
    segs = lg.Segments(ctx, 16, vertex_colors=True)
    S.draw(ctx, mods, segs)
    pos, c0, c1, _ = segs.download()
    img = lg.Image(ctx, 256, 256)
    img.set_blend(*BLENDS[name])
    img.accumulate(segs)
    got = img.accum().astype(np.float64)
    ref = orc.accumulate_blend_f64(pos.astype(np.float64), c0, c1, img.view, blend_struct(*BLENDS[name]))
    segs.close()
    img.close()
    err = rel_l1(got, ref)
    print(name, err)
    assert err < 1e-3


def test_blends_in_one_batch_equal_separate_renders(ctx):
    scenes = [configs.plain_small(num_rays=3000 + 1000 * i) for i in range(3)]
    blends = [None, BLENDS["src"], BLENDS["reverse_subtract"]]
    alone = [render(ctx, sc, bl)[0] for sc, bl in zip(scenes, blends)]
    trs, segs, imgs = [], [], []
    for sc, bl in zip(scenes, blends):
        tr = lg.Tracer(ctx, sc.canvas)
        for o in sc.objects:
            tr.push_object(o)
        for l in sc.lights:
            tr.push_light(l)
        img = lg.Image(ctx, sc.width, sc.height, sc.canvas)
        if bl is not None:
            img.set_blend(*bl)
        trs.append(tr)
        segs.append(lg.Segments(ctx))
        imgs.append(img)
    lg.batch_render(ctx, trs, segs, imgs)
    for a, img in zip(alone, imgs):
        np.testing.assert_array_equal(img.rgba16f(), a)
    for x in trs + segs + imgs:
        x.close()


def test_blend_rejections(ctx):
    img = lg.Image(ctx, 8, 8)
    lib = abi.load()
    unsupported = [((ONE, abi.BF_ONE_MINUS_SRC_ALPHA, ADD), (ONE, ONE, ADD)),
                   ((ONE, ONE, abi.BO_SUBTRACT), (ONE, ONE, ADD)),
                   ((ONE, ONE, abi.BO_MIN), (ONE, ONE, ADD)),
                   ((ONE, ONE, ADD), (abi.BF_DST_ALPHA, ONE, ADD)),
                   ((abi.BF_SRC_ALPHA_SATURATED, ONE, ADD), (ONE, ONE, ADD))]
    for color, alpha in unsupported:
        b = blend_struct(color, alpha)
        assert lib.lg_image_set_blend(img.handle, C.byref(b)) == abi.LG_E_UNSUPPORTED
    b = blend_struct((13, ONE, ADD), (ONE, ONE, ADD))
    assert lib.lg_image_set_blend(img.handle, C.byref(b)) == abi.LG_E_INVALID
    b = blend_struct((ONE, ONE, 5), (ONE, ONE, ADD))
    assert lib.lg_image_set_blend(img.handle, C.byref(b)) == abi.LG_E_INVALID
    b = blend_struct((ONE, ONE, ADD), (ONE, ONE, ADD), (float("nan"), 0, 0, 0))
    assert lib.lg_image_set_blend(img.handle, C.byref(b)) == abi.LG_E_INVALID
    # a rejected state leaves the image's blend as it was
    assert img.blend() == ((ONE, ONE, ADD), (abi.BF_SRC_ALPHA, ONE, ADD), (0.0, 0.0, 0.0, 0.0))
    out = np.empty(8 * 8 * 4, np.uint8)
    assert lib.lg_image_encode_8bit(img.handle, 2, out.ctypes.data, out.size) == abi.LG_E_INVALID
    assert lib.lg_image_encode_8bit(img.handle, 0, out.ctypes.data, out.size - 1) == abi.LG_E_CAPACITY
    img.close()

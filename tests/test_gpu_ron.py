"""A RON scene loaded with lg_scene_load_ron traces like the same scene built by hand, checked
against the FP64 oracle; lg_scene_save_ron writes the file back (tracer.rs:183-204)."""
import ctypes as C
import os

import numpy as np
import pytest

import light_garden_amd as lg
from light_garden_amd import abi, configs
from light_garden_amd import geo as G

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_RON = os.path.join(HERE, "golden", "default.ron")


def default_scene_by_hand():
    """default.ron's scene, typed from the file's values."""
    lens = lg.Object(G.Logic(abi.LOGIC_AND, G.Circle((1.9, 0.0), 2.0), G.Circle((-1.9, 0.0), 2.0),
                             (-0.022772240638732733, -0.09999999999999998),
                             (0.00000000000000006123233995736766, -1.0, 1.0, 0.00000000000000006123233995736766)),
                     lg.Material(1.05))
    curved = lg.Object(G.CubicBezier([(-0.622772240638733, 0.40000000000000013), (-0.3227722406387328, 0.8),
                                      (0.2772277593612673, 0.8), (0.5772277593612682, 0.40000000000000013)]), None)
    rect = lg.Object(G.Rect((-0.022772240638732733, -0.5), (1.0, 0.0, 0.0, 1.0), 0.40000000000000036,
                            0.3999999999999999), lg.Material(1.73))
    point = lg.PointLight((-0.022772240638732733, -0.5), 10000, [0.009721218, 0.009721218, 0.009721218, 0.011764706])
    spot = lg.SpotLight((1.0772277593612674, -0.09999999999999998), 0.17453292519943295,
                        (-0.9999922358557027, 0.003940587305547067), 2000,
                        [0.0036765062, 0.020288562, 0.016807375, 0.03529412])
    return configs.Scene([lens, curved, rect], [point, spot], max_bounce=5, width=512, height=512, name="default_ron")


def test_loaded_scene_matches_oracle(ctx, orc):
    text = open(DEFAULT_RON, "rb").read()
    scene = default_scene_by_hand()
    h = C.c_void_p()
    abi.call("lg_scene_create", ctx.handle, C.byref(h))
    segs = lg.Segments(ctx)
    try:
        abi.call("lg_scene_load_ron", h, text, len(text))
        no, nl, nr = C.c_uint32(), C.c_uint32(), C.c_uint64()
        abi.call("lg_scene_counts", h, C.byref(no), C.byref(nl), C.byref(nr))
        assert (no.value, nl.value, nr.value) == (3, 2, 12000)
        p = scene.params(fp64=True)
        abi.call("lg_trace", h, C.byref(p), segs.handle)
        pos, col, _, key = segs.download()
        pos64 = segs.download_pos64()
        rpos, rcol, rkey = orc.trace_scene(scene, threads=8)
        m = orc.match_segments(key, pos64, rkey, rpos)
        print(m)
        assert m["n_a"] == m["n_b"] > 12000
        assert m["matched"] == m["n_b"]
        # the point light sits at the centre of the 0.4 x 0.4 rect, so its 45-degree rays meet
        # the rect's corners, where the exit edge (and the child rays) is decided by rounding
        assert m["diverged"] <= 1e-4 * m["n_b"]
        assert m["max_err"] < 1e-9
        np.testing.assert_allclose(col[np.argsort(key)], rcol[np.argsort(rkey)], rtol=1e-6, atol=0)

        # save writes the file back byte for byte
        n = C.c_uint64()
        buf = C.create_string_buffer(len(text) + 1)
        abi.call("lg_scene_save_ron", h, buf, len(text) + 1, C.byref(n))
        assert n.value == len(text) and buf.value == text

        # a text that does not parse leaves the scene as it was
        bad = b"([(object_enum: Circle((origin: [0, 0])), material_opt: None, moved: false)], [])"
        assert lg.abi.load().lg_scene_load_ron(h, bad, len(bad)) == abi.LG_E_INVALID
        abi.call("lg_scene_counts", h, C.byref(no), C.byref(nl), C.byref(nr))
        assert (no.value, nl.value, nr.value) == (3, 2, 12000)
    finally:
        segs.close()
        abi.load().lg_scene_destroy(h)


def test_tracer_load_renders_like_pushed_scene(ctx):
    """Tracer.load then render equals the hand-built scene's render bit for bit."""
    scene = default_scene_by_hand()
    a = lg.Tracer(ctx, scene.canvas)
    a.load(open(DEFAULT_RON).read())
    b = lg.Tracer(ctx, scene.canvas)
    for o in scene.objects:
        b.push_object(o)
    for l in scene.lights:
        b.push_light(l)
    imgs = []
    for tr in (a, b):
        img = lg.Image(ctx, scene.width, scene.height, scene.canvas)
        segs = lg.Segments(ctx)
        tr.render(segs, img)
        imgs.append(img.rgba16f())
        segs.close()
        img.close()
    a.close()
    b.close()
    assert imgs[0].any()
    np.testing.assert_array_equal(imgs[0], imgs[1])

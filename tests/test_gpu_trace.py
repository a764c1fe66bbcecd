"""GPU trace parity against the FP64 oracle (through the C ABI)."""
import ctypes as C

import numpy as np
import pytest

import light_garden_amd as lg
from light_garden_amd import abi, configs

pytestmark = pytest.mark.gpu

PATH_MASK = np.uint64((1 << 24) - 1)


def make_tracer(ctx, scene, fp64=False):
    tr = lg.Tracer(ctx, scene.canvas)
    for o in scene.objects:
        tr.push_object(o)
    for l in scene.lights:
        tr.push_light(l)
    tr.max_bounce = scene.max_bounce
    tr.cutoff_color = list(scene.cutoff)
    tr.fp64 = fp64
    return tr


def depth_of(key):
    return np.floor(np.log2((np.asarray(key, np.uint64) & PATH_MASK).astype(np.float64))).astype(int)


def gpu_trace(ctx, scene, fp64):
    tr = make_tracer(ctx, scene, fp64)
    segs = tr.trace()
    pos, col, _, key = segs.download()
    pos64 = segs.download_pos64() if fp64 else pos.astype(np.float64)
    return pos64, col, key


@pytest.mark.parametrize("scene_fn", [lambda: configs.plain_small(),
                                      lambda: configs.large_single(num_rays=30000, width=512)],
                         ids=["plain_small", "large_single_30k"])
def test_trace_fp64_matches_oracle(ctx, orc, scene_fn):
    scene = scene_fn()
    pos, col, key = gpu_trace(ctx, scene, fp64=True)
    rpos, rcol, rkey = orc.trace_scene(scene, threads=8)
    m = orc.match_segments(key, pos, rkey, rpos)
    print(scene.name, m)
    # same tree: every segment present on both sides, end points to FP64 rounding
    assert m["n_a"] == m["n_b"]
    assert m["matched"] == m["n_b"]
    assert m["diverged"] == 0
    assert m["max_err"] < 1e-9
    order_a, order_b = np.argsort(key), np.argsort(rkey)
    np.testing.assert_allclose(col[order_a], rcol[order_b], rtol=1e-6, atol=0)


@pytest.mark.parametrize("scene_fn", [lambda: configs.plain_small(),
                                      lambda: configs.large_single(num_rays=30000, width=512)],
                         ids=["plain_small", "large_single_30k"])
def test_trace_fp32_matches_oracle(ctx, orc, scene_fn):
    """FP32 bulk trace: end points of matched segments within 1e-5 of the world extent;
    branch flips (segments present on one side only, or matched keys whose geometry
    diverged after a flipped decision upstream) stay rare."""
    scene = scene_fn()
    pos, col, key = gpu_trace(ctx, scene, fp64=False)
    rpos, rcol, rkey = orc.trace_scene(scene, threads=8)
    m = orc.match_segments(key, pos, rkey, rpos)
    print(scene.name, m)
    n = m["n_b"]
    assert m["matched"] >= 0.995 * n
    assert m["diverged"] <= 0.005 * n
    assert m["p99_err"] < 1e-5
    # per depth: the primary generation has no upstream flips
    d = depth_of(key)
    prim = d == 0
    mp = orc.match_segments(key[prim], pos[prim], rkey[depth_of(rkey) == 0], rpos[depth_of(rkey) == 0])
    assert mp["matched"] == mp["n_b"]


def test_trace_all_reference_order(ctx, orc):
    """lg_trace_all returns vertex pairs in trace_all's order (tracer.rs:276-330)."""
    scene = configs.plain_small(num_rays=500)
    tr = make_tracer(ctx, scene, fp64=True)
    tr.sync()
    p = tr.params()
    n = C.c_uint64()
    rc = abi.load().lg_trace_all(tr._h, C.byref(p), None, None, 0, C.byref(n))
    assert rc == abi.LG_E_CAPACITY
    xy = np.empty((n.value, 2), np.float64)
    rgba = np.empty((n.value, 4), np.float32)
    abi.call("lg_trace_all", tr._h, C.byref(p), xy.ctypes.data, rgba.ctypes.data, n.value, C.byref(n))
    rpos, rcol, rkey = orc.trace_scene(scene)
    ref_xy = rpos.reshape(-1, 2)
    assert xy.shape == ref_xy.shape
    np.testing.assert_allclose(xy, ref_xy, atol=1e-9, rtol=0)
    np.testing.assert_allclose(rgba, np.repeat(rcol, 2, axis=0), rtol=1e-6)


def test_trace_empty_and_zero_bounce(ctx):
    scene = configs.plain_small(num_rays=100)
    tr = make_tracer(ctx, scene)
    tr.max_bounce = 0
    assert tr.trace().count() == 0
    tr.max_bounce = 5
    tr.clear()
    assert tr.trace().count() == 0


def test_trace_rejects_bad_scene(ctx):
    tr = lg.Tracer(ctx)
    tr.push_object(lg.Object.new_circle((0.0, 0.0), -1.0))
    tr.push_light(lg.PointLight((0.5, 0.5), 10, [1, 1, 1, 1]))
    with pytest.raises(abi.LgError) as e:
        tr.trace()
    assert e.value.code == abi.LG_E_INVALID


def test_trace_capacity_regrow(ctx, orc):
    """A segment buffer created far too small is regrown and the trace completes."""
    scene = configs.plain_small(num_rays=3000)
    tr = make_tracer(ctx, scene)
    segs = lg.Segments(ctx, capacity=8)
    tr.trace(segs)
    _, _, rkey = orc.trace_scene(scene)
    assert abs(segs.count() - len(rkey)) <= 0.01 * len(rkey)


def shape_zoo():
    """Every shape kind and boolean op the GPU path handles, overlapping each other."""
    from light_garden_amd import geo as G
    rng = np.random.default_rng(31)
    objs = []
    prims = lambda: [G.Circle((rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1)), rng.uniform(0.1, 0.25)),
                     G.Rect((rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1)), G.rotation(rng.uniform(0, 3)),
                            rng.uniform(0.1, 0.4), rng.uniform(0.1, 0.4)),
                     G.Ellipse((rng.uniform(-0.1, 0.1), 0.0), rng.uniform(0.1, 0.3), rng.uniform(0.05, 0.2),
                               G.rotation(rng.uniform(0, 3))),
                     G.ConvexPolygon.new_convex_hull([(rng.uniform(-0.25, 0.25), rng.uniform(-0.25, 0.25))
                                                      for _ in range(6)])]
    centres = [(x, y) for x in (-0.6, -0.2, 0.2, 0.6) for y in (-0.6, -0.2, 0.2, 0.6)]
    k = 0
    for op in (abi.LOGIC_AND, abi.LOGIC_OR, abi.LOGIC_AND_NOT):
        for i in range(4):
            a, b = prims()[i], prims()[(i + 1 + op) % 4]
            c = centres[k % len(centres)]
            k += 1
            objs.append(lg.Object(G.Logic(op, a, b, c, G.rotation(rng.uniform(0, 3))),
                                  lg.Material(rng.uniform(1.3, 1.8))))
    for p in prims():
        c = centres[k % len(centres)]
        k += 1
        objs.append(lg.Object(p.translated(*c), lg.Material(rng.uniform(1.3, 1.8))))
    objs.append(lg.Object.new_mirror((-0.9, 0.85), (0.9, 0.8)))
    objs.append(lg.Object.new_curved_mirror(configs.arc_bezier((0.0, 0.0), 0.95, -0.6, 0.6)))
    mirrored_circle = lg.Object.new_circle((0.0, -0.85), 0.08)
    mirrored_circle.material_opt = None  # a circle without material reflects
    objs.append(mirrored_circle)
    lights = [lg.PointLight((0.01, 0.02), 4000, [0.02, 0.02, 0.02, 0.02]),
              lg.SpotLight((-0.95, -0.95), 0.6, (1.0, 1.0), 3000, [0.03, 0.01, 0.02, 0.02]),
              lg.DirectionalLight([0.01, 0.03, 0.02, 0.02], 3000, lg.LineSegment((-0.9, 0.95), (0.9, 0.95)))]
    # the left normal of b - a = (1.8, 0) points up, out of the canvas: reverse the band
    lights[2].start = lg.LineSegment((0.9, 0.95), (-0.9, 0.95))
    return configs.Scene(objs, lights, max_bounce=8, width=256, height=256, name="shape_zoo")


@pytest.mark.parametrize("fp64", [True, False], ids=["fp64", "fp32"])
def test_trace_shape_zoo(ctx, orc, fp64):
    """And/Or/AndNot over circle/rect/ellipse/polygon operands, every primitive kind, a
    mirror without material: the kernel's interval boolean rule and chords agree with the
    oracle's hit-filter rule."""
    scene = shape_zoo()
    pos, col, key = gpu_trace(ctx, scene, fp64=fp64)
    rpos, rcol, rkey = orc.trace_scene(scene, threads=8)
    m = orc.match_segments(key, pos, rkey, rpos)
    print(m)
    assert m["n_b"] > 20000
    assert m["matched"] >= 0.999 * m["n_b"]
    assert m["diverged"] <= 0.001 * m["n_b"]
    assert m["max_err"] < (1e-9 if fp64 else 1e-6)


def test_python_trace_all_matches_c_entry(ctx, orc):
    scene = configs.plain_small(num_rays=700)
    tr = make_tracer(ctx, scene, fp64=True)
    xy, rgba = tr.trace_all()
    rpos, rcol, _ = orc.trace_scene(scene)
    np.testing.assert_allclose(xy, rpos.reshape(-1, 2), atol=1e-9, rtol=0)
    np.testing.assert_array_equal(rgba, np.repeat(rcol, 2, axis=0))

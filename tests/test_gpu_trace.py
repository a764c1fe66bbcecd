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

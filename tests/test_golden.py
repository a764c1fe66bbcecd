"""Golden vectors (tests/golden/*.npz, written by tests/golden/make_golden.py): the oracle
must reproduce them exactly (CPU), and the GPU path must match them (GPU)."""
import os
import sys

import numpy as np
import pytest

import light_garden_amd as lg
from light_garden_amd import abi, configs

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import golden_scenes  # noqa: E402


def load(name):
    return dict(np.load(os.path.join(HERE, "golden", name + ".npz")))


@pytest.mark.parametrize("name", ["plain_small_1500", "large_single_1200"])
def test_oracle_reproduces_trace_golden(orc, name):
    g = load(name)
    sc = golden_scenes()[name]
    pos, col, key = orc.trace_scene(sc)
    assert np.array_equal(key, g["key"])
    np.testing.assert_allclose(pos, g["pos"], rtol=0, atol=1e-13)
    assert np.array_equal(col, g["col"])
    img = orc.resolve_f64(orc.accumulate_f64(pos, col, None, sc.view()))
    assert np.array_equal(img, g["img"])


def test_oracle_reproduces_string_mod_golden(orc):
    g = load("string_mod_2000")
    spec = configs.StringModSpec(modulo=2000, mults=(2, 3, 7), width=128, height=128)
    pos, col = configs.string_mod_segments(spec)
    np.testing.assert_allclose(pos, g["pos"], atol=1e-15)
    assert np.array_equal(col, g["col"])
    img = orc.resolve_f64(orc.accumulate_f64(pos, col, col, abi.LgView(-1, 1, -1, 1, 128, 128)))
    assert np.array_equal(img, g["img"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["plain_small_1500", "large_single_1200"])
def test_gpu_matches_trace_golden(ctx, name):
    g = load(name)
    sc = golden_scenes()[name]
    tr = lg.Tracer(ctx, sc.canvas)
    for o in sc.objects:
        tr.push_object(o)
    for l in sc.lights:
        tr.push_light(l)
    tr.max_bounce = sc.max_bounce
    segs = lg.Segments(ctx)
    img = lg.Image(ctx, sc.width, sc.height, sc.canvas)
    tr.render(segs, img)
    pos, col, _, key = segs.download()
    order = np.argsort(key)
    assert np.array_equal(key[order], g["key"])
    np.testing.assert_allclose(pos[order], g["pos"], atol=1e-7, rtol=0)
    ref = g["img"].view(np.float16).astype(np.float64)
    got = img.rgba().astype(np.float64)
    err = np.abs(got - ref)[..., :3].sum() / np.abs(ref[..., :3]).sum()
    assert err < 1e-3


@pytest.mark.gpu
def test_gpu_matches_string_mod_golden(ctx):
    g = load("string_mod_2000")
    segs = lg.Segments(ctx, len(g["pos"]), vertex_colors=True)
    segs.upload(g["pos"], g["col"], g["col"])
    img = lg.Image(ctx, 128, 128)
    img.accumulate(segs)
    img.resolve()
    ref = g["img"].view(np.float16).astype(np.float64)
    got = img.rgba().astype(np.float64)
    err = np.abs(got - ref)[..., :3].sum() / np.abs(ref[..., :3]).sum()
    assert err < 1e-3

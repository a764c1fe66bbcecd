"""Serving loop (lg_server_*): results alone and in a mix, edits, HBM-budget evictions."""
import copy

import numpy as np
import pytest

import light_garden_amd as lg
from light_garden_amd import abi, configs

pytestmark = pytest.mark.gpu


def tracer_for(ctx, sc):
    tr = lg.Tracer(ctx, sc.canvas)
    for o in sc.objects:
        tr.push_object(o)
    for l in sc.lights:
        tr.push_light(l)
    tr.max_bounce = sc.max_bounce
    tr.cutoff_color = list(sc.cutoff)
    return tr


def small_scenes(n, seed=9):
    scenes = configs.batched_uneven(n_scenes=n, seed=seed, ray_scale=0.01)
    for s in scenes:
        s.width = s.height = min(s.width, 512)
    return scenes


def direct_render(ctx, tr, sc):
    im = lg.Image(ctx, sc.width, sc.height, sc.canvas)
    tr.render(lg.Segments(ctx), im)
    return im.rgba16f()


def test_server_alone_equals_mix(ctx):
    """Each request's image is the same bits rendered alone, in a mixed continuous batch and
    by a plain lg_render."""
    scenes = small_scenes(12)
    trs = [tracer_for(ctx, s) for s in scenes]
    with lg.Server(ctx) as srv:
        alone = []
        for tr, sc in zip(trs, scenes):
            im = lg.Image(ctx, sc.width, sc.height, sc.canvas)
            r = srv.wait(srv.submit(tr, im), 60)
            assert r["batch_size"] == 1
            alone.append(im.rgba16f())
        mixed = [lg.Image(ctx, sc.width, sc.height, sc.canvas) for sc in scenes]
        tickets = [srv.submit(tr, im) for tr, im in zip(trs, mixed)]
        res = [srv.wait(t, 60) for t in tickets]
        assert max(r["batch_size"] for r in res) > 1  # requests shared frames
        for a, im in zip(alone, mixed):
            assert np.array_equal(a, im.rgba16f())
        st = srv.stats()
        assert st["completed"] == 24 and st["submitted"] == 24
    for tr, sc, a in zip(trs, scenes, alone):
        assert np.array_equal(direct_render(ctx, tr, sc), a)


def test_server_edits_in_order(ctx):
    """Edits apply right before their render, in submission order per scene, while other
    scenes' requests interleave."""
    sc = configs.large_single(num_rays=20000, width=512)
    others = small_scenes(6, seed=11)
    tr = tracer_for(ctx, sc)
    o_trs = [tracer_for(ctx, s) for s in others]
    rng = configs.SplitMix64(17)
    edits = [(3, configs.edit_object(sc.objects[3], rng)), (10, configs.edit_object(sc.objects[10], rng)),
             (3, configs.edit_object(sc.objects[3], rng))]
    o_imgs = [lg.Image(ctx, s.width, s.height, s.canvas) for s in others]
    with lg.Server(ctx, slots=2) as srv:
        imgs, tickets = [], []
        for ix, obj in edits:
            im = lg.Image(ctx, sc.width, sc.height, sc.canvas)
            imgs.append(im)
            tickets.append(srv.submit(tr, im, edit=(ix, obj)))
            for otr, oim in zip(o_trs, o_imgs):
                tickets.append(srv.submit(otr, oim))
        for t in tickets:
            srv.wait(t, 60)
    ref = tracer_for(ctx, sc)
    for (ix, obj), im in zip(edits, imgs):
        ref.replace_object(ix, obj)
        assert np.array_equal(direct_render(ctx, ref, sc), im.rgba16f())


def test_server_budget_evicts_and_stays_exact(ctx):
    scenes = small_scenes(8, seed=13)
    trs = [tracer_for(ctx, s) for s in scenes]
    refs = [direct_render(ctx, tr, sc) for tr, sc in zip(trs, scenes)]
    imgs = [lg.Image(ctx, sc.width, sc.height, sc.canvas) for sc in scenes]
    with lg.Server(ctx, hbm_budget=24 << 20, slots=1) as srv:
        for rep in range(3):
            for tr, im in zip(trs, imgs):
                srv.wait(srv.submit(tr, im), 60)
            for im, ref in zip(imgs, refs):
                assert np.array_equal(im.rgba16f(), ref)
        st = srv.stats()
    assert st["evictions"] > 0
    assert st["completed"] == 24


def test_server_rejects_bad_requests(ctx):
    sc = small_scenes(1)[0]
    tr = tracer_for(ctx, sc)
    im = lg.Image(ctx, sc.width, sc.height, sc.canvas)
    with lg.Server(ctx) as srv:
        with pytest.raises(abi.LgError):
            srv.submit(tr, im, edit=(len(sc.objects) + 5, sc.objects[0]))
        tr.fp64 = True
        with pytest.raises(abi.LgError):
            srv.submit(tr, im)
        tr.fp64 = False
        with pytest.raises(abi.LgError) as e:
            srv.wait(123456, timeout=0.01)
        assert e.value.code == abi.LG_E_TIMEOUT

"""The BVH culling pass (lg_tuning.bvh_min_objects) returns the same segments, bit for bit,
as the walk over every object, which is the reference's brute-force candidate selection
(tracer.rs:413-423); and images stay the same."""
import numpy as np
import pytest

import light_garden_amd as lg
from light_garden_amd import configs

pytestmark = pytest.mark.gpu

NEVER = 0xFFFFFFFF


def trace_sorted(ctx, scene, min_objects):
    ctx.set_tuning(bvh_min_objects=min_objects)
    tr = lg.Tracer(ctx, scene.canvas)
    for o in scene.objects:
        tr.push_object(o)
    for l in scene.lights:
        tr.push_light(l)
    tr.max_bounce = scene.max_bounce
    segs = tr.trace()
    pos, col, _, key = segs.download()
    img = lg.Image(ctx, 256, 256, scene.canvas)
    tr.render(segs, img)
    out = img.rgba16f()
    segs.close()
    img.close()
    tr.close()
    order = np.argsort(key, kind="stable")
    return key[order], pos[order], col[order], out


def scenes():
    from test_gpu_trace import shape_zoo
    out = [("large_single_64", configs.large_single(num_rays=30000, width=256)), ("shape_zoo", shape_zoo())]
    many = [sc for sc in configs.batched_uneven(n_scenes=40, seed=4, ray_scale=0.1) if len(sc.objects) >= 32]
    out += [("batched_%d_objs" % len(sc.objects), sc) for sc in many[:4]]
    return out


@pytest.mark.parametrize("name,scene", scenes(), ids=lambda v: v if isinstance(v, str) else "")
def test_bvh_equals_full_walk(ctx, name, scene):
    before = ctx.tuning()
    try:
        full = trace_sorted(ctx, scene, NEVER)
        bvh = trace_sorted(ctx, scene, 1)
    finally:
        ctx.set_tuning(**before)
    assert len(full[0]) > 100
    for a, b in zip(full, bvh):
        np.testing.assert_array_equal(a, b)

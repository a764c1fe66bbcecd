"""Host-side geometry helpers (no GPU): the transforms serving edits use."""
import math

from light_garden_amd import configs, geo as G


def test_scaled_shapes():
    c = G.Circle((0.2, 0.1), 0.5).scaled(2.0)
    assert c.origin == (0.2, 0.1) and c.radius == 1.0
    s = G.LineSegment((0.0, 0.0), (2.0, 0.0)).scaled(0.5)
    assert s.a == (0.5, 0.0) and s.b == (1.5, 0.0)
    r = G.Rect((0.0, 0.0), G.rotation(0.3), 0.2, 0.4).scaled(1.5)
    assert math.isclose(r.width, 0.3) and math.isclose(r.height, 0.6)
    lens = G.lens((0.3, 0.3), 0.1, 0.05).scaled(2.0)
    assert lens.origin == (0.3, 0.3)
    assert math.isclose(lens.a.radius, 0.2) and math.isclose(lens.a.origin[0], 0.05)
    b = G.CubicBezier([(0, 0), (1, 0), (1, 1), (0, 1)]).scaled(2.0)
    assert b.points[0] == (-0.5, -0.5) and b.points[2] == (1.5, 1.5)


def test_edit_object_moves_or_resizes():
    sc = configs.large_single(num_rays=10)
    rng = configs.SplitMix64(5)
    kinds = set()
    for i in range(40):
        o = sc.objects[i % len(sc.objects)]
        e = configs.edit_object(o, rng)
        assert type(e.geometry) is type(o.geometry)
        assert e.material_opt is o.material_opt
        kinds.add(e.geometry != o.geometry)
    assert kinds == {True}

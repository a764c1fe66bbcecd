"""Pins the oracle's geometry (the collision2d restatement) to closed forms, invariants and
brute force. CPU only."""
import math

import numpy as np
import pytest

import light_garden_amd as lg
from light_garden_amd import abi
from light_garden_amd import geo as G

T_MIN = 1e-5


def obj(geometry, n=None):
    o = lg.Object(geometry, lg.Material(n) if n else None)
    return o.to_abi()


def unit(a):
    return (math.cos(a), math.sin(a))


# ---- closed forms ---------------------------------------------------------------------------
def test_circle_known_answers(orc):
    c = obj(G.Circle((0.0, 0.0), 1.0), 1.5)
    hits = orc.intersect(c, (-2.0, 0.0), (1.0, 0.0))
    assert len(hits) == 2
    (t1, p1, n1), (t2, p2, n2) = hits
    assert (t1, t2) == pytest.approx((1.0, 3.0), abs=1e-15)
    assert p1 == pytest.approx((-1.0, 0.0), abs=1e-15) and n1 == pytest.approx((-1.0, 0.0))
    assert p2 == pytest.approx((1.0, 0.0), abs=1e-15) and n2 == pytest.approx((1.0, 0.0))
    # off-centre chord: y = 0.6 -> x = +-0.8
    hits = orc.intersect(c, (-2.0, 0.6), (1.0, 0.0))
    assert [h[1][0] for h in hits] == pytest.approx([-0.8, 0.8], abs=1e-14)
    assert hits[0][2] == pytest.approx((-0.8, 0.6), abs=1e-14)
    # from inside: only the exit
    hits = orc.intersect(c, (0.0, 0.0), unit(0.3))
    assert len(hits) == 1 and hits[0][0] == pytest.approx(1.0, abs=1e-15)
    # misses: passing above, and pointing away
    assert orc.intersect(c, (-2.0, 1.5), (1.0, 0.0)) == []
    assert orc.intersect(c, (2.0, 0.0), (1.0, 0.0)) == []
    # a ray starting on the surface does not hit its own start (t <= T_MIN)
    hits = orc.intersect(c, (-1.0, 0.0), (1.0, 0.0))
    assert len(hits) == 1 and hits[0][0] == pytest.approx(2.0)


def test_segment_known_answers(orc):
    m = obj(G.LineSegment((0.0, -1.0), (0.0, 1.0)))
    hits = orc.intersect(m, (-1.0, 0.5), (1.0, 0.0))
    assert len(hits) == 1
    t, p, n = hits[0]
    assert t == pytest.approx(1.0) and p == pytest.approx((0.0, 0.5))
    # left normal of b - a = (0, 2) is (-1, 0)
    assert n == pytest.approx((-1.0, 0.0))
    # end points are included
    assert len(orc.intersect(m, (-1.0, 1.0), (1.0, 0.0))) == 1
    assert orc.intersect(m, (-1.0, 1.0 + 1e-9), (1.0, 0.0)) == []
    # parallel ray misses
    assert orc.intersect(m, (-1.0, 0.0), (0.0, 1.0)) == []


def test_rect_known_answers(orc):
    # centre (1, 2), width 2, height 1, rotated 90 degrees: world extent x in [0.5, 1.5],
    # y in [1, 3]
    r = obj(G.Rect((1.0, 2.0), G.rotation(math.pi / 2), 2.0, 1.0), 1.3)
    hits = orc.intersect(r, (0.0, 2.5), (1.0, 0.0))
    assert [h[1][0] for h in sorted(hits)] == pytest.approx([0.5, 1.5], abs=1e-14)
    ns = [h[2] for h in sorted(hits)]
    assert ns[0] == pytest.approx((-1.0, 0.0), abs=1e-14)
    assert ns[1] == pytest.approx((1.0, 0.0), abs=1e-14)
    hits = orc.intersect(r, (1.2, 0.0), (0.0, 1.0))
    assert sorted(h[1][1] for h in hits) == pytest.approx([1.0, 3.0], abs=1e-14)
    assert orc.contains(r, (1.0, 2.9)) and not orc.contains(r, (1.6, 2.0))
    # boundary is outside (strict)
    assert not orc.contains(r, (1.5, 2.0))


def test_ellipse_polygon_known_answers(orc):
    e = obj(G.Ellipse((0.0, 0.0), 2.0, 1.0), 1.4)
    hits = sorted(orc.intersect(e, (-3.0, 0.0), (1.0, 0.0)))
    assert [h[1][0] for h in hits] == pytest.approx([-2.0, 2.0])
    hits = sorted(orc.intersect(e, (0.0, -3.0), (0.0, 1.0)))
    assert [h[1][1] for h in hits] == pytest.approx([-1.0, 1.0])
    sq = G.ConvexPolygon.new_convex_hull([(1, 1), (-1, 1), (-1, -1), (1, -1), (0, 0)])
    assert len(sq.vertices) == 4
    p = obj(sq, 1.5)
    hits = sorted(orc.intersect(p, (-2.0, 0.25), (1.0, 0.0)))
    assert [h[1][0] for h in hits] == pytest.approx([-1.0, 1.0])
    assert [h[2] for h in hits] == [pytest.approx((-1.0, 0.0)), pytest.approx((1.0, 0.0))]
    assert orc.contains(p, (0.5, 0.5)) and not orc.contains(p, (1.5, 0.0))


def test_lens_known_answers(orc):
    # Lens::new(origin, radius, distance): circles of radius 1 at (+-0.5, 0); their
    # intersection spans x in [-0.5, 0.5] on the axis
    ln = lg.Object.new_lens((0.0, 0.0), 1.0, 1.0).to_abi()
    hits = sorted(orc.intersect(ln, (-2.0, 0.0), (1.0, 0.0)))
    assert [h[1][0] for h in hits] == pytest.approx([-0.5, 0.5])
    # the entry point lies on the circle centred at +0.5, normal points away from that centre
    by_x = sorted(hits, key=lambda h: h[1][0])
    assert by_x[0][2] == pytest.approx((-1.0, 0.0))
    assert orc.contains(ln, (0.0, 0.5)) and not orc.contains(ln, (0.0, 0.9))


def test_bezier_straight_and_arc(orc):
    # collinear control points: the curve is the segment (0,-1)-(0,1)
    b = obj(G.CubicBezier([(0.0, -1.0), (0.0, -1 / 3), (0.0, 1 / 3), (0.0, 1.0)]))
    hits = orc.intersect(b, (-1.0, 0.25), (1.0, 0.0))
    assert len(hits) == 1 and hits[0][1] == pytest.approx((0.0, 0.25), abs=1e-12)
    # the 60-degree arc approximation stays within 2.5e-5 * r of the circle of radius 0.8 (the
    # standard handle length deviates by about 2.1e-5 r at this angle)
    from light_garden_amd.configs import arc_bezier
    arc = obj(arc_bezier((0.0, 0.2), 0.8, math.radians(-30), math.radians(30)))
    for ang in np.linspace(-0.45, 0.45, 21):
        d = unit(ang)
        hits = orc.intersect(arc, (0.0, 0.2), d)
        assert len(hits) == 1
        t, p, n = hits[0]
        assert abs(t - 0.8) < 2.5e-5 * 0.8
        # normal is radial (up to sign)
        assert abs(abs(n[0] * d[0] + n[1] * d[1]) - 1.0) < 1e-4


# ---- brute force ------------------------------------------------------------------------------
def crossings_brute(contains_fn, o, d, t_max=4.0, n=200001):
    """Boundary crossings along the ray from dense inside/outside sampling."""
    ts = np.linspace(0.0, t_max, n)
    inside = np.array([contains_fn((o[0] + t * d[0], o[1] + t * d[1])) for t in ts])
    idx = np.nonzero(inside[1:] != inside[:-1])[0]
    return [0.5 * (ts[i] + ts[i + 1]) for i in idx]


def random_closed_shapes(rng):
    shapes = []
    for _ in range(3):
        shapes.append(G.Circle((rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5)), rng.uniform(0.1, 0.6)))
        shapes.append(G.Rect((rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5)), G.rotation(rng.uniform(0, math.pi)),
                             rng.uniform(0.1, 1.0), rng.uniform(0.1, 1.0)))
        shapes.append(G.Ellipse((rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5)), rng.uniform(0.1, 0.6),
                                rng.uniform(0.1, 0.6), G.rotation(rng.uniform(0, math.pi))))
        pts = [(rng.uniform(-0.6, 0.6), rng.uniform(-0.6, 0.6)) for _ in range(7)]
        shapes.append(G.ConvexPolygon.new_convex_hull(pts))
        a = G.Circle((rng.uniform(-0.3, 0.3), 0.0), rng.uniform(0.3, 0.6))
        b = G.Rect((rng.uniform(-0.3, 0.3), 0.1), G.rotation(rng.uniform(0, 3)), 0.8, 0.5)
        for op in (abi.LOGIC_AND, abi.LOGIC_OR, abi.LOGIC_AND_NOT):
            shapes.append(G.Logic(op, a, b, (rng.uniform(-0.2, 0.2), rng.uniform(-0.2, 0.2)),
                                  G.rotation(rng.uniform(0, math.pi))))
    return shapes


def test_closed_shapes_vs_brute_force(orc):
    """Every closed shape: the hits equal the inside/outside transitions found by dense
    sampling with contains(), lie on the ray, and carry outward unit normals."""
    rng = np.random.default_rng(11)
    for g in random_closed_shapes(rng):
        o_abi = obj(g, 1.5)
        for _ in range(6):
            o = (rng.uniform(-1.5, 1.5), rng.uniform(-1.5, 1.5))
            d = unit(rng.uniform(0, 2 * math.pi))
            hits = orc.intersect(o_abi, o, d)
            brute = crossings_brute(lambda p: orc.contains(o_abi, p), o, d, n=20001)
            ts = sorted(h[0] for h in hits)
            assert len(ts) == len(brute), (g, o, d, ts, brute)
            for t, b in zip(ts, brute):
                assert abs(t - b) < 4.0 / 20000
            for t, p, n in hits:
                assert p == pytest.approx((o[0] + t * d[0], o[1] + t * d[1]), abs=1e-12)
                assert math.hypot(*n) == pytest.approx(1.0, abs=1e-12)
                # outward: a small step along +n leaves the shape, along -n enters it
                eps = 1e-6
                out_p = (p[0] + eps * n[0], p[1] + eps * n[1])
                in_p = (p[0] - eps * n[0], p[1] - eps * n[1])
                assert not orc.contains(o_abi, out_p) and orc.contains(o_abi, in_p)


def test_bezier_vs_brute_force(orc):
    rng = np.random.default_rng(12)
    for _ in range(30):
        P = [(rng.uniform(-1, 1), rng.uniform(-1, 1)) for _ in range(4)]
        b = obj(G.CubicBezier(P))
        o = (rng.uniform(-1.5, 1.5), rng.uniform(-1.5, 1.5))
        d = unit(rng.uniform(0, 2 * math.pi))
        hits = orc.intersect(b, o, d)
        # brute: sign changes of the side of B(s) w.r.t. the ray's line, s in [0,1]
        s = np.linspace(0.0, 1.0, 200001)
        Pa = np.array(P)
        B = ((1 - s) ** 3)[:, None] * Pa[0] + (3 * (1 - s) ** 2 * s)[:, None] * Pa[1] + \
            (3 * (1 - s) * s ** 2)[:, None] * Pa[2] + (s ** 3)[:, None] * Pa[3]
        side = (B[:, 0] - o[0]) * (-d[1]) + (B[:, 1] - o[1]) * d[0]
        idx = np.nonzero(np.sign(side[1:]) != np.sign(side[:-1]))[0]
        cand = []
        for i in idx:
            pt = B[i]
            t = (pt[0] - o[0]) * d[0] + (pt[1] - o[1]) * d[1]
            if t > 1e-3:
                cand.append(t)
        got = sorted(h[0] for h in hits if h[0] > 1e-3)
        assert len(got) == len(cand), (P, o, d, got, cand)
        for g, c in zip(got, sorted(cand)):
            assert abs(g - c) < 1e-4
        for t, p, n in hits:
            # hit point is on the ray and on the curve (distance to the sampled curve)
            assert p == pytest.approx((o[0] + t * d[0], o[1] + t * d[1]), abs=1e-9)
            assert np.min(np.hypot(B[:, 0] - p[0], B[:, 1] - p[1])) < 1e-4


def test_canvas_exit(orc):
    cv = (-1.0, 1.0, -1.0, 1.0)
    assert orc.canvas_exit(cv, (0.0, 0.0), (1.0, 0.0)) == pytest.approx((1.0, 0.0))
    d = unit(math.radians(30))
    p = orc.canvas_exit(cv, (0.0, 0.0), d)
    assert p[0] == pytest.approx(1.0) and p[1] == pytest.approx(math.tan(math.radians(30)))
    # from outside, pointing in: the nearest boundary point
    assert orc.canvas_exit(cv, (-2.0, 0.0), (1.0, 0.0)) == pytest.approx((-1.0, 0.0))
    assert orc.canvas_exit(cv, (-2.0, 0.0), (-1.0, 0.0)) is None


# ---- optics ----------------------------------------------------------------------------------
def fresnel_textbook(ti, n1, n2):
    """Unpolarised Fresnel reflectance from the angle form (Born & Wolf):
    Rs = sin^2(ti - tt) / sin^2(ti + tt), Rp = tan^2(ti - tt) / tan^2(ti + tt)."""
    st = n1 / n2 * math.sin(ti)
    if st >= 1.0:
        return 1.0
    tt = math.asin(st)
    if ti == 0.0:
        return ((n1 - n2) / (n1 + n2)) ** 2
    rs = math.sin(ti - tt) ** 2 / math.sin(ti + tt) ** 2
    rp = math.tan(ti - tt) ** 2 / math.tan(ti + tt) ** 2
    return 0.5 * (rs + rp)


@pytest.mark.parametrize("n1,n2", [(1.0, 1.5), (1.5, 1.0), (1.33, 1.8), (1.8, 1.33), (1.0, 1.0)])
def test_refract_snell_fresnel(orc, n1, n2):
    n = (0.0, 1.0)
    for deg in [0.0, 5.0, 20.0, 35.0, 41.0, 42.0, 48.0, 60.0, 80.0, 89.0]:
        ti = math.radians(deg)
        d = (math.sin(ti), -math.cos(ti))  # coming down onto the surface y = 0
        for nn in (n, (0.0, -1.0)):        # either normal orientation
            refl, refr, R = orc.refract(d, nn, n1, n2)
            assert refl == pytest.approx((d[0], -d[1]), abs=1e-12)       # reflection law
            tir = n1 / n2 * math.sin(ti) > 1.0
            if tir:
                assert refr is None and R == 1.0
                continue
            assert refr is not None
            assert math.hypot(*refr) == pytest.approx(1.0, abs=1e-12)
            assert n1 * math.sin(ti) == pytest.approx(n2 * abs(refr[0]), abs=1e-12)  # Snell
            assert refr[1] < 0 or deg == 0 and refr[1] <= 0                            # transmitted
            assert R == pytest.approx(fresnel_textbook(ti, n1, n2), abs=1e-12)


def test_refract_normal_incidence_and_grazing(orc):
    refl, refr, R = orc.refract((0.0, -1.0), (0.0, 1.0), 1.0, 1.5)
    assert R == pytest.approx(0.04)
    assert refr == pytest.approx((0.0, -1.0))
    # grazing incidence reflects (almost) everything
    ti = math.radians(89.999)
    _, _, R = orc.refract((math.sin(ti), -math.cos(ti)), (0.0, 1.0), 1.0, 1.5)
    assert R > 0.999


def test_reflect(orc):
    rng = np.random.default_rng(3)
    for _ in range(50):
        d = unit(rng.uniform(0, 2 * math.pi))
        n = unit(rng.uniform(0, 2 * math.pi))
        r = orc.reflect(d, n)
        assert math.hypot(*r) == pytest.approx(1.0)
        dn = d[0] * n[0] + d[1] * n[1]
        rn = r[0] * n[0] + r[1] * n[1]
        assert rn == pytest.approx(-dn, abs=1e-12)
        # tangential component kept
        assert (d[0] * -n[1] + d[1] * n[0]) == pytest.approx(r[0] * -n[1] + r[1] * n[0], abs=1e-12)

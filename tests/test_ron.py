"""RON scene files: the reference's Tracer::load / Tracer::serialize format
(src/light_garden/tracer.rs:183-204), through lg_ron_parse_scene / lg_ron_write_scene.

Pinned by the reference's own scene file (tests/golden/default.ron, the reference's
default.ron): its values as read off the text, and a byte-exact write-back, which fixes the
pretty layout (indent, trailing commas, tuple vs sequence brackets) and the float form
(Rust's shortest Display). Host only: no device calls."""
import ctypes as C
import math
import os

import pytest

import light_garden_amd as lg
from light_garden_amd import abi
from light_garden_amd import geo as G
from light_garden_amd import tracer as T

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_RON = os.path.join(HERE, "golden", "default.ron")


@pytest.fixture(scope="module")
def default_text(lglib):
    with open(DEFAULT_RON) as f:
        return f.read()


def test_default_ron_values(default_text):
    objs, lights = T.parse_ron(default_text)
    assert len(objs) == 3 and len(lights) == 2
    lens, curved, rect = objs
    g = lens.geometry
    assert isinstance(g, G.Logic) and g.op == abi.LOGIC_AND
    assert g.a == G.Circle((1.9, 0.0), 2.0) and g.b == G.Circle((-1.9, 0.0), 2.0)
    assert g.origin == (-0.022772240638732733, -0.09999999999999998)
    assert g.rotation == (0.00000000000000006123233995736766, -1.0, 1.0, 0.00000000000000006123233995736766)
    assert lens.material_opt.refractive_index == 1.05
    assert curved.geometry == G.CubicBezier([(-0.622772240638733, 0.40000000000000013),
                                             (-0.3227722406387328, 0.8),
                                             (0.2772277593612673, 0.8),
                                             (0.5772277593612682, 0.40000000000000013)])
    assert curved.material_opt is None
    assert rect.geometry == G.Rect((-0.022772240638732733, -0.5), (1.0, 0.0, 0.0, 1.0),
                                   0.40000000000000036, 0.3999999999999999)
    assert rect.material_opt.refractive_index == 1.73
    p, s = lights
    assert isinstance(p, lg.PointLight) and isinstance(s, lg.SpotLight)
    assert p.position == (-0.022772240638732733, -0.5) and p.num_rays == 10000
    f32 = lambda v: float(C.c_float(v).value)
    assert p.color == [f32(0.009721218), f32(0.009721218), f32(0.009721218), f32(0.011764706)]
    assert s.position == (1.0772277593612674, -0.09999999999999998) and s.num_rays == 2000
    assert s.color == [f32(0.0036765062), f32(0.020288562), f32(0.016807375), f32(0.03529412)]
    assert s.spot_angle == 0.17453292519943295 == math.radians(10.0)
    assert s.spot_direction == (-0.9999922358557027, 0.003940587305547067)


def test_default_ron_writes_back_byte_exact(default_text):
    objs, lights = T.parse_ron(default_text)
    assert T.write_ron(objs, lights) == default_text


def zoo():
    objs = [
        lg.Object.new_mirror((-0.9, 0.85), (0.9, 0.8)),
        lg.Object.new_curved_mirror(G.CubicBezier([(0.1, 0.2), (0.3, 0.45), (0.5, 0.4), (0.7, 0.1)])),
        lg.Object.new_circle((0.25, -0.125), 0.3),
        lg.Object(G.Rect((0.1, 0.2), G.rotation(0.7), 0.4, 0.25), lg.Material(1.41)),
        lg.Object.new_lens((0.3, -0.2), 0.9, 1.6),
        lg.Object.new_convex_polygon([(0.0, 0.0), (0.5, 0.1), (0.4, 0.6), (-0.1, 0.3)]),
        lg.Object(G.Ellipse((-0.3, 0.4), 0.3, 0.1, G.rotation(1.1)), lg.Material(1.9)),
        # boolean trees beyond Lens: Geo(GeoLogic(...)), a nested Logic, a polygon operand
        lg.Object.new_geo(G.Logic(abi.LOGIC_OR, G.Rect((0.0, 0.0), G.IDENTITY, 0.2, 0.1),
                                  G.Ellipse((0.05, 0.0), 0.1, 0.2, G.rotation(0.3)), (0.4, 0.4), G.rotation(2.0))),
        lg.Object.new_geo(G.Logic(abi.LOGIC_AND_NOT,
                                  G.Logic(abi.LOGIC_AND, G.Circle((0.1, 0.0), 0.3), G.Circle((-0.1, 0.0), 0.3)),
                                  G.ConvexPolygon([(-0.05, -0.05), (0.05, -0.05), (0.0, 0.1)]),
                                  (-0.5, -0.5), G.rotation(-0.4))),
        lg.Object.new_geo(G.LineSegment((0.0, -0.9), (0.3, -0.95))),
    ]
    objs[2].material_opt = None  # a mirrored circle: material_opt None
    lights = [lg.PointLight((0.01, 0.02), 4000, [0.02, 0.02, 0.02, 0.02]),
              lg.SpotLight((-0.95, -0.95), 0.6, (0.7071067811865476, 0.7071067811865476), 3000,
                           [0.03, 0.01, 0.02, 0.5]),
              lg.DirectionalLight([0.01, 0.03, 0.02, 0.02], 3000, G.LineSegment((0.9, 0.95), (-0.9, 0.95)))]
    return objs, lights


def abi_bytes(objs, lights):
    oa, la = T.objects_array(objs), T.lights_array(lights)
    return (bytes(oa)[:len(objs) * C.sizeof(abi.LgObject)], bytes(la)[:len(lights) * C.sizeof(abi.LgLight)])


def test_every_kind_round_trips(lglib):
    objs, lights = zoo()
    text = T.write_ron(objs, lights)
    objs2, lights2 = T.parse_ron(text)
    assert len(objs2) == len(objs) and len(lights2) == len(lights)
    # the lg_object/lg_light encodings agree exactly (vertices of polygons with the identity
    # frame are stored as written)
    assert abi_bytes(objs2, lights2) == abi_bytes(objs, lights)
    assert T.write_ron(objs2, lights2) == text
    # variants chosen by shape: Lens for the lens, Geo(GeoLogic) for other trees
    assert text.count("object_enum: Lens((") == 1
    assert text.count("object_enum: Geo(GeoLogic((") == 2
    assert "object_enum: Geo(GeoLineSegment((" not in text  # a bare segment is a StraightMirror
    assert text.count("object_enum: StraightMirror((") == 2
    assert "DirectionalLight((" in text and "start: (" in text


def test_polygon_frame_is_baked_into_points(lglib):
    poly = G.ConvexPolygon([(0.0, 0.0), (1.0, 0.0), (0.0, 1.0)], (0.5, -0.25), G.rotation(math.pi / 2))
    objs2, _ = T.parse_ron(T.write_ron([lg.Object.new_geo(poly)], []))
    g = objs2[0].geometry
    assert g.origin == (0.0, 0.0) and g.rotation == G.IDENTITY
    want = [(0.5, -0.25), (0.5, 0.75), (-0.5, -0.25)]
    for (x, y), (wx, wy) in zip(g.vertices, want):
        assert abs(x - wx) < 1e-15 and abs(y - wy) < 1e-15


@pytest.mark.parametrize("value,text", [
    (2.0, "2"), (-1.0, "-1"), (0.0, "0"), (-0.0, "-0"), (0.1 + 0.2, "0.30000000000000004"),
    (1e21, "1000000000000000000000"), (1e-7, "0.0000001"), (123456.789, "123456.789"),
    (5e-324, "0." + "0" * 323 + "5"), (1.7976931348623157e308, "17976931348623157" + "0" * 292),
])
def test_float_text_is_rust_display(lglib, value, text):
    """Rust's Display for f64: shortest round-trip digits, always positional."""
    out = T.write_ron([lg.Object.new_circle((value, 0.0), 1.0)], [])
    assert "\n                %s,\n" % text in out
    objs, _ = T.parse_ron(out)
    got = objs[0].geometry.origin[0]
    assert got == value and math.copysign(1.0, got) == math.copysign(1.0, value)


def test_compact_text_and_comments(lglib):
    """ron::from_str accepts any whitespace, comments, exponents and optional trailing
    commas; the compact form reads the same scene as the pretty one."""
    text = ("// a scene\n([(object_enum:Circle((origin:[1e-1,-2.5E0],radius:0.5)),"
            "material_opt:Some((refractive_index:1.5)),moved:true)],"
            "[PointLight((position:[0,0],color:(1,0.5,0.25,1),num_rays:7)) /* light */])")
    objs, lights = T.parse_ron(text)
    assert objs[0].geometry == G.Circle((0.1, -2.5), 0.5)
    assert objs[0].material_opt.refractive_index == 1.5
    assert lights[0].num_rays == 7 and lights[0].color == [1.0, 0.5, 0.25, 1.0]


@pytest.mark.parametrize("bad", [
    "", "(", "([], [])x", "([])", "([], [], [])",
    "([(object_enum: Blob((origin: [0, 0], radius: 1)), material_opt: None, moved: false)], [])",
    "([(object_enum: Circle((origin: [0], radius: 1)), material_opt: None, moved: false)], [])",
    "([(object_enum: Circle((origin: [0, 0], radius: x)), material_opt: None, moved: false)], [])",
    "([(object_enum: Circle((origin: [0, 0], radius: 1)), material_opt: Maybe, moved: false)], [])",
    "([(object_enum: Lens((l: (op: Xor, a: GeoCircle((origin: [0, 0], radius: 1)), "
    "b: GeoCircle((origin: [0, 0], radius: 1)), origin: [0, 0], rotation: [1, 0, 0, 1]))), "
    "material_opt: None, moved: false)], [])",
    "([], [PointLight((position: [0, 0], color: (1, 1, 1), num_rays: 5))])",
    "([], [PointLight((position: [0, 0], color: (1, 1, 1, 1), num_rays: -5))])",
    "([], [PointLight((position: [0, 0], color: (1, 1, 1, 1), num_rays: 2.5))])",
    "([], [LaserLight((position: [0, 0]))])",
    "(" * 200 + ")" * 200,
])
def test_malformed_text_is_invalid(lglib, bad):
    with pytest.raises(abi.LgError) as e:
        T.parse_ron(bad)
    assert e.value.code == abi.LG_E_INVALID


def test_too_deep_logic_tree_is_invalid(lglib):
    c = "GeoCircle((origin: [0, 0], radius: 1))"
    t = c
    for _ in range(4):  # 9 nodes > LG_MAX_SHAPES
        t = "GeoLogic((op: Or, a: %s, b: %s, origin: [0, 0], rotation: [1, 0, 0, 1]))" % (t, c)
    with pytest.raises(abi.LgError) as e:
        T.parse_ron("([(object_enum: Geo(%s), material_opt: None, moved: false)], [])" % t)
    assert e.value.code == abi.LG_E_INVALID


def test_capacity_reports_counts(lglib, default_text):
    data = default_text.encode()
    no, nl = C.c_uint32(0), C.c_uint32(0)
    one = (abi.LgObject * 1)()
    rc = lglib.lg_ron_parse_scene(data, len(data), one, 1, C.byref(no), None, 0, C.byref(nl))
    assert rc == abi.LG_E_CAPACITY and (no.value, nl.value) == (3, 2)
    objs, lights = T.parse_ron(default_text)
    oa, la = T.objects_array(objs), T.lights_array(lights)
    n = C.c_uint64(0)
    small = C.create_string_buffer(16)
    rc = lglib.lg_ron_write_scene(oa, 3, la, 2, small, 16, C.byref(n))
    assert rc == abi.LG_E_CAPACITY and n.value == len(default_text)

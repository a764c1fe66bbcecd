"""Geometry values of the reference's scene API (collision2d::Geo as used by
src/light_garden/object.rs:6-120) and their lg_shape encoding.

Rotations are 2x2 matrices stored column-major [m00, m10, m01, m11], the layout of the RON
scene files (default.ron:24-29); rotation by angle a is [cos a, sin a, -sin a, cos a].
"""
import math
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

from . import abi

Point = Tuple[float, float]
IDENTITY = (1.0, 0.0, 0.0, 1.0)


def _scale_pt(p, c, f):
    return (c[0] + (p[0] - c[0]) * f, c[1] + (p[1] - c[1]) * f)


def rotation(angle: float) -> Tuple[float, float, float, float]:
    c, s = math.cos(angle), math.sin(angle)
    return (c, s, -s, c)


@dataclass
class LineSegment:
    """LineSegment::from_ab(a, b)."""
    a: Point
    b: Point

    def get_origin(self):
        return self.a

    def translated(self, dx, dy):
        return LineSegment((self.a[0] + dx, self.a[1] + dy), (self.b[0] + dx, self.b[1] + dy))

    def scaled(self, f, about=None):
        """Scaled by f about `about` (default: the midpoint)."""
        c = about or ((self.a[0] + self.b[0]) * 0.5, (self.a[1] + self.b[1]) * 0.5)
        return LineSegment(_scale_pt(self.a, c, f), _scale_pt(self.b, c, f))


@dataclass
class CubicBezier:
    """CubicBezier{points: [P2; 4]} (the reference's curved mirror, object.rs:345-385)."""
    points: Sequence[Point]

    def get_origin(self):
        return self.points[0]

    def get_control_lines(self):
        """CurvedMirror::get_control_lines (object.rs:355-365): the red control polygon."""
        p = self.points
        return [(p[0], p[1]), (p[1], p[2]), (p[2], p[3])]

    def translated(self, dx, dy):
        return CubicBezier([(x + dx, y + dy) for x, y in self.points])

    def scaled(self, f, about=None):
        """Scaled by f about `about` (default: the centroid of the control points)."""
        n = len(self.points)
        c = about or (sum(p[0] for p in self.points) / n, sum(p[1] for p in self.points) / n)
        return CubicBezier([_scale_pt(p, c, f) for p in self.points])


@dataclass
class Circle:
    origin: Point
    radius: float

    def get_origin(self):
        return self.origin

    def translated(self, dx, dy):
        return Circle((self.origin[0] + dx, self.origin[1] + dy), self.radius)

    def scaled(self, f, about=None):
        return Circle(_scale_pt(self.origin, about or self.origin, f), self.radius * f)


@dataclass
class Rect:
    """Rect::new(origin = centre, rotation, width, height)."""
    origin: Point
    rotation: Tuple[float, float, float, float]
    width: float
    height: float

    def get_origin(self):
        return self.origin

    def translated(self, dx, dy):
        return Rect((self.origin[0] + dx, self.origin[1] + dy), self.rotation, self.width, self.height)

    def scaled(self, f, about=None):
        return Rect(_scale_pt(self.origin, about or self.origin, f), self.rotation, self.width * f, self.height * f)


@dataclass
class Ellipse:
    origin: Point
    a: float
    b: float
    rotation: Tuple[float, float, float, float] = IDENTITY

    def get_origin(self):
        return self.origin

    def translated(self, dx, dy):
        return Ellipse((self.origin[0] + dx, self.origin[1] + dy), self.a, self.b, self.rotation)

    def scaled(self, f, about=None):
        return Ellipse(_scale_pt(self.origin, about or self.origin, f), self.a * f, self.b * f, self.rotation)


def convex_hull(points: Sequence[Point]) -> List[Point]:
    """Counter-clockwise convex hull (Andrew's monotone chain), no collinear points; the
    ConvexPolygon::new_convex_hull constructor (object.rs:35)."""
    pts = sorted(set((float(x), float(y)) for x, y in points))
    if len(pts) <= 2:
        return pts

    def cross(o, a, b):
        return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])

    lower, upper = [], []
    for p in pts:
        while len(lower) >= 2 and cross(lower[-2], lower[-1], p) <= 0:
            lower.pop()
        lower.append(p)
    for p in reversed(pts):
        while len(upper) >= 2 and cross(upper[-2], upper[-1], p) <= 0:
            upper.pop()
        upper.append(p)
    return lower[:-1] + upper[:-1]


@dataclass
class ConvexPolygon:
    """Vertices in the polygon's local frame (counter-clockwise), placed by origin/rotation."""
    vertices: Sequence[Point]
    origin: Point = (0.0, 0.0)
    rotation: Tuple[float, float, float, float] = IDENTITY

    @staticmethod
    def new_convex_hull(points: Sequence[Point]) -> "ConvexPolygon":
        return ConvexPolygon(convex_hull(points))

    def get_origin(self):
        return self.origin

    def translated(self, dx, dy):
        return ConvexPolygon(self.vertices, (self.origin[0] + dx, self.origin[1] + dy), self.rotation)

    def scaled(self, f, about=None):
        return ConvexPolygon([(x * f, y * f) for x, y in self.vertices],
                             _scale_pt(self.origin, about or self.origin, f), self.rotation)


@dataclass
class Logic:
    """Logic{op, a, b, origin, rotation}: a boolean combination of two shapes placed in a
    local frame (Lens::new builds And of two circles, object.rs:393-411)."""
    op: int
    a: object
    b: object
    origin: Point = (0.0, 0.0)
    rotation: Tuple[float, float, float, float] = IDENTITY

    def get_origin(self):
        return self.origin

    def translated(self, dx, dy):
        return Logic(self.op, self.a, self.b, (self.origin[0] + dx, self.origin[1] + dy), self.rotation)

    def scaled(self, f, about=None):
        """Both operands scaled by f about the local frame's origin, the frame placed as before
        (scaled about `about` when given)."""
        return Logic(self.op, self.a.scaled(f, (0.0, 0.0)), self.b.scaled(f, (0.0, 0.0)),
                     _scale_pt(self.origin, about or self.origin, f), self.rotation)


def lens(origin: Point, radius: float, distance: float) -> Logic:
    """Lens::new (object.rs:393-411): And of two circles at (+-distance/2, 0)."""
    return Logic(abi.LOGIC_AND, Circle((distance * 0.5, 0.0), radius),
                 Circle((-distance * 0.5, 0.0), radius), origin, IDENTITY)


def encode(geo, obj: abi.LgObject):
    """Writes geo's node tree into obj.shapes (root at 0) and polygon vertices into
    obj.vertices; sets obj.n_shapes."""
    nodes: List[abi.LgShape] = []
    verts: List[float] = []

    def add(g) -> int:
        idx = len(nodes)
        s = abi.LgShape()
        nodes.append(s)
        if isinstance(g, LineSegment):
            s.kind = abi.SHAPE_LINE_SEGMENT
            s.p[0], s.p[1], s.p[2], s.p[3] = g.a[0], g.a[1], g.b[0], g.b[1]
        elif isinstance(g, CubicBezier):
            s.kind = abi.SHAPE_CUBIC_BEZIER
            for i, (x, y) in enumerate(g.points):
                s.p[2 * i], s.p[2 * i + 1] = x, y
        elif isinstance(g, Circle):
            s.kind = abi.SHAPE_CIRCLE
            s.origin[0], s.origin[1] = g.origin
            s.rotation[:] = IDENTITY
            s.p[0] = g.radius
        elif isinstance(g, Rect):
            s.kind = abi.SHAPE_RECT
            s.origin[0], s.origin[1] = g.origin
            s.rotation[:] = g.rotation
            s.p[0], s.p[1] = g.width, g.height
        elif isinstance(g, Ellipse):
            s.kind = abi.SHAPE_ELLIPSE
            s.origin[0], s.origin[1] = g.origin
            s.rotation[:] = g.rotation
            s.p[0], s.p[1] = g.a, g.b
        elif isinstance(g, ConvexPolygon):
            s.kind = abi.SHAPE_CONVEX_POLYGON
            s.origin[0], s.origin[1] = g.origin
            s.rotation[:] = g.rotation
            s.vertex_offset = len(verts) // 2
            s.n_vertices = len(g.vertices)
            for x, y in g.vertices:
                verts.extend((x, y))
        elif isinstance(g, Logic):
            s.kind = abi.SHAPE_LOGIC
            s.op = g.op
            s.origin[0], s.origin[1] = g.origin
            s.rotation[:] = g.rotation
            s.child_a = add(g.a)
            s.child_b = add(g.b)
        else:
            raise TypeError("unsupported geometry %r" % (g,))
        return idx

    add(geo)
    if len(nodes) > abi.MAX_SHAPES:
        raise ValueError("shape tree has %d nodes, at most %d" % (len(nodes), abi.MAX_SHAPES))
    if len(verts) > 2 * abi.MAX_POLY_VERTICES:
        raise ValueError("more than %d polygon vertices" % abi.MAX_POLY_VERTICES)
    for i, s in enumerate(nodes):
        obj.shapes[i] = s
    for i, v in enumerate(verts):
        obj.vertices[i] = v
    obj.n_shapes = len(nodes)


def decode(obj: abi.LgObject):
    """The inverse of encode: obj's node tree (root at 0) as geometry values."""
    n = int(obj.n_shapes)
    if not 1 <= n <= abi.MAX_SHAPES:
        raise ValueError("object has %d shape nodes" % n)

    def node(i: int):
        if not 0 <= i < n:
            raise ValueError("shape node %d out of range" % i)
        s = obj.shapes[i]
        origin = (s.origin[0], s.origin[1])
        rot = tuple(s.rotation[k] for k in range(4))
        if s.kind == abi.SHAPE_LINE_SEGMENT:
            return LineSegment((s.p[0], s.p[1]), (s.p[2], s.p[3]))
        if s.kind == abi.SHAPE_CUBIC_BEZIER:
            return CubicBezier([(s.p[2 * k], s.p[2 * k + 1]) for k in range(4)])
        if s.kind == abi.SHAPE_CIRCLE:
            return Circle(origin, s.p[0])
        if s.kind == abi.SHAPE_RECT:
            return Rect(origin, rot, s.p[0], s.p[1])
        if s.kind == abi.SHAPE_ELLIPSE:
            return Ellipse(origin, s.p[0], s.p[1], rot)
        if s.kind == abi.SHAPE_CONVEX_POLYGON:
            o = s.vertex_offset
            verts = [(obj.vertices[2 * (o + k)], obj.vertices[2 * (o + k) + 1]) for k in range(s.n_vertices)]
            return ConvexPolygon(verts, origin, rot)
        if s.kind == abi.SHAPE_LOGIC:
            if s.child_a <= i or s.child_b <= i:
                raise ValueError("Logic node %d has children before it" % i)
            return Logic(int(s.op), node(s.child_a), node(s.child_b), origin, rot)
        raise ValueError("unknown shape kind %d" % s.kind)

    return node(0)

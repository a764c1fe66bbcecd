// lg_ron.hip — the reference's RON scene format: Tracer::serialize / Tracer::load
// (src/light_garden/tracer.rs:183-204), i.e. ron's `(Vec<Object>, Vec<Light>)` as written by
// ron::ser::to_string_pretty and read by ron::from_str. Host code only.
//
// Schema (default.ron and the serde derives of object.rs / light.rs):
//   ([ Object, ... ], [ Light, ... ])
//   Object  = (object_enum: ObjectE, material_opt: Some((refractive_index: f)) | None, moved: bool)
//   ObjectE = StraightMirror((line_segment: LineSegment)) | CurvedMirror((cubic: CubicBezier))
//           | Circle(Circle) | Rect(Rect) | Lens((l: Logic)) | ConvexPolygon(ConvexPolygon)
//           | Ellipse(Ellipse) | Geo(Geo)
//   Geo     = GeoCircle(Circle) | GeoRect(Rect) | GeoLogic(Logic) | GeoLineSegment(LineSegment)
//           | GeoCubicBezier(CubicBezier) | GeoConvexPolygon(ConvexPolygon) | GeoEllipse(Ellipse)
//   Circle = (origin: [x, y], radius: r)            Rect = (origin, rotation: [4], width, height)
//   Logic  = (op: And|Or|AndNot, a: Geo, b: Geo, origin, rotation)
//   CubicBezier = (points: ([x, y], [x, y], [x, y], [x, y]))    Ellipse = (origin, a, b, rot: [4])
//   Light  = PointLight((position, color: (r, g, b, a), num_rays))
//          | SpotLight((position, color, num_rays, spot_angle, spot_direction: [x, y]))
//          | DirectionalLight((color, num_rays, start: LineSegment))
// Rotations are nalgebra's column-major [m00, m10, m01, m11]. The serialisations of
// collision2d's LineSegment and ConvexPolygon are not in any file the reference holds; they
// are defined here as LineSegment = (a: [x, y], b: [x, y]) and ConvexPolygon =
// (points: [[x, y], ...]) (world coordinates, counter-clockwise).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "lg_host.h"

using lg::fail;

namespace {

// ---- a generic RON value ------------------------------------------------------------------
struct Value {
  enum Kind { NUM, IDENT, SEQ, STRUCT, CALL } kind = NUM;
  double num = 0.0;
  std::string name;                                      // IDENT text, CALL name
  std::vector<Value> items;                              // SEQ items, CALL arguments
  std::vector<std::pair<std::string, Value>> fields;     // STRUCT fields in order
  const Value* field(const char* f) const {
    for (const auto& kv : fields)
      if (kv.first == f) return &kv.second;
    return nullptr;
  }
};

class Parser {
 public:
  Parser(const char* s, size_t n) : p_(s), end_(s + n) {}
  bool parse(Value& v) {
    if (!value(v)) return false;
    ws();
    if (p_ != end_) return error("trailing characters");
    return true;
  }
  std::string err;

 private:
  const char* p_;
  const char* end_;
  int depth_ = 0;

  bool error(const char* what) {
    if (err.empty()) err = what;
    return false;
  }
  void ws() {
    for (;;) {
      while (p_ < end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) p_++;
      if (end_ - p_ >= 2 && p_[0] == '/' && p_[1] == '/') {
        while (p_ < end_ && *p_ != '\n') p_++;
        continue;
      }
      if (end_ - p_ >= 2 && p_[0] == '/' && p_[1] == '*') {
        p_ += 2;
        while (end_ - p_ >= 2 && !(p_[0] == '*' && p_[1] == '/')) p_++;
        p_ = end_ - p_ >= 2 ? p_ + 2 : end_;
        continue;
      }
      return;
    }
  }
  bool peek(char c) {
    ws();
    return p_ < end_ && *p_ == c;
  }
  bool eat(char c) {
    if (!peek(c)) return false;
    p_++;
    return true;
  }
  static bool ident_start(char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || c == '_'; }
  static bool ident_char(char c) { return ident_start(c) || (c >= '0' && c <= '9'); }
  bool ident(std::string& out) {
    ws();
    if (p_ >= end_ || !ident_start(*p_)) return false;
    const char* b = p_;
    while (p_ < end_ && ident_char(*p_)) p_++;
    out.assign(b, p_);
    return true;
  }
  // Whether the next tokens are `ident :` (a struct field), without consuming them.
  bool at_field() {
    ws();
    const char* save = p_;
    std::string id;
    bool r = ident(id) && peek(':');
    p_ = save;
    return r;
  }
  bool number(Value& v) {
    ws();
    const char* b = p_;
    {
      const char* t = p_;
      const bool neg = t < end_ && (*t == '-' || *t == '+');
      if (neg) t++;
      if (end_ - t >= 3 && (std::strncmp(t, "inf", 3) == 0 || std::strncmp(t, "NaN", 3) == 0)) {
        v.kind = Value::NUM;
        v.num = t[0] == 'N' ? NAN : (*p_ == '-' ? -INFINITY : INFINITY);
        p_ = t + 3;
        return true;
      }
    }
    std::string buf;  // positional f64 text runs to ~330 characters (5e-324, f64::MAX)
    while (p_ < end_ && buf.size() < 1024 &&
           ((*p_ >= '0' && *p_ <= '9') || *p_ == '-' || *p_ == '+' || *p_ == '.' || *p_ == 'e' || *p_ == 'E'))
      buf += *p_++;
    if (buf.empty()) return error("expected a value");
    char* e = nullptr;
    v.kind = Value::NUM;
    v.num = std::strtod(buf.c_str(), &e);
    if (e != buf.c_str() + buf.size()) {
      p_ = b;
      return error("malformed number");
    }
    return true;
  }
  // Comma-separated values up to `close` (trailing comma allowed).
  bool seq(char close, std::vector<Value>& out) {
    for (;;) {
      if (eat(close)) return true;
      Value v;
      if (!value(v)) return false;
      out.push_back(std::move(v));
      if (eat(',')) continue;
      if (eat(close)) return true;
      return error("expected ',' or a closing bracket");
    }
  }
  bool fields(std::vector<std::pair<std::string, Value>>& out) {
    for (;;) {
      if (eat(')')) return true;
      std::string name;
      if (!ident(name)) return error("expected a field name");
      if (!eat(':')) return error("expected ':'");
      Value v;
      if (!value(v)) return false;
      out.emplace_back(name, std::move(v));
      if (eat(',')) continue;
      if (eat(')')) return true;
      return error("expected ',' or ')'");
    }
  }
  // `( ... )`: a struct when it starts with `name:`, else a tuple.
  bool paren(Value& v) {
    if (at_field()) {
      v.kind = Value::STRUCT;
      return fields(v.fields);
    }
    v.kind = Value::SEQ;
    return seq(')', v.items);
  }
  bool value(Value& v) {
    if (++depth_ > 64) return error("nesting too deep");
    bool ok = value_inner(v);
    depth_--;
    return ok;
  }
  bool value_inner(Value& v) {
    ws();
    if (p_ >= end_) return error("unexpected end of text");
    if (eat('(')) return paren(v);
    if (eat('[')) {
      v.kind = Value::SEQ;
      return seq(']', v.items);
    }
    std::string id;
    if (ident(id)) {
      if (eat('(')) {  // enum variant or named struct: Name(args) / Name(field: ...)
        v.kind = Value::CALL;
        v.name = id;
        if (at_field()) {
          Value s;
          s.kind = Value::STRUCT;
          if (!fields(s.fields)) return false;
          v.items.push_back(std::move(s));
          return true;
        }
        return seq(')', v.items);
      }
      if (id == "inf" || id == "NaN") {
        v.kind = Value::NUM;
        v.num = id == "inf" ? INFINITY : NAN;
        return true;
      }
      v.kind = Value::IDENT;
      v.name = id;
      return true;
    }
    return number(v);
  }
};

// ---- value -> lg structs ---------------------------------------------------------------------
struct Conv {
  std::string err;
  bool error(const std::string& what) {
    if (err.empty()) err = what;
    return false;
  }
  bool num(const Value* v, double& out, const char* what) {
    if (!v || v->kind != Value::NUM) return error(std::string("expected a number for ") + what);
    out = v->num;
    return true;
  }
  bool field_num(const Value& s, const char* f, double& out) { return num(s.field(f), out, f); }
  bool vec2(const Value* v, double out[2], const char* what) {
    if (!v || v->kind != Value::SEQ || v->items.size() != 2)
      return error(std::string("expected [x, y] for ") + what);
    return num(&v->items[0], out[0], what) && num(&v->items[1], out[1], what);
  }
  bool rot(const Value* v, double out[4], const char* what) {
    if (!v || v->kind != Value::SEQ || v->items.size() != 4)
      return error(std::string("expected [m00, m10, m01, m11] for ") + what);
    for (int i = 0; i < 4; i++)
      if (!num(&v->items[i], out[i], what)) return false;
    return true;
  }
  // The struct payload of `Name((...))` / `Name(...)`.
  const Value* payload(const Value& call) {
    if (call.kind != Value::CALL || call.items.size() != 1) return nullptr;
    const Value* p = &call.items[0];
    if (p->kind == Value::SEQ && p->items.size() == 1) p = &p->items[0];  // tolerate an extra tuple
    return p->kind == Value::STRUCT ? p : nullptr;
  }
  static void identity(double r[4]) {
    r[0] = 1.0;
    r[1] = 0.0;
    r[2] = 0.0;
    r[3] = 1.0;
  }

  bool line_segment(const Value* s, lg_shape& sh) {
    double a[2], b[2];
    if (!s || s->kind != Value::STRUCT) return error("expected a LineSegment (a: [x, y], b: [x, y])");
    if (!vec2(s->field("a"), a, "LineSegment.a") || !vec2(s->field("b"), b, "LineSegment.b")) return false;
    sh.kind = LG_SHAPE_LINE_SEGMENT;
    identity(sh.rotation);
    sh.p[0] = a[0];
    sh.p[1] = a[1];
    sh.p[2] = b[0];
    sh.p[3] = b[1];
    return true;
  }
  bool cubic(const Value* s, lg_shape& sh) {
    if (!s || s->kind != Value::STRUCT) return error("expected a CubicBezier (points: (...))");
    const Value* pts = s->field("points");
    if (!pts || pts->kind != Value::SEQ || pts->items.size() != 4) return error("CubicBezier needs 4 points");
    sh.kind = LG_SHAPE_CUBIC_BEZIER;
    identity(sh.rotation);
    for (int i = 0; i < 4; i++) {
      double q[2];
      if (!vec2(&pts->items[i], q, "CubicBezier.points")) return false;
      sh.p[2 * i] = q[0];
      sh.p[2 * i + 1] = q[1];
    }
    return true;
  }
  bool circle(const Value* s, lg_shape& sh) {
    if (!s) return error("expected a Circle (origin, radius)");
    sh.kind = LG_SHAPE_CIRCLE;
    identity(sh.rotation);
    return vec2(s->field("origin"), sh.origin, "Circle.origin") && field_num(*s, "radius", sh.p[0]);
  }
  bool rect(const Value* s, lg_shape& sh) {
    if (!s) return error("expected a Rect (origin, rotation, width, height)");
    sh.kind = LG_SHAPE_RECT;
    return vec2(s->field("origin"), sh.origin, "Rect.origin") && rot(s->field("rotation"), sh.rotation, "Rect.rotation") &&
           field_num(*s, "width", sh.p[0]) && field_num(*s, "height", sh.p[1]);
  }
  bool ellipse(const Value* s, lg_shape& sh) {
    if (!s) return error("expected an Ellipse (origin, a, b, rot)");
    sh.kind = LG_SHAPE_ELLIPSE;
    return vec2(s->field("origin"), sh.origin, "Ellipse.origin") && field_num(*s, "a", sh.p[0]) &&
           field_num(*s, "b", sh.p[1]) && rot(s->field("rot"), sh.rotation, "Ellipse.rot");
  }
  bool polygon(const Value* s, lg_object& o, lg_shape& sh) {
    if (!s || s->kind != Value::STRUCT) return error("expected a ConvexPolygon (points: [...])");
    const Value* pts = s->field("points");
    if (!pts || pts->kind != Value::SEQ) return error("ConvexPolygon needs points");
    int used = 0;
    for (int i = 0; i < o.n_shapes; i++)
      if (o.shapes[i].kind == LG_SHAPE_CONVEX_POLYGON) used += o.shapes[i].n_vertices;
    if (used + (int)pts->items.size() > LG_MAX_POLY_VERTICES) return error("ConvexPolygon: too many vertices");
    sh.kind = LG_SHAPE_CONVEX_POLYGON;
    identity(sh.rotation);
    sh.vertex_offset = used;
    sh.n_vertices = (int)pts->items.size();
    for (size_t i = 0; i < pts->items.size(); i++) {
      double q[2];
      if (!vec2(&pts->items[i], q, "ConvexPolygon.points")) return false;
      o.vertices[2 * (used + i)] = q[0];
      o.vertices[2 * (used + i) + 1] = q[1];
    }
    return true;
  }
  bool logic(const Value* s, lg_object& o, int node) {
    if (!s || s->kind != Value::STRUCT) return error("expected a Logic (op, a, b, origin, rotation)");
    const Value* op = s->field("op");
    if (!op || op->kind != Value::IDENT) return error("Logic.op must be And, Or or AndNot");
    lg_shape& sh = o.shapes[node];
    sh.kind = LG_SHAPE_LOGIC;
    if (op->name == "And")
      sh.op = LG_LOGIC_AND;
    else if (op->name == "Or")
      sh.op = LG_LOGIC_OR;
    else if (op->name == "AndNot")
      sh.op = LG_LOGIC_AND_NOT;
    else
      return error("Logic.op must be And, Or or AndNot");
    if (!vec2(s->field("origin"), sh.origin, "Logic.origin") || !rot(s->field("rotation"), sh.rotation, "Logic.rotation"))
      return false;
    const Value* a = s->field("a");
    const Value* b = s->field("b");
    if (!a || !b) return error("Logic needs operands a and b");
    int ia = add_geo(*a, o);
    if (ia < 0) return false;
    int ib = add_geo(*b, o);
    if (ib < 0) return false;
    o.shapes[node].child_a = ia;
    o.shapes[node].child_b = ib;
    return true;
  }
  int new_node(lg_object& o) {
    if (o.n_shapes >= LG_MAX_SHAPES) {
      error("shape tree deeper than LG_MAX_SHAPES nodes");
      return -1;
    }
    std::memset(&o.shapes[o.n_shapes], 0, sizeof(lg_shape));
    return o.n_shapes++;
  }
  // A Geo value (GeoCircle(...), ...) as a new node; returns its index or -1.
  int add_geo(const Value& g, lg_object& o) {
    if (g.kind != Value::CALL) {
      error("expected a Geo variant");
      return -1;
    }
    int n = new_node(o);
    if (n < 0) return -1;
    const Value* s = payload(g);
    bool ok;
    if (g.name == "GeoCircle")
      ok = circle(s, o.shapes[n]);
    else if (g.name == "GeoRect")
      ok = rect(s, o.shapes[n]);
    else if (g.name == "GeoEllipse")
      ok = ellipse(s, o.shapes[n]);
    else if (g.name == "GeoLineSegment")
      ok = line_segment(s, o.shapes[n]);
    else if (g.name == "GeoCubicBezier")
      ok = cubic(s, o.shapes[n]);
    else if (g.name == "GeoConvexPolygon")
      ok = polygon(s, o, o.shapes[n]);
    else if (g.name == "GeoLogic")
      ok = logic(s, o, n);
    else
      ok = error("unknown Geo variant " + g.name);
    return ok ? n : -1;
  }

  bool object(const Value& v, lg_object& o) {
    std::memset(&o, 0, sizeof(o));
    if (v.kind != Value::STRUCT) return error("expected an Object (object_enum, material_opt, moved)");
    const Value* e = v.field("object_enum");
    if (!e || e->kind != Value::CALL) return error("Object needs object_enum");
    int n = new_node(o);
    const Value* s = payload(*e);
    bool ok;
    if (e->name == "StraightMirror")
      ok = s && line_segment(s->field("line_segment"), o.shapes[n]);
    else if (e->name == "CurvedMirror")
      ok = s && cubic(s->field("cubic"), o.shapes[n]);
    else if (e->name == "Circle")
      ok = circle(s, o.shapes[n]);
    else if (e->name == "Rect")
      ok = rect(s, o.shapes[n]);
    else if (e->name == "Ellipse")
      ok = ellipse(s, o.shapes[n]);
    else if (e->name == "ConvexPolygon")
      ok = polygon(s, o, o.shapes[n]);
    else if (e->name == "Lens")
      ok = s && logic(s->field("l"), o, n);
    else if (e->name == "Geo") {
      o.n_shapes = 0;
      ok = e->items.size() == 1 && add_geo(e->items[0], o) == 0;
    } else
      ok = error("unknown object_enum variant " + e->name);
    if (!ok) return error("bad " + e->name);
    const Value* m = v.field("material_opt");
    if (!m) return error("Object needs material_opt");
    if (m->kind == Value::IDENT && m->name == "None") {
      o.has_material = 0;
      o.refractive_index = 1.0;
    } else if (m->kind == Value::CALL && m->name == "Some") {
      const Value* ms = payload(*m);
      if (!ms || !field_num(*ms, "refractive_index", o.refractive_index)) return error("bad material");
      o.has_material = 1;
    } else {
      return error("material_opt must be Some((refractive_index: n)) or None");
    }
    return true;
  }

  bool color(const Value* v, float c[4]) {
    if (!v || v->kind != Value::SEQ || v->items.size() != 4) return error("expected color: (r, g, b, a)");
    for (int i = 0; i < 4; i++) {
      double d;
      if (!num(&v->items[i], d, "color")) return false;
      c[i] = (float)d;
    }
    return true;
  }
  bool count(const Value* v, uint64_t& n) {
    double d;
    if (!num(v, d, "num_rays")) return false;
    if (!(d >= 0.0) || d != std::floor(d) || d > 1.8e19) return error("num_rays must be a non-negative integer");
    n = (uint64_t)d;
    return true;
  }
  bool light(const Value& v, lg_light& l) {
    std::memset(&l, 0, sizeof(l));
    const Value* s = payload(v);
    if (v.kind != Value::CALL || !s) return error("expected a Light variant");
    if (v.name == "PointLight") {
      l.kind = LG_LIGHT_POINT;
      return vec2(s->field("position"), l.position, "position") && color(s->field("color"), l.color) &&
             count(s->field("num_rays"), l.num_rays);
    }
    if (v.name == "SpotLight") {
      l.kind = LG_LIGHT_SPOT;
      return vec2(s->field("position"), l.position, "position") && color(s->field("color"), l.color) &&
             count(s->field("num_rays"), l.num_rays) && field_num(*s, "spot_angle", l.spot_angle) &&
             vec2(s->field("spot_direction"), l.direction, "spot_direction");
    }
    if (v.name == "DirectionalLight") {
      l.kind = LG_LIGHT_DIRECTIONAL;
      lg_shape seg{};
      if (!color(s->field("color"), l.color) || !count(s->field("num_rays"), l.num_rays) ||
          !line_segment(s->field("start"), seg))
        return false;
      l.position[0] = seg.p[0];
      l.position[1] = seg.p[1];
      l.direction[0] = seg.p[2];
      l.direction[1] = seg.p[3];
      return true;
    }
    return error("unknown Light variant " + v.name);
  }
};

// ---- lg structs -> RON text (ron's pretty style: 4-space indent, trailing commas) -------------
// Rust's Display for floats, which ron's serializer uses: the shortest digit string that
// reads back to the same value, always positional (no exponent), integral values without a
// decimal point ("2", "-1", "0.00000000000000006123233995736766" as in default.ron).
std::string rust_display(double v, bool single) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
  if (v == 0.0) return std::signbit(v) ? "-0" : "0";
  char buf[48];
  const int max_p = single ? 8 : 16;
  for (int prec = 0; prec <= max_p; prec++) {
    std::snprintf(buf, sizeof(buf), "%.*e", prec, v);
    if (single ? std::strtof(buf, nullptr) == (float)v : std::strtod(buf, nullptr) == v) break;
  }
  // buf = [-]d[.ddd]e(+|-)xx
  std::string digits;
  const char* q = buf;
  const bool neg = *q == '-';
  if (neg) q++;
  for (; *q && *q != 'e'; q++)
    if (*q != '.') digits += *q;
  const int exp10 = std::atoi(q + 1);
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  std::string out = neg ? "-" : "";
  if (exp10 >= 0) {
    const size_t int_len = (size_t)exp10 + 1;
    if (digits.size() <= int_len) {
      out += digits + std::string(int_len - digits.size(), '0');
    } else {
      out += digits.substr(0, int_len) + "." + digits.substr(int_len);
    }
  } else {
    out += "0." + std::string((size_t)(-exp10 - 1), '0') + digits;
  }
  return out;
}

std::string fmt_f64(double v) { return rust_display(v, false); }
std::string fmt_f32(float v) { return rust_display((double)v, true); }

struct Writer {
  std::string s;
  int d = 0;
  void nl() {
    s += '\n';
    s.append((size_t)(4 * d), ' ');
  }
  void open(const char* head) {  // "name: Variant((" or "(" ...
    s += head;
    d++;
  }
  void close(const char* tail) {
    d--;
    nl();
    s += tail;
  }
  void num_field(const char* name, double v) {
    nl();
    s += name;
    s += ": " + fmt_f64(v) + ",";
  }
  void seq_field(const char* name, const double* v, int n) {
    nl();
    s += name;
    s += ": [";
    d++;
    for (int i = 0; i < n; i++) {
      nl();
      s += fmt_f64(v[i]) + ",";
    }
    d--;
    nl();
    s += "],";
  }
};

void world_vertex(const lg_object& o, const lg_shape& sh, int i, double& x, double& y) {
  const double lx = o.vertices[2 * (sh.vertex_offset + i)], ly = o.vertices[2 * (sh.vertex_offset + i) + 1];
  x = sh.origin[0] + sh.rotation[0] * lx + sh.rotation[2] * ly;
  y = sh.origin[1] + sh.rotation[1] * lx + sh.rotation[3] * ly;
}

bool write_struct_body(Writer& w, const lg_object& o, int node, std::string& err);

// A shape node as a Geo variant: `GeoCircle((...))`.
bool write_geo(Writer& w, const char* field, const lg_object& o, int node, std::string& err) {
  const lg_shape& sh = o.shapes[node];
  static const char* names[] = {"GeoLineSegment", "GeoCubicBezier", "GeoCircle", "GeoRect",
                                "GeoLogic", "GeoConvexPolygon", "GeoEllipse"};
  if (sh.kind < 0 || sh.kind > 6) {
    err = "unknown shape kind";
    return false;
  }
  w.nl();
  std::string head = std::string(field) + ": " + names[sh.kind] + "((";
  w.open(head.c_str());
  if (!write_struct_body(w, o, node, err)) return false;
  w.close(")),");
  return true;
}

// The fields of a node's struct (without the surrounding parentheses).
bool write_struct_body(Writer& w, const lg_object& o, int node, std::string& err) {
  const lg_shape& sh = o.shapes[node];
  switch (sh.kind) {
    case LG_SHAPE_LINE_SEGMENT:
      w.seq_field("a", sh.p, 2);
      w.seq_field("b", sh.p + 2, 2);
      return true;
    case LG_SHAPE_CUBIC_BEZIER: {
      w.nl();
      w.s += "points: (";
      for (int i = 0; i < 4; i++) {
        w.s += i ? ", [" : "[";
        w.d++;
        w.nl();
        w.s += fmt_f64(sh.p[2 * i]) + ",";
        w.nl();
        w.s += fmt_f64(sh.p[2 * i + 1]) + ",";
        w.d--;
        w.nl();
        w.s += "]";
      }
      w.s += "),";
      return true;
    }
    case LG_SHAPE_CIRCLE:
      w.seq_field("origin", sh.origin, 2);
      w.num_field("radius", sh.p[0]);
      return true;
    case LG_SHAPE_RECT:
      w.seq_field("origin", sh.origin, 2);
      w.seq_field("rotation", sh.rotation, 4);
      w.num_field("width", sh.p[0]);
      w.num_field("height", sh.p[1]);
      return true;
    case LG_SHAPE_ELLIPSE:
      w.seq_field("origin", sh.origin, 2);
      w.num_field("a", sh.p[0]);
      w.num_field("b", sh.p[1]);
      w.seq_field("rot", sh.rotation, 4);
      return true;
    case LG_SHAPE_CONVEX_POLYGON: {
      w.nl();
      w.s += "points: [";
      w.d++;
      for (int i = 0; i < sh.n_vertices; i++) {
        double xy[2];
        world_vertex(o, sh, i, xy[0], xy[1]);
        w.nl();
        w.s += "[";
        w.d++;
        w.nl();
        w.s += fmt_f64(xy[0]) + ",";
        w.nl();
        w.s += fmt_f64(xy[1]) + ",";
        w.d--;
        w.nl();
        w.s += "],";
      }
      w.d--;
      w.nl();
      w.s += "],";
      return true;
    }
    case LG_SHAPE_LOGIC: {
      static const char* ops[] = {"And", "Or", "AndNot"};
      if (sh.op < 0 || sh.op > 2 || sh.child_a <= node || sh.child_b <= node || sh.child_a >= o.n_shapes ||
          sh.child_b >= o.n_shapes) {
        err = "bad Logic node";
        return false;
      }
      w.nl();
      w.s += std::string("op: ") + ops[sh.op] + ",";
      if (!write_geo(w, "a", o, sh.child_a, err) || !write_geo(w, "b", o, sh.child_b, err)) return false;
      w.seq_field("origin", sh.origin, 2);
      w.seq_field("rotation", sh.rotation, 4);
      return true;
    }
  }
  err = "unknown shape kind";
  return false;
}

// Lens::new's shape (object.rs:393-411): And of two equal circles at (+-d/2, 0).
bool is_lens(const lg_object& o) {
  const lg_shape& r = o.shapes[0];
  if (r.kind != LG_SHAPE_LOGIC || r.op != LG_LOGIC_AND || o.n_shapes != 3 || r.child_a != 1 || r.child_b != 2)
    return false;
  const lg_shape& a = o.shapes[1];
  const lg_shape& b = o.shapes[2];
  return a.kind == LG_SHAPE_CIRCLE && b.kind == LG_SHAPE_CIRCLE && a.p[0] == b.p[0] && a.origin[1] == 0.0 &&
         b.origin[1] == 0.0 && a.origin[0] == -b.origin[0];
}

bool write_object(Writer& w, const lg_object& o, std::string& err) {
  if (o.n_shapes < 1 || o.n_shapes > LG_MAX_SHAPES) {
    err = "object without shapes";
    return false;
  }
  w.nl();
  w.open("(");
  const lg_shape& root = o.shapes[0];
  w.nl();
  switch (root.kind) {
    case LG_SHAPE_LINE_SEGMENT:
      w.open("object_enum: StraightMirror((");
      w.nl();
      w.open("line_segment: (");
      if (!write_struct_body(w, o, 0, err)) return false;
      w.close("),");
      w.close(")),");
      break;
    case LG_SHAPE_CUBIC_BEZIER:
      w.open("object_enum: CurvedMirror((");
      w.nl();
      w.open("cubic: (");
      if (!write_struct_body(w, o, 0, err)) return false;
      w.close("),");
      w.close(")),");
      break;
    case LG_SHAPE_CIRCLE:
    case LG_SHAPE_RECT:
    case LG_SHAPE_ELLIPSE:
    case LG_SHAPE_CONVEX_POLYGON: {
      static const char* names[] = {"", "", "Circle", "Rect", "", "ConvexPolygon", "Ellipse"};
      std::string head = std::string("object_enum: ") + names[root.kind] + "((";
      w.open(head.c_str());
      if (!write_struct_body(w, o, 0, err)) return false;
      w.close(")),");
      break;
    }
    case LG_SHAPE_LOGIC:
      if (is_lens(o)) {
        w.open("object_enum: Lens((");
        w.nl();
        w.open("l: (");
        if (!write_struct_body(w, o, 0, err)) return false;
        w.close("),");
        w.close(")),");
      } else {
        w.open("object_enum: Geo(GeoLogic((");
        if (!write_struct_body(w, o, 0, err)) return false;
        w.close("))),");
      }
      break;
    default:
      err = "unknown shape kind";
      return false;
  }
  w.nl();
  if (o.has_material) {
    w.open("material_opt: Some((");
    w.num_field("refractive_index", o.refractive_index);
    w.close(")),");
  } else {
    w.s += "material_opt: None,";
  }
  w.nl();
  w.s += "moved: false,";
  w.close("),");
  return true;
}

bool write_light(Writer& w, const lg_light& l, std::string& err) {
  static const char* names[] = {"PointLight", "DirectionalLight", "SpotLight"};
  if (l.kind < 0 || l.kind > 2) {
    err = "unknown light kind";
    return false;
  }
  w.nl();
  std::string head = std::string(names[l.kind]) + "((";
  w.open(head.c_str());
  std::string color = "color: (" + fmt_f32(l.color[0]) + ", " + fmt_f32(l.color[1]) + ", " + fmt_f32(l.color[2]) +
                      ", " + fmt_f32(l.color[3]) + "),";
  char nr[32];
  std::snprintf(nr, sizeof(nr), "num_rays: %llu,", (unsigned long long)l.num_rays);
  if (l.kind == LG_LIGHT_DIRECTIONAL) {
    w.nl();
    w.s += color;
    w.nl();
    w.s += nr;
    w.nl();
    w.open("start: (");
    w.seq_field("a", l.position, 2);
    w.seq_field("b", l.direction, 2);
    w.close("),");
  } else {
    w.seq_field("position", l.position, 2);
    w.nl();
    w.s += color;
    w.nl();
    w.s += nr;
    if (l.kind == LG_LIGHT_SPOT) {
      w.num_field("spot_angle", l.spot_angle);
      w.seq_field("spot_direction", l.direction, 2);
    }
  }
  w.close(")),");
  return true;
}

// Text -> objects and lights; on failure `err` says where.
bool parse_scene(const char* text, size_t len, std::vector<lg_object>& objects, std::vector<lg_light>& lights,
                 std::string& err) {
  Value root;
  Parser p(text, len);
  if (!p.parse(root)) {
    err = p.err;
    return false;
  }
  if (root.kind != Value::SEQ || root.items.size() != 2 || root.items[0].kind != Value::SEQ ||
      root.items[1].kind != Value::SEQ) {
    err = "expected ([objects], [lights])";
    return false;
  }
  const auto& objs = root.items[0].items;
  const auto& ls = root.items[1].items;
  Conv c;
  objects.assign(objs.size(), lg_object{});
  lights.assign(ls.size(), lg_light{});
  for (size_t i = 0; i < objs.size(); i++)
    if (!c.object(objs[i], objects[i])) {
      err = "object " + std::to_string(i) + ": " + c.err;
      return false;
    }
  for (size_t i = 0; i < ls.size(); i++)
    if (!c.light(ls[i], lights[i])) {
      err = "light " + std::to_string(i) + ": " + c.err;
      return false;
    }
  return true;
}

bool write_scene(const lg_object* objects, uint32_t n_objects, const lg_light* lights, uint32_t n_lights,
                 std::string& out, std::string& err) {
  Writer w;
  w.s = "([";
  w.d = 1;
  for (uint32_t i = 0; i < n_objects; i++)
    if (!write_object(w, objects[i], err)) {
      err = "object " + std::to_string(i) + ": " + err;
      return false;
    }
  w.d = 0;
  w.nl();
  w.s += "], [";
  w.d = 1;
  for (uint32_t i = 0; i < n_lights; i++)
    if (!write_light(w, lights[i], err)) {
      err = "light " + std::to_string(i) + ": " + err;
      return false;
    }
  w.d = 0;
  w.nl();
  w.s += "])";
  out.swap(w.s);
  return true;
}

int copy_text(const std::string& t, char* buf, uint64_t cap, uint64_t* len) {
  *len = t.size();
  if (!buf || t.size() + 1 > cap) return fail(LG_E_CAPACITY, "ron: buffer too small (length returned)");
  std::memcpy(buf, t.c_str(), t.size() + 1);
  return LG_OK;
}

}  // namespace

extern "C" {

int lg_ron_parse_scene(const char* text, uint64_t len, lg_object* objects, uint32_t cap_objects,
                       uint32_t* n_objects, lg_light* lights, uint32_t cap_lights, uint32_t* n_lights) {
  if (!text || !n_objects || !n_lights) return fail(LG_E_INVALID, "ron: text or counts are NULL");
  std::vector<lg_object> ov;
  std::vector<lg_light> lv;
  std::string err;
  if (!parse_scene(text, (size_t)len, ov, lv, err)) return fail(LG_E_INVALID, "ron: " + err);
  *n_objects = (uint32_t)ov.size();
  *n_lights = (uint32_t)lv.size();
  if (ov.size() > cap_objects || lv.size() > cap_lights || (!ov.empty() && !objects) || (!lv.empty() && !lights))
    return fail(LG_E_CAPACITY, "ron: arrays too small for the scene (counts returned)");
  if (!ov.empty()) std::memcpy(objects, ov.data(), ov.size() * sizeof(lg_object));
  if (!lv.empty()) std::memcpy(lights, lv.data(), lv.size() * sizeof(lg_light));
  return LG_OK;
}

int lg_ron_write_scene(const lg_object* objects, uint32_t n_objects, const lg_light* lights, uint32_t n_lights,
                       char* buf, uint64_t cap, uint64_t* len) {
  if (!len || (n_objects && !objects) || (n_lights && !lights)) return fail(LG_E_INVALID, "ron: NULL argument");
  std::string text, err;
  if (!write_scene(objects, n_objects, lights, n_lights, text, err)) return fail(LG_E_INVALID, "ron: " + err);
  return copy_text(text, buf, cap, len);
}

int lg_scene_load_ron(lg_scene* scene, const char* text, uint64_t len) {
  if (!scene || !text) return fail(LG_E_INVALID, "scene or text is NULL");
  std::vector<lg_object> ov;
  std::vector<lg_light> lv;
  std::string err;
  if (!parse_scene(text, (size_t)len, ov, lv, err)) return fail(LG_E_INVALID, "ron: " + err);
  scene->objects.swap(ov);
  scene->lights.swap(lv);
  scene->dirty = true;
  return LG_OK;
}

int lg_scene_save_ron(const lg_scene* scene, char* buf, uint64_t cap, uint64_t* len) {
  if (!scene || !len) return fail(LG_E_INVALID, "scene or len is NULL");
  std::string text, err;
  if (!write_scene(scene->objects.data(), (uint32_t)scene->objects.size(), scene->lights.data(),
                   (uint32_t)scene->lights.size(), text, err))
    return fail(LG_E_INVALID, "ron: " + err);
  return copy_text(text, buf, cap, len);
}

}  // extern "C"

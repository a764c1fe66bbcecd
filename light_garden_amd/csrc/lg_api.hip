// lg_api.hip — the C ABI of include/lg.h: contexts, scenes, buffers and the host side of
// trace / accumulate / resolve / render.
//
// The scene API mirrors Tracer (src/light_garden/tracer.rs:4-17, 48-116). A scene keeps
// the caller's lg_object / lg_light values on the host (FP64, as the reference does) and
// compiles them into world-space device records on the first trace after an edit.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "lg_internal.h"

namespace lg {

// ---- launchers defined in lg_trace.hip / lg_accum.hip ------------------------------------
template <typename T>
hipError_t launch_trace_gen(const TraceArgs<T>& args, int grid, hipStream_t stream);
template <typename T>
size_t trace_lds_bytes(int n_recs, int n_objs, int n_verts);
int tile_size();
hipError_t launch_bin_count(const float4* pos, uint64_t n, const AccumView& v, uint32_t chunk,
                            uint32_t n_chunks, uint32_t* counts, hipStream_t st);
hipError_t launch_bin_fill(const float4* pos, uint64_t n, const AccumView& v, uint32_t chunk,
                           uint32_t n_chunks, const uint32_t* offsets, uint32_t* pairs,
                           hipStream_t st);
hipError_t launch_tile_starts(const uint32_t* offsets, uint32_t n_chunks, int n_tiles,
                              const unsigned long long* total64, uint32_t* tile_start,
                              hipStream_t st);
hipError_t launch_sum_u64(const uint32_t* in, uint64_t n, unsigned long long* total64,
                          hipStream_t st);
uint64_t scan_scratch_elems(uint64_t n);
hipError_t launch_scan_exclusive(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* scratch,
                                 hipStream_t st);
hipError_t launch_tile_accum(const float4* pos, const float4* col0, const float4* col1,
                             const uint32_t* pairs, const WorkItem* items, uint32_t n_items,
                             uint32_t* item_counter, const AccumView& v, float4* accum, int grid,
                             int fixed_bits, hipStream_t st);
int tile_block();
int weight_bits();
hipError_t launch_resolve(const float4* accum, ushort4* out, uint64_t n, hipStream_t st);
hipError_t launch_fill_zero(float4* p, uint64_t n, hipStream_t st);

// ---- errors --------------------------------------------------------------------------------
static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }

int hip_fail(hipError_t e, const char* what) {
  g_error = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? LG_E_OOM : LG_E_HIP;
}

static int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

// ---- device buffer ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  // Grows to at least `need` bytes (contents are not kept).
  int ensure(size_t need) {
    if (need <= bytes && p) return LG_OK;
    reset();
    size_t b = need ? need : 16;
    LG_HIP(hipMalloc(&p, b));
    bytes = b;
    return LG_OK;
  }
  template <typename U>
  U* as() const {
    return reinterpret_cast<U*>(p);
  }
};

// ---- per-stage timing ---------------------------------------------------------------------
struct StageRec {
  std::string name;
  hipEvent_t start, stop;
  uint64_t work;
};

}  // namespace lg

using namespace lg;

struct lg_ctx {
  int device = 0;
  int children = 0;      // live scenes, segment buffers and images
  bool closing = false;  // lg_ctx_destroy was called while children were alive
  int n_cu = 256;
  hipStream_t stream = nullptr;
  bool profiling = false;
  std::vector<StageRec> pending;        // events of the current call
  std::vector<hipEvent_t> event_pool;
  std::vector<lg_stage_time> last_times;
  // accumulation scratch
  DevBuf counts, offsets, scan_scratch, tile_start, pairs, items, item_counter, total64;
  std::vector<uint32_t> h_tile_start;
  std::vector<WorkItem> h_items;
  WorkItem* pinned_items = nullptr;
  size_t pinned_items_cap = 0;

  hipEvent_t get_event() {
    if (!event_pool.empty()) {
      hipEvent_t e = event_pool.back();
      event_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
};

struct lg_scene {
  lg_ctx* ctx = nullptr;
  std::vector<lg_object> objects;
  std::vector<lg_light> lights;
  bool dirty = true;
  // compiled scene (host copies kept for the starting refractive index)
  std::vector<Rec<double>> h_recs64;
  std::vector<int32_t> h_obj_rec;
  std::vector<double> h_verts64;
  int n_recs = 0, n_objs = 0, n_verts = 0;
  DevBuf recs32, recs64, obj_n32, obj_n64, obj_flags, obj_rec, verts32, verts64, lights_dev;
  // trace scratch
  DevBuf q0, q1, counters;
  uint64_t q_cap = 0;       // rays per queue
  uint64_t seg_hint = 0;    // high-water segment count
  uint64_t n_primary = 0;
};

struct lg_segments {
  lg_ctx* ctx = nullptr;
  uint32_t flags = 0;
  uint64_t count = 0, cap = 0;
  float max_abs = 0.0f;   // largest |colour channel| of the stored segments (host-side bound)
  DevBuf pos, col0, col1, key, pos64;
  bool has_pos64 = false;
};

struct lg_image {
  lg_ctx* ctx = nullptr;
  lg_view view{};
  DevBuf accum, out16;
};

namespace lg {

// ---- profiling helpers -------------------------------------------------------------------
struct StageScope {
  lg_ctx* ctx;
  size_t idx = (size_t)-1;
  StageScope(lg_ctx* c, const char* name, uint64_t work) : ctx(c) {
    if (!ctx->profiling) return;
    StageRec r;
    r.name = name;
    r.work = work;
    r.start = ctx->get_event();
    r.stop = ctx->get_event();
    if (!r.start || !r.stop) return;
    (void)hipEventRecord(r.start, ctx->stream);
    ctx->pending.push_back(r);
    idx = ctx->pending.size() - 1;
  }
  ~StageScope() {
    if (idx != (size_t)-1) (void)hipEventRecord(ctx->pending[idx].stop, ctx->stream);
  }
};

static void begin_call(lg_ctx* ctx) {
  for (auto& r : ctx->pending) {
    ctx->event_pool.push_back(r.start);
    ctx->event_pool.push_back(r.stop);
  }
  ctx->pending.clear();
}

// Resolves event times after the stream has been synchronised.
static void end_call(lg_ctx* ctx) {
  if (!ctx->profiling) return;
  std::vector<lg_stage_time> out;
  for (auto& r : ctx->pending) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, r.start, r.stop);
    auto it = std::find_if(out.begin(), out.end(), [&](const lg_stage_time& t) {
      return std::strncmp(t.name, r.name.c_str(), sizeof(t.name)) == 0;
    });
    if (it == out.end()) {
      lg_stage_time t{};
      std::snprintf(t.name, sizeof(t.name), "%s", r.name.c_str());
      out.push_back(t);
      it = out.end() - 1;
    }
    it->ms += ms;
    it->launches += 1;
    it->work += r.work;
  }
  ctx->last_times = out;
}

// ---- scene compilation -------------------------------------------------------------------
struct Frame {
  double o[2];
  double r[4];  // column-major
};

static Frame frame_identity() { return Frame{{0, 0}, {1, 0, 0, 1}}; }

static Frame frame_compose(const Frame& parent, const double origin[2], const double rot[4]) {
  Frame f;
  f.o[0] = parent.o[0] + parent.r[0] * origin[0] + parent.r[2] * origin[1];
  f.o[1] = parent.o[1] + parent.r[1] * origin[0] + parent.r[3] * origin[1];
  // parent.R * rot (column-major)
  f.r[0] = parent.r[0] * rot[0] + parent.r[2] * rot[1];
  f.r[1] = parent.r[1] * rot[0] + parent.r[3] * rot[1];
  f.r[2] = parent.r[0] * rot[2] + parent.r[2] * rot[3];
  f.r[3] = parent.r[1] * rot[2] + parent.r[3] * rot[3];
  return f;
}

static void frame_point(const Frame& f, double x, double y, double& ox, double& oy) {
  ox = f.o[0] + f.r[0] * x + f.r[2] * y;
  oy = f.o[1] + f.r[1] * x + f.r[3] * y;
}

static bool finite_all(const double* v, int n) {
  for (int i = 0; i < n; i++)
    if (!std::isfinite(v[i])) return false;
  return true;
}

static bool rotation_ok(const double r[4]) {
  if (!finite_all(r, 4)) return false;
  // columns unit length and orthogonal
  double c0 = r[0] * r[0] + r[1] * r[1], c1 = r[2] * r[2] + r[3] * r[3];
  double d = r[0] * r[2] + r[1] * r[3];
  return std::fabs(c0 - 1.0) < 1e-6 && std::fabs(c1 - 1.0) < 1e-6 && std::fabs(d) < 1e-6;
}

static bool is_closed_prim(int kind) {
  return kind == LG_SHAPE_CIRCLE || kind == LG_SHAPE_RECT || kind == LG_SHAPE_ELLIPSE ||
         kind == LG_SHAPE_CONVEX_POLYGON;
}

static bool shape_is_framed(int kind) { return is_closed_prim(kind) || kind == LG_SHAPE_LOGIC; }

// One primitive shape node -> world-space record. Polygon vertices are appended to verts.
static int compile_prim(const lg_object& obj, int oi, const lg_shape& s, const Frame& parent,
                        std::vector<Rec<double>>& out, std::vector<double>& verts) {
  Rec<double> R;
  std::memset(&R, 0, sizeof(R));
  R.kind = s.kind;
  R.obj = oi;
  switch (s.kind) {
    case LG_SHAPE_LINE_SEGMENT: {
      if (!finite_all(s.p, 4)) return fail(LG_E_INVALID, "line segment: non-finite point");
      double ax, ay, bx, by;
      frame_point(parent, s.p[0], s.p[1], ax, ay);
      frame_point(parent, s.p[2], s.p[3], bx, by);
      double ex = bx - ax, ey = by - ay;
      double l = std::sqrt(ex * ex + ey * ey);
      if (!(l > 0)) return fail(LG_E_INVALID, "line segment: zero length");
      R.p[0] = ax; R.p[1] = ay; R.p[2] = ex; R.p[3] = ey;
      R.p[4] = -ey / l; R.p[5] = ex / l;
      break;
    }
    case LG_SHAPE_CUBIC_BEZIER: {
      if (!finite_all(s.p, 8)) return fail(LG_E_INVALID, "bezier: non-finite point");
      double P[8];
      for (int i = 0; i < 4; i++) frame_point(parent, s.p[2 * i], s.p[2 * i + 1], P[2 * i], P[2 * i + 1]);
      for (int c = 0; c < 2; c++) {
        double p0 = P[c], p1 = P[2 + c], p2 = P[4 + c], p3 = P[6 + c];
        R.p[0 + c] = -p0 + 3.0 * p1 - 3.0 * p2 + p3;
        R.p[2 + c] = 3.0 * p0 - 6.0 * p1 + 3.0 * p2;
        R.p[4 + c] = 3.0 * (p1 - p0);
        R.p[6 + c] = p0;
      }
      break;
    }
    case LG_SHAPE_CIRCLE: {
      if (!finite_all(s.origin, 2) || !std::isfinite(s.p[0]) || !(s.p[0] > 0))
        return fail(LG_E_INVALID, "circle: bad centre or radius");
      double cx, cy;
      frame_point(parent, s.origin[0], s.origin[1], cx, cy);
      R.p[0] = cx; R.p[1] = cy; R.p[2] = s.p[0]; R.p[3] = s.p[0] * s.p[0];
      break;
    }
    case LG_SHAPE_RECT:
    case LG_SHAPE_ELLIPSE:
    case LG_SHAPE_CONVEX_POLYGON: {
      if (!finite_all(s.origin, 2) || !rotation_ok(s.rotation))
        return fail(LG_E_INVALID, "shape: bad origin or rotation (must be a 2x2 rotation)");
      Frame f = frame_compose(parent, s.origin, s.rotation);
      R.p[0] = f.o[0]; R.p[1] = f.o[1];
      R.p[2] = f.r[0]; R.p[3] = f.r[1];  // local x axis
      R.p[4] = f.r[2]; R.p[5] = f.r[3];  // local y axis
      if (s.kind == LG_SHAPE_RECT || s.kind == LG_SHAPE_ELLIPSE) {
        if (!std::isfinite(s.p[0]) || !std::isfinite(s.p[1]) || !(s.p[0] > 0) || !(s.p[1] > 0))
          return fail(LG_E_INVALID, "rect/ellipse: sizes must be positive");
        if (s.kind == LG_SHAPE_RECT) {
          R.p[6] = 0.5 * s.p[0];
          R.p[7] = 0.5 * s.p[1];
        } else {
          R.p[6] = s.p[0];
          R.p[7] = s.p[1];
        }
      } else {
        int nv = s.n_vertices, first = s.vertex_offset;
        if (nv < 3 || first < 0 || first + nv > LG_MAX_POLY_VERTICES)
          return fail(LG_E_INVALID, "polygon: bad vertex range");
        const double* v = obj.vertices + 2 * first;
        if (!finite_all(v, 2 * nv)) return fail(LG_E_INVALID, "polygon: non-finite vertex");
        for (int i = 0; i < nv; i++) {
          const double* a = v + 2 * i;
          const double* b = v + 2 * ((i + 1) % nv);
          const double* c = v + 2 * ((i + 2) % nv);
          double cr = (b[0] - a[0]) * (c[1] - b[1]) - (b[1] - a[1]) * (c[0] - b[0]);
          if (!(cr > 0)) return fail(LG_E_INVALID, "polygon: vertices must be convex and counter-clockwise");
        }
        R.aux = (int32_t)(verts.size() / 2);
        R.op = nv;
        if ((int)(verts.size() / 2) + nv > kMaxPolyVertices)
          return fail(LG_E_UNSUPPORTED, "scene: too many polygon vertices");
        for (int i = 0; i < 2 * nv; i++) verts.push_back(v[i]);
      }
      break;
    }
    default: return fail(LG_E_INVALID, "unknown shape kind");
  }
  out.push_back(R);
  return LG_OK;
}

static int compile_object(const lg_object& obj, int oi, std::vector<Rec<double>>& recs,
                          std::vector<double>& verts) {
  if (obj.n_shapes < 1 || obj.n_shapes > LG_MAX_SHAPES)
    return fail(LG_E_INVALID, "object " + std::to_string(oi) + ": n_shapes out of range");
  if (obj.has_material && (!std::isfinite(obj.refractive_index) || !(obj.refractive_index > 0)))
    return fail(LG_E_INVALID, "object " + std::to_string(oi) + ": bad refractive index");
  const lg_shape& root = obj.shapes[0];
  Frame id = frame_identity();
  if (root.kind == LG_SHAPE_LOGIC) {
    if (root.op < LG_LOGIC_AND || root.op > LG_LOGIC_AND_NOT)
      return fail(LG_E_INVALID, "logic: bad op");
    if (root.child_a <= 0 || root.child_b <= 0 || root.child_a >= obj.n_shapes ||
        root.child_b >= obj.n_shapes || root.child_a == root.child_b)
      return fail(LG_E_INVALID, "logic: bad operand index");
    if (!finite_all(root.origin, 2) || !rotation_ok(root.rotation))
      return fail(LG_E_INVALID, "logic: bad origin or rotation");
    const lg_shape& a = obj.shapes[root.child_a];
    const lg_shape& b = obj.shapes[root.child_b];
    if (!is_closed_prim(a.kind) || !is_closed_prim(b.kind))
      return fail(LG_E_UNSUPPORTED,
                  "logic: operands must be circles, rects, ellipses or convex polygons");
    Frame f = frame_compose(id, root.origin, root.rotation);
    Rec<double> L;
    std::memset(&L, 0, sizeof(L));
    L.kind = REC_LOGIC;
    L.obj = oi;
    L.op = root.op;
    recs.push_back(L);
    int rc = compile_prim(obj, oi, a, f, recs, verts);
    if (rc) return rc;
    return compile_prim(obj, oi, b, f, recs, verts);
  }
  return compile_prim(obj, oi, root, id, recs, verts);
}

template <typename U>
static int upload(DevBuf& b, const U* data, size_t n, hipStream_t st) {
  int rc = b.ensure(std::max<size_t>(n * sizeof(U), 16));
  if (rc) return rc;
  if (n) LG_HIP(hipMemcpyAsync(b.p, data, n * sizeof(U), hipMemcpyHostToDevice, st));
  return LG_OK;
}

static int compile_scene(lg_scene* sc) {
  if (!sc->dirty) return LG_OK;
  hipStream_t st = sc->ctx->stream;
  std::vector<Rec<double>> recs;
  std::vector<double> verts;
  std::vector<int32_t> obj_rec;
  std::vector<uint32_t> flags;
  std::vector<double> n64;
  if (sc->objects.size() > (size_t)kMaxObjects) return fail(LG_E_UNSUPPORTED, "scene: too many objects");
  for (size_t i = 0; i < sc->objects.size(); i++) {
    const lg_object& o = sc->objects[i];
    obj_rec.push_back((int32_t)recs.size());
    int rc = compile_object(o, (int)i, recs, verts);
    if (rc) return rc;
    uint32_t f = 0;
    if (o.has_material) f |= OBJ_MATERIAL;
    if (shape_is_framed(o.shapes[0].kind)) f |= OBJ_CLOSED;
    flags.push_back(f);
    n64.push_back(o.has_material ? o.refractive_index : 1.0);
  }
  if ((int)recs.size() > kMaxRecords) return fail(LG_E_UNSUPPORTED, "scene: too many shape records");
  std::vector<Rec<float>> recs32(recs.size());
  for (size_t i = 0; i < recs.size(); i++) {
    recs32[i].kind = recs[i].kind;
    recs32[i].obj = recs[i].obj;
    recs32[i].op = recs[i].op;
    recs32[i].aux = recs[i].aux;
    for (int k = 0; k < 12; k++) recs32[i].p[k] = (float)recs[i].p[k];
  }
  std::vector<float> n32(n64.begin(), n64.end()), verts32(verts.begin(), verts.end());
  int rc;
  if ((rc = upload(sc->recs32, recs32.data(), recs32.size(), st))) return rc;
  if ((rc = upload(sc->recs64, recs.data(), recs.size(), st))) return rc;
  if ((rc = upload(sc->obj_n32, n32.data(), n32.size(), st))) return rc;
  if ((rc = upload(sc->obj_n64, n64.data(), n64.size(), st))) return rc;
  if ((rc = upload(sc->obj_flags, flags.data(), flags.size(), st))) return rc;
  if ((rc = upload(sc->obj_rec, obj_rec.data(), obj_rec.size(), st))) return rc;
  if ((rc = upload(sc->verts32, verts32.data(), verts32.size(), st))) return rc;
  if ((rc = upload(sc->verts64, verts.data(), verts.size(), st))) return rc;
  // lights
  if (sc->lights.size() > (size_t)kMaxLights) return fail(LG_E_UNSUPPORTED, "scene: too many lights");
  std::vector<DevLight> dl;
  uint64_t base = 0;
  for (size_t l = 0; l < sc->lights.size(); l++) {
    const lg_light& L = sc->lights[l];
    DevLight d;
    std::memset(&d, 0, sizeof(d));
    d.kind = L.kind;
    d.num_rays = L.num_rays;
    d.ray_base = base;
    base += L.num_rays;
    std::memcpy(d.color, L.color, sizeof(d.color));
    if (!finite_all(L.position, 2)) return fail(LG_E_INVALID, "light: non-finite position");
    d.a[0] = L.position[0];
    d.a[1] = L.position[1];
    if (L.kind == LG_LIGHT_SPOT) {
      double dx = L.direction[0], dy = L.direction[1];
      double ang = std::fabs(dx) < kEpsilon ? (dy >= 0.0 ? M_PI * 0.5 : -M_PI * 0.5) : std::atan(dy / dx);
      d.min_angle = ang - 0.5 * L.spot_angle;
      d.spot_angle = L.spot_angle;
      d.sign = std::signbit(dx) ? -1.0 : 1.0;
    } else if (L.kind == LG_LIGHT_DIRECTIONAL) {
      double ex = L.direction[0] - L.position[0], ey = L.direction[1] - L.position[1];
      double l = std::sqrt(ex * ex + ey * ey);
      if (!(l > 0)) return fail(LG_E_INVALID, "directional light: zero-length band");
      d.e[0] = ex;
      d.e[1] = ey;
      d.nrm[0] = -ey / l;
      d.nrm[1] = ex / l;
    } else if (L.kind != LG_LIGHT_POINT) {
      return fail(LG_E_INVALID, "light: unknown kind");
    }
    // starting refractive index: last object with a material containing the origin
    double ior = 1.0;
    for (size_t o = 0; o < sc->objects.size(); o++)
      if ((flags[o] & OBJ_MATERIAL) && contains_object<double>(recs.data(), verts.data(), obj_rec[o],
                                                               L.position[0], L.position[1]))
        ior = n64[o];
    d.ior = ior;
    dl.push_back(d);
  }
  if (base >= (1ull << (64 - kPathBits))) return fail(LG_E_UNSUPPORTED, "scene: too many rays");
  if ((rc = upload(sc->lights_dev, dl.data(), dl.size(), st))) return rc;
  sc->n_primary = base;
  sc->h_recs64 = recs;
  sc->h_obj_rec = obj_rec;
  sc->h_verts64 = verts;
  sc->n_recs = (int)recs.size();
  sc->n_objs = (int)sc->objects.size();
  sc->n_verts = (int)(verts.size() / 2);
  sc->dirty = false;
  return LG_OK;
}

// ---- trace -----------------------------------------------------------------------------------
static int seg_ensure(lg_segments* s, uint64_t cap, bool need64) {
  int rc;
  if (cap > s->cap || !s->pos.p) {
    if ((rc = s->pos.ensure(cap * sizeof(float4)))) return rc;
    if ((rc = s->col0.ensure(cap * sizeof(float4)))) return rc;
    if ((rc = s->key.ensure(cap * sizeof(uint64_t)))) return rc;
    if (s->flags & LG_SEG_VERTEX_COLORS)
      if ((rc = s->col1.ensure(cap * sizeof(float4)))) return rc;
    s->cap = cap;
    s->pos64.reset();
  }
  if (need64 && s->pos64.bytes < cap * 4 * sizeof(double))
    if ((rc = s->pos64.ensure(s->cap * 4 * sizeof(double)))) return rc;
  return LG_OK;
}

template <typename T>
static int trace_impl(lg_scene* sc, const lg_params* P, lg_segments* out) {
  lg_ctx* ctx = sc->ctx;
  hipStream_t st = ctx->stream;
  const bool fp64 = sizeof(T) == 8;
  const uint64_t n_primary = sc->n_primary;
  out->count = 0;
  out->has_pos64 = false;
  if (n_primary == 0 || P->max_bounce == 0) return LG_OK;

  size_t lds = trace_lds_bytes<T>(sc->n_recs, sc->n_objs, sc->n_verts);
  if (lds > 64 * 1024) return fail(LG_E_UNSUPPORTED, "scene too large for the LDS-resident trace");

  uint64_t q_need = std::max<uint64_t>(sc->q_cap, std::min<uint64_t>(2 * n_primary, 0xffffffffull));
  uint64_t seg_need = std::max<uint64_t>(sc->seg_hint, n_primary * std::min<uint32_t>(P->max_bounce, 4u) + 1024);
  for (int attempt = 0; attempt < 4; attempt++) {
    int rc;
    if (q_need > 0xffffffffull) return fail(LG_E_UNSUPPORTED, "ray queue exceeds 2^32 entries");
    if ((rc = sc->q0.ensure(q_need * sizeof(QRay<double>)))) return rc;
    if ((rc = sc->q1.ensure(q_need * sizeof(QRay<double>)))) return rc;
    sc->q_cap = q_need;
    if ((rc = seg_ensure(out, seg_need, fp64))) return rc;
    size_t cnt_bytes = 8 + 4 * (size_t)(P->max_bounce + 1);
    if ((rc = sc->counters.ensure(cnt_bytes))) return rc;
    LG_HIP(hipMemsetAsync(sc->counters.p, 0, cnt_bytes, st));
    unsigned long long* seg_count = sc->counters.as<unsigned long long>();
    uint32_t* gen_counts = reinterpret_cast<uint32_t*>(sc->counters.as<char>() + 8);

    TraceArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    a.recs = fp64 ? (const Rec<T>*)sc->recs64.p : (const Rec<T>*)sc->recs32.p;
    a.recs64 = sc->recs64.as<const Rec<double>>();
    a.obj_n64 = sc->obj_n64.as<const double>();
    a.obj_flags = sc->obj_flags.as<const uint32_t>();
    a.obj_rec = sc->obj_rec.as<const int32_t>();
    a.verts = fp64 ? (const T*)sc->verts64.p : (const T*)sc->verts32.p;
    a.verts64 = sc->verts64.as<const double>();
    a.n_recs = sc->n_recs;
    a.n_objs = sc->n_objs;
    a.n_verts = sc->n_verts;
    a.n_lights = (int)sc->lights.size();
    a.lights = sc->lights_dev.as<const DevLight>();
    a.n_primary = n_primary;
    a.q_cap = (uint32_t)sc->q_cap;
    for (int c = 0; c < 4; c++) a.cutoff[c] = P->cutoff_color[c];
    for (int c = 0; c < 4; c++) a.canvas[c] = P->canvas[c];
    a.seg_pos = out->pos.as<float4>();
    a.seg_col = out->col0.as<float4>();
    a.seg_key = out->key.as<uint64_t>();
    a.seg_pos64 = fp64 ? out->pos64.as<double>() : nullptr;
    a.seg_count = seg_count;
    a.seg_cap = out->cap;
    const int persistent = ctx->n_cu * 8;
    for (uint32_t g = 0; g < P->max_bounce; g++) {
      a.gen = g;
      a.last_gen = (g + 1 == P->max_bounce) ? 1u : 0u;
      a.in_q = (g & 1) ? sc->q0.as<const QRay<double>>() : sc->q1.as<const QRay<double>>();
      a.out_q = (g & 1) ? sc->q1.as<QRay<double>>() : sc->q0.as<QRay<double>>();
      a.in_count = g ? gen_counts + (g - 1) : nullptr;
      a.out_count = gen_counts + g;
      uint64_t want = g == 0 ? (n_primary + 255) / 256 : (uint64_t)persistent;
      int grid = (int)std::min<uint64_t>(std::max<uint64_t>(want, 1), (uint64_t)persistent);
      StageScope sc_t(ctx, fp64 ? "trace_fp64" : "trace", 0);
      LG_HIP(launch_trace_gen<T>(a, grid, st));
    }
    std::vector<uint32_t> hc(P->max_bounce + 1 + 2);
    LG_HIP(hipMemcpyAsync(hc.data(), sc->counters.p, cnt_bytes, hipMemcpyDeviceToHost, st));
    LG_HIP(hipStreamSynchronize(st));
    uint64_t segs = (uint64_t)hc[0] | ((uint64_t)hc[1] << 32);
    uint64_t qmax = 0;
    for (uint32_t g = 0; g < P->max_bounce; g++) qmax = std::max<uint64_t>(qmax, hc[2 + g]);
    bool over = segs > out->cap || qmax > sc->q_cap;
    if (!over) {
      out->count = segs;
      // children only scale RGB by R or 1 - R, so light colours bound every segment colour
      float mx = 0.0f;
      for (const lg_light& L : sc->lights)
        for (int c = 0; c < 4; c++) mx = std::max(mx, std::fabs(L.color[c]));
      out->max_abs = mx;
      out->has_pos64 = fp64;
      sc->seg_hint = std::max(sc->seg_hint, segs);
      return LG_OK;
    }
    seg_need = std::max<uint64_t>(seg_need, segs + segs / 4 + 1024);
    q_need = std::max<uint64_t>(q_need, qmax + qmax / 4 + 1024);
  }
  return fail(LG_E_CAPACITY, "trace: buffers kept overflowing");
}

// ---- accumulation --------------------------------------------------------------------------
static AccumView make_view(const lg_view& v) {
  AccumView a;
  a.left = v.left;
  a.top = v.top;
  a.sx = (double)v.width / (v.right - v.left);
  a.sy = (double)v.height / (v.top - v.bottom);
  a.width = (int32_t)v.width;
  a.height = (int32_t)v.height;
  int T = tile_size();
  a.tiles_x = (int32_t)((v.width + T - 1) / T);
  a.tiles_y = (int32_t)((v.height + T - 1) / T);
  return a;
}

static int accumulate_impl(lg_ctx* ctx, const lg_segments* segs, lg_image* img) {
  hipStream_t st = ctx->stream;
  uint64_t n = segs->count;
  if (n == 0) return LG_OK;
  if (n > 0xffffffffull) return fail(LG_E_UNSUPPORTED, "accumulate: more than 2^32 segments");
  AccumView v = make_view(img->view);
  const int n_tiles = v.tiles_x * v.tiles_y;
  if (n_tiles > 16384) return fail(LG_E_UNSUPPORTED, "accumulate: image above 8192x8192");
  // about two chunks per CU: few enough that the [tile][chunk] count matrix stays small
  uint64_t chunk = std::max<uint64_t>(1024, (n + 2 * (uint64_t)ctx->n_cu - 1) / (2 * (uint64_t)ctx->n_cu));
  while (((n + chunk - 1) / chunk) * (uint64_t)n_tiles > (64ull << 20)) chunk *= 2;
  uint32_t n_chunks = (uint32_t)((n + chunk - 1) / chunk);
  uint64_t n_cells = (uint64_t)n_tiles * n_chunks;
  int rc;
  if ((rc = ctx->counts.ensure(n_cells * 4))) return rc;
  if ((rc = ctx->offsets.ensure(n_cells * 4))) return rc;
  if ((rc = ctx->scan_scratch.ensure(scan_scratch_elems(n_cells) * 4))) return rc;
  if ((rc = ctx->tile_start.ensure((n_tiles + 1) * 4))) return rc;
  if ((rc = ctx->total64.ensure(16))) return rc;
  const float4* pos = segs->pos.as<const float4>();
  {
    StageScope s(ctx, "bin_count", n);
    LG_HIP(launch_bin_count(pos, n, v, (uint32_t)chunk, n_chunks, ctx->counts.as<uint32_t>(), st));
  }
  {
    StageScope s(ctx, "bin_scan", n_cells);
    LG_HIP(hipMemsetAsync(ctx->total64.p, 0, 8, st));
    LG_HIP(launch_sum_u64(ctx->counts.as<uint32_t>(), n_cells, ctx->total64.as<unsigned long long>(), st));
    LG_HIP(launch_scan_exclusive(ctx->counts.as<uint32_t>(), ctx->offsets.as<uint32_t>(), n_cells,
                                 ctx->scan_scratch.as<uint32_t>(), st));
    LG_HIP(launch_tile_starts(ctx->offsets.as<uint32_t>(), n_chunks, n_tiles,
                              ctx->total64.as<unsigned long long>(), ctx->tile_start.as<uint32_t>(), st));
  }
  ctx->h_tile_start.resize(n_tiles + 1);
  unsigned long long total = 0;
  LG_HIP(hipMemcpyAsync(ctx->h_tile_start.data(), ctx->tile_start.p, (n_tiles + 1) * 4,
                        hipMemcpyDeviceToHost, st));
  LG_HIP(hipMemcpyAsync(&total, ctx->total64.p, 8, hipMemcpyDeviceToHost, st));
  LG_HIP(hipStreamSynchronize(st));
  if (total == 0) return LG_OK;
  if (total > 0xffffffffull) return fail(LG_E_UNSUPPORTED, "accumulate: more than 2^32 segment-tile pairs");
  if ((rc = ctx->pairs.ensure(total * 4))) return rc;
  {
    StageScope s(ctx, "bin_fill", total);
    LG_HIP(launch_bin_fill(pos, n, v, (uint32_t)chunk, n_chunks, ctx->offsets.as<uint32_t>(),
                           ctx->pairs.as<uint32_t>(), st));
  }
  // work items: tiles split into pieces of at most `piece` pairs, largest first
  const int grid = ctx->n_cu;
  const uint64_t kMaxPiece = 1u << 20;
  uint64_t piece = std::min<uint64_t>(kMaxPiece, std::max<uint64_t>(4096, total / ((uint64_t)grid * 2) + 1));
  // fixed point of the LDS tile: colour channel c -> round(c * 2^G), weights in 2^-20. A
  // pixel takes at most one deposit per pair, so an item of <= kMaxPiece pairs stays below
  // kMaxPiece * 2^20 * 2^31 < 2^63 when |c| * 2^G < 2^31.
  double bound = std::max<double>(segs->max_abs, (double)segs->max_abs * segs->max_abs);
  int G = 60;
  if (bound > 0.0 && std::isfinite(bound)) {
    double g = std::min(62.0 - weight_bits() - std::log2((double)kMaxPiece), 30.0) - std::log2(bound);
    G = (int)std::floor(g) - 1;
    G = std::max(-30, std::min(60, G));
  }
  ctx->h_items.clear();
  for (int t = 0; t < n_tiles; t++) {
    uint32_t b = ctx->h_tile_start[t], e = ctx->h_tile_start[t + 1];
    if (e <= b) continue;
    uint32_t cnt = e - b;
    uint32_t split = cnt > piece ? 1u : 0u;
    for (uint32_t p = b; p < e; p += (uint32_t)piece) {
      WorkItem w;
      w.tile = (uint32_t)t;
      w.split = split;
      w.begin = p;
      w.end = (uint32_t)std::min<uint64_t>(e, (uint64_t)p + piece);
      ctx->h_items.push_back(w);
    }
  }
  std::stable_sort(ctx->h_items.begin(), ctx->h_items.end(), [](const WorkItem& a, const WorkItem& b) {
    return (a.end - a.begin) > (b.end - b.begin);
  });
  uint32_t n_items = (uint32_t)ctx->h_items.size();
  if (ctx->pinned_items_cap < n_items) {
    if (ctx->pinned_items) (void)hipHostFree(ctx->pinned_items);
    ctx->pinned_items = nullptr;
    LG_HIP(hipHostMalloc((void**)&ctx->pinned_items, n_items * sizeof(WorkItem)));
    ctx->pinned_items_cap = n_items;
  }
  std::memcpy(ctx->pinned_items, ctx->h_items.data(), n_items * sizeof(WorkItem));
  if ((rc = ctx->items.ensure(n_items * sizeof(WorkItem)))) return rc;
  if ((rc = ctx->item_counter.ensure(16))) return rc;
  LG_HIP(hipMemcpyAsync(ctx->items.p, ctx->pinned_items, n_items * sizeof(WorkItem),
                        hipMemcpyHostToDevice, st));
  LG_HIP(hipMemsetAsync(ctx->item_counter.p, 0, 4, st));
  {
    StageScope s(ctx, "tile_accum", total);
    int g = (int)std::min<uint32_t>(n_items, (uint32_t)grid);
    LG_HIP(launch_tile_accum(pos, segs->col0.as<const float4>(),
                             (segs->flags & LG_SEG_VERTEX_COLORS) ? segs->col1.as<const float4>() : nullptr,
                             ctx->pairs.as<const uint32_t>(), ctx->items.as<const WorkItem>(), n_items,
                             ctx->item_counter.as<uint32_t>(), v, img->accum.as<float4>(), g, G, st));
  }
  return LG_OK;
}

}  // namespace lg

// =============================================================================================
// C ABI
// =============================================================================================
extern "C" {

const char* lg_last_error(void) { return g_error.c_str(); }

int lg_version(void) { return LG_ABI_VERSION; }

int lg_ctx_create(int device, lg_ctx** out) {
  if (!out) return fail(LG_E_INVALID, "lg_ctx_create: out is NULL");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(LG_E_NODEVICE, "no HIP device");
  if (device < 0 || device >= n) return fail(LG_E_NODEVICE, "device ordinal out of range");
  LG_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  LG_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(LG_E_NODEVICE, std::string("device is ") + prop.gcnArchName + ", this build targets gfx950");
  lg_ctx* c = new lg_ctx();
  c->device = device;
  c->n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "hipStreamCreate");
  }
  *out = c;
  return LG_OK;
}

static void ctx_free(lg_ctx* ctx) {
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  begin_call(ctx);
  for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
  if (ctx->pinned_items) (void)hipHostFree(ctx->pinned_items);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

// A context outlives its children: destroying it while scenes, segment buffers or images
// are alive defers the release to the last child's destroy.
int lg_ctx_destroy(lg_ctx* ctx) {
  if (!ctx) return LG_OK;
  if (ctx->children > 0) {
    ctx->closing = true;
    return LG_OK;
  }
  ctx_free(ctx);
  return LG_OK;
}

static void ctx_child_gone(lg_ctx* ctx) {
  if (--ctx->children == 0 && ctx->closing) ctx_free(ctx);
}

int lg_ctx_synchronize(lg_ctx* ctx) {
  if (!ctx) return fail(LG_E_INVALID, "ctx is NULL");
  LG_HIP(hipStreamSynchronize(ctx->stream));
  return LG_OK;
}

int lg_ctx_set_profiling(lg_ctx* ctx, int enable) {
  if (!ctx) return fail(LG_E_INVALID, "ctx is NULL");
  ctx->profiling = enable != 0;
  return LG_OK;
}

int lg_ctx_stage_times(lg_ctx* ctx, lg_stage_time* out, int cap, int* n_out) {
  if (!ctx || !n_out) return fail(LG_E_INVALID, "ctx or n_out is NULL");
  *n_out = (int)ctx->last_times.size();
  if (cap < *n_out) return fail(LG_E_CAPACITY, "stage array too small");
  for (int i = 0; i < *n_out; i++) out[i] = ctx->last_times[i];
  return LG_OK;
}

// ---- scene ---------------------------------------------------------------------------------
int lg_scene_create(lg_ctx* ctx, lg_scene** out) {
  if (!ctx || !out) return fail(LG_E_INVALID, "ctx or out is NULL");
  lg_scene* s = new lg_scene();
  s->ctx = ctx;
  ctx->children++;
  *out = s;
  return LG_OK;
}

int lg_scene_destroy(lg_scene* scene) {
  if (!scene) return LG_OK;
  lg_ctx* ctx = scene->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete scene;
  ctx_child_gone(ctx);
  return LG_OK;
}

int lg_scene_clear(lg_scene* s) {
  if (!s) return fail(LG_E_INVALID, "scene is NULL");
  s->objects.clear();
  s->lights.clear();
  s->dirty = true;
  return LG_OK;
}

int lg_scene_clear_objects(lg_scene* s) {
  if (!s) return fail(LG_E_INVALID, "scene is NULL");
  s->objects.clear();
  s->dirty = true;
  return LG_OK;
}

int lg_scene_push_object(lg_scene* s, const lg_object* obj) {
  if (!s || !obj) return fail(LG_E_INVALID, "scene or object is NULL");
  s->objects.push_back(*obj);
  s->dirty = true;
  return LG_OK;
}

int lg_scene_push_light(lg_scene* s, const lg_light* light) {
  if (!s || !light) return fail(LG_E_INVALID, "scene or light is NULL");
  s->lights.push_back(*light);
  s->dirty = true;
  return LG_OK;
}

int lg_scene_set_objects(lg_scene* s, const lg_object* objs, uint32_t n) {
  if (!s || (n && !objs)) return fail(LG_E_INVALID, "scene or objects is NULL");
  s->objects.assign(objs, objs + n);
  s->dirty = true;
  return LG_OK;
}

int lg_scene_set_lights(lg_scene* s, const lg_light* lights, uint32_t n) {
  if (!s || (n && !lights)) return fail(LG_E_INVALID, "scene or lights is NULL");
  s->lights.assign(lights, lights + n);
  s->dirty = true;
  return LG_OK;
}

int lg_scene_replace_object(lg_scene* s, uint32_t ix, const lg_object* obj) {
  if (!s || !obj) return fail(LG_E_INVALID, "scene or object is NULL");
  if (ix >= s->objects.size()) return fail(LG_E_INVALID, "object index out of range");
  s->objects[ix] = *obj;
  s->dirty = true;
  return LG_OK;
}

int lg_scene_remove_object(lg_scene* s, uint32_t ix) {
  if (!s) return fail(LG_E_INVALID, "scene is NULL");
  if (ix >= s->objects.size()) return fail(LG_E_INVALID, "object index out of range");
  s->objects.erase(s->objects.begin() + ix);
  s->dirty = true;
  return LG_OK;
}

int lg_scene_remove_light(lg_scene* s, uint32_t ix) {
  if (!s) return fail(LG_E_INVALID, "scene is NULL");
  if (ix >= s->lights.size()) return fail(LG_E_INVALID, "light index out of range");
  s->lights.erase(s->lights.begin() + ix);
  s->dirty = true;
  return LG_OK;
}

int lg_scene_counts(const lg_scene* s, uint32_t* n_objects, uint32_t* n_lights, uint64_t* n_rays) {
  if (!s) return fail(LG_E_INVALID, "scene is NULL");
  if (n_objects) *n_objects = (uint32_t)s->objects.size();
  if (n_lights) *n_lights = (uint32_t)s->lights.size();
  if (n_rays) {
    uint64_t r = 0;
    for (auto& l : s->lights) r += l.num_rays;
    *n_rays = r;
  }
  return LG_OK;
}

// ---- segments ------------------------------------------------------------------------------
int lg_segments_create(lg_ctx* ctx, uint64_t capacity, uint32_t flags, lg_segments** out) {
  if (!ctx || !out) return fail(LG_E_INVALID, "ctx or out is NULL");
  lg_segments* s = new lg_segments();
  s->ctx = ctx;
  s->flags = flags;
  int rc = seg_ensure(s, std::max<uint64_t>(capacity, 1), false);
  if (rc) {
    delete s;
    return rc;
  }
  ctx->children++;
  *out = s;
  return LG_OK;
}

int lg_segments_destroy(lg_segments* s) {
  if (!s) return LG_OK;
  lg_ctx* ctx = s->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete s;
  ctx_child_gone(ctx);
  return LG_OK;
}

int lg_segments_count(const lg_segments* s, uint64_t* count) {
  if (!s || !count) return fail(LG_E_INVALID, "segments or count is NULL");
  *count = s->count;
  return LG_OK;
}

int lg_segments_upload(lg_segments* s, const float* pos, const float* color0, const float* color1,
                       uint64_t n) {
  if (!s || (n && (!pos || !color0))) return fail(LG_E_INVALID, "segments, pos or color0 is NULL");
  if (color1 && !(s->flags & LG_SEG_VERTEX_COLORS))
    return fail(LG_E_INVALID, "color1 given but the buffer has no per-vertex colours");
  hipStream_t st = s->ctx->stream;
  int rc = seg_ensure(s, std::max<uint64_t>(n, 1), false);
  if (rc) return rc;
  if (n) {
    LG_HIP(hipMemcpyAsync(s->pos.p, pos, n * 16, hipMemcpyHostToDevice, st));
    LG_HIP(hipMemcpyAsync(s->col0.p, color0, n * 16, hipMemcpyHostToDevice, st));
    if (s->flags & LG_SEG_VERTEX_COLORS)
      LG_HIP(hipMemcpyAsync(s->col1.p, color1 ? color1 : color0, n * 16, hipMemcpyHostToDevice, st));
    std::vector<uint64_t> keys(n);
    std::iota(keys.begin(), keys.end(), 0ull);
    float mx = 0.0f;
    for (uint64_t i = 0; i < 4 * n; i++) mx = std::max(mx, std::fabs(color0[i]));
    if (color1)
      for (uint64_t i = 0; i < 4 * n; i++) mx = std::max(mx, std::fabs(color1[i]));
    s->max_abs = mx;
    LG_HIP(hipMemcpyAsync(s->key.p, keys.data(), n * 8, hipMemcpyHostToDevice, st));
    LG_HIP(hipStreamSynchronize(st));
  }
  s->count = n;
  s->has_pos64 = false;
  return LG_OK;
}

int lg_segments_download(lg_segments* s, float* pos, float* color0, float* color1, uint64_t* key,
                         uint64_t cap, uint64_t* count) {
  if (!s) return fail(LG_E_INVALID, "segments is NULL");
  if (count) *count = s->count;
  if (cap < s->count) return fail(LG_E_CAPACITY, "download: capacity below segment count");
  hipStream_t st = s->ctx->stream;
  uint64_t n = s->count;
  if (n) {
    if (pos) LG_HIP(hipMemcpyAsync(pos, s->pos.p, n * 16, hipMemcpyDeviceToHost, st));
    if (color0) LG_HIP(hipMemcpyAsync(color0, s->col0.p, n * 16, hipMemcpyDeviceToHost, st));
    if (color1)
      LG_HIP(hipMemcpyAsync(color1, (s->flags & LG_SEG_VERTEX_COLORS) ? s->col1.p : s->col0.p, n * 16,
                            hipMemcpyDeviceToHost, st));
    if (key) LG_HIP(hipMemcpyAsync(key, s->key.p, n * 8, hipMemcpyDeviceToHost, st));
  }
  LG_HIP(hipStreamSynchronize(st));
  return LG_OK;
}

int lg_segments_download_pos64(lg_segments* s, double* pos64, uint64_t cap, uint64_t* count) {
  if (!s || !pos64) return fail(LG_E_INVALID, "segments or pos64 is NULL");
  if (count) *count = s->count;
  if (!s->has_pos64) return fail(LG_E_INVALID, "no FP64 end points: trace with LG_TRACE_FP64");
  if (cap < s->count) return fail(LG_E_CAPACITY, "download: capacity below segment count");
  if (s->count)
    LG_HIP(hipMemcpyAsync(pos64, s->pos64.p, s->count * 32, hipMemcpyDeviceToHost, s->ctx->stream));
  LG_HIP(hipStreamSynchronize(s->ctx->stream));
  return LG_OK;
}

// ---- trace ---------------------------------------------------------------------------------
static int check_params(const lg_params* p) {
  if (!p) return fail(LG_E_INVALID, "params is NULL");
  if (p->max_bounce > kMaxBounce) return fail(LG_E_UNSUPPORTED, "max_bounce above 23");
  if (!finite_all(p->canvas, 4) || !(p->canvas[1] > p->canvas[0]) || !(p->canvas[3] > p->canvas[2]))
    return fail(LG_E_INVALID, "canvas must be {left < right, bottom < top}");
  return LG_OK;
}

int lg_trace(lg_scene* scene, const lg_params* params, lg_segments* out) {
  if (!scene || !out) return fail(LG_E_INVALID, "scene or out is NULL");
  int rc = check_params(params);
  if (rc) return rc;
  lg_ctx* ctx = scene->ctx;
  LG_HIP(hipSetDevice(ctx->device));
  begin_call(ctx);
  if ((rc = compile_scene(scene))) return rc;
  if (params->flags & LG_TRACE_FP64)
    rc = trace_impl<double>(scene, params, out);
  else
    rc = trace_impl<float>(scene, params, out);
  if (rc) return rc;
  LG_HIP(hipStreamSynchronize(ctx->stream));
  end_call(ctx);
  return LG_OK;
}

int lg_trace_all(lg_scene* scene, const lg_params* params, double* xy, float* rgba,
                 uint64_t cap_vertices, uint64_t* n_vertices) {
  if (!scene || !n_vertices) return fail(LG_E_INVALID, "scene or n_vertices is NULL");
  lg_segments* segs = nullptr;
  int rc = lg_segments_create(scene->ctx, 1024, 0, &segs);
  if (rc) return rc;
  rc = lg_trace(scene, params, segs);
  if (rc) {
    lg_segments_destroy(segs);
    return rc;
  }
  uint64_t n = segs->count;
  *n_vertices = 2 * n;
  if (cap_vertices < 2 * n || (n && (!xy || !rgba))) {
    lg_segments_destroy(segs);
    return fail(LG_E_CAPACITY, "trace_all: output too small");
  }
  std::vector<float> pos(4 * n), col(4 * n);
  std::vector<double> pos64;
  std::vector<uint64_t> key(n);
  rc = lg_segments_download(segs, pos.data(), col.data(), nullptr, key.data(), n, nullptr);
  if (!rc && segs->has_pos64) {
    pos64.resize(4 * n);
    rc = lg_segments_download_pos64(segs, pos64.data(), n, nullptr);
  }
  lg_segments_destroy(segs);
  if (rc) return rc;
  std::vector<uint64_t> order(n);
  std::iota(order.begin(), order.end(), 0ull);
  std::sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return key[a] < key[b]; });
  for (uint64_t i = 0; i < n; i++) {
    uint64_t s = order[i];
    for (int v = 0; v < 2; v++) {
      xy[4 * i + 2 * v] = pos64.empty() ? (double)pos[4 * s + 2 * v] : pos64[4 * s + 2 * v];
      xy[4 * i + 2 * v + 1] = pos64.empty() ? (double)pos[4 * s + 2 * v + 1] : pos64[4 * s + 2 * v + 1];
      std::memcpy(rgba + 8 * i + 4 * v, &col[4 * s], 16);
    }
  }
  return LG_OK;
}

// ---- image ---------------------------------------------------------------------------------
int lg_image_create(lg_ctx* ctx, const lg_view* view, lg_image** out) {
  if (!ctx || !view || !out) return fail(LG_E_INVALID, "ctx, view or out is NULL");
  if (view->width == 0 || view->height == 0 || view->width > 8192 || view->height > 8192)
    return fail(LG_E_INVALID, "image size must be 1..8192");
  double b[4] = {view->left, view->right, view->bottom, view->top};
  if (!finite_all(b, 4) || !(view->right > view->left) || !(view->top > view->bottom))
    return fail(LG_E_INVALID, "view must be {left < right, bottom < top}");
  lg_image* im = new lg_image();
  im->ctx = ctx;
  im->view = *view;
  size_t px = (size_t)view->width * view->height;
  int rc = im->accum.ensure(px * 16);
  if (!rc) rc = im->out16.ensure(px * 8);
  if (rc) {
    delete im;
    return rc;
  }
  ctx->children++;
  *out = im;
  return lg_image_clear(im);
}

int lg_image_destroy(lg_image* img) {
  if (!img) return LG_OK;
  lg_ctx* ctx = img->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete img;
  ctx_child_gone(ctx);
  return LG_OK;
}

int lg_image_clear(lg_image* img) {
  if (!img) return fail(LG_E_INVALID, "image is NULL");
  size_t px = (size_t)img->view.width * img->view.height;
  LG_HIP(launch_fill_zero(img->accum.as<float4>(), px, img->ctx->stream));
  return LG_OK;
}

int lg_image_device_ptrs(lg_image* img, void** accum, void** rgba16f) {
  if (!img) return fail(LG_E_INVALID, "image is NULL");
  if (accum) *accum = img->accum.p;
  if (rgba16f) *rgba16f = img->out16.p;
  return LG_OK;
}

int lg_image_download_rgba16f(lg_image* img, uint16_t* host, uint64_t cap_halves) {
  if (!img || !host) return fail(LG_E_INVALID, "image or host is NULL");
  size_t n = (size_t)img->view.width * img->view.height * 4;
  if (cap_halves < n) return fail(LG_E_CAPACITY, "download: capacity below 4*W*H");
  LG_HIP(hipMemcpyAsync(host, img->out16.p, n * 2, hipMemcpyDeviceToHost, img->ctx->stream));
  LG_HIP(hipStreamSynchronize(img->ctx->stream));
  return LG_OK;
}

int lg_image_download_accum(lg_image* img, float* host, uint64_t cap_floats) {
  if (!img || !host) return fail(LG_E_INVALID, "image or host is NULL");
  size_t n = (size_t)img->view.width * img->view.height * 4;
  if (cap_floats < n) return fail(LG_E_CAPACITY, "download: capacity below 4*W*H");
  LG_HIP(hipMemcpyAsync(host, img->accum.p, n * 4, hipMemcpyDeviceToHost, img->ctx->stream));
  LG_HIP(hipStreamSynchronize(img->ctx->stream));
  return LG_OK;
}

// ---- accumulate / resolve / render --------------------------------------------------------------
int lg_accumulate(lg_ctx* ctx, const lg_segments* segs, lg_image* img) {
  if (!ctx || !segs || !img) return fail(LG_E_INVALID, "ctx, segments or image is NULL");
  LG_HIP(hipSetDevice(ctx->device));
  begin_call(ctx);
  int rc = accumulate_impl(ctx, segs, img);
  if (rc) return rc;
  LG_HIP(hipStreamSynchronize(ctx->stream));
  end_call(ctx);
  return LG_OK;
}

int lg_resolve(lg_ctx* ctx, lg_image* img) {
  if (!ctx || !img) return fail(LG_E_INVALID, "ctx or image is NULL");
  size_t px = (size_t)img->view.width * img->view.height;
  LG_HIP(launch_resolve(img->accum.as<const float4>(), img->out16.as<ushort4>(), px, ctx->stream));
  return LG_OK;
}

int lg_render(lg_scene* scene, const lg_params* params, lg_segments* segs, lg_image* img) {
  if (!scene || !segs || !img) return fail(LG_E_INVALID, "scene, segments or image is NULL");
  int rc = check_params(params);
  if (rc) return rc;
  lg_ctx* ctx = scene->ctx;
  LG_HIP(hipSetDevice(ctx->device));
  begin_call(ctx);
  size_t px = (size_t)img->view.width * img->view.height;
  {
    StageScope s(ctx, "clear", px);
    LG_HIP(launch_fill_zero(img->accum.as<float4>(), px, ctx->stream));
  }
  if ((rc = compile_scene(scene))) return rc;
  if (params->flags & LG_TRACE_FP64)
    rc = trace_impl<double>(scene, params, segs);
  else
    rc = trace_impl<float>(scene, params, segs);
  if (rc) return rc;
  if ((rc = accumulate_impl(ctx, segs, img))) return rc;
  {
    StageScope s(ctx, "resolve", px);
    LG_HIP(launch_resolve(img->accum.as<const float4>(), img->out16.as<ushort4>(), px, ctx->stream));
  }
  LG_HIP(hipStreamSynchronize(ctx->stream));
  end_call(ctx);
  return LG_OK;
}

}  // extern "C"

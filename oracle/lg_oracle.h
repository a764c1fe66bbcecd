/*
 * lg_oracle.h — FP64 CPU restatement of the light_garden hot path.
 *
 * TEST INFRASTRUCTURE ONLY. The oracle is the checker for the HIP path: only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg load it. The product library
 * (light_garden_amd) never links, loads or falls back to it.
 *
 * It restates, in plain C99 + pthreads:
 *   - ray generation        light.rs:103-115 (directional), 163-174 (point), 225-249 (spot)
 *   - the trace             tracer.rs:276-330 (trace_all), 360-493 (trace)
 *   - the geometry kernel   collision2d @a7b471b (absent: restated from the mathematics,
 *                           each defined choice marked DEFINED below)
 *   - the accumulation rule this build defines for the LineList draw + One/One blend
 *                           (shader.wgsl:14-28, mod.rs:57-73, renderer.rs:120-124)
 *   - the RGBA16F store     texture_renderer.rs:5, renderer.rs:175 (clear alpha 1)
 *   - string-mod segments   string_mod.rs:33-158
 */
#ifndef LG_ORACLE_H
#define LG_ORACLE_H

#include <stdint.h>
#include "../include/lg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* DEFINED: self-hit epsilon. Intersections with t <= OR_T_MIN along the ray are ignored,
 * so a child ray starting on the surface it left does not hit that point again. */
#define OR_T_MIN 1e-5
/* DEFINED: collision2d::EPSILON as used by SpotLight::set_num_rays (light.rs:230). */
#define OR_EPSILON 1e-9

typedef struct or_hit {
  double t;
  double p[2];
  double n[2];
} or_hit;

#define OR_MAX_HITS 48

/* Growable segment list: pos {x0,y0,x1,y1} doubles, color RGBA floats, key. */
typedef struct or_segs {
  double* pos;
  float* color;
  float* color1;   /* per-vertex end colour (string mod); NULL for trace output */
  uint64_t* key;
  uint64_t n, cap;
} or_segs;

/* ---- geometry -------------------------------------------------------------------- */
int or_intersect_object(const lg_object* obj, const double o[2], const double d[2],
                        or_hit* hits, int cap);
int or_contains_object(const lg_object* obj, const double p[2]);
int or_object_is_closed(const lg_object* obj);
/* Ray::refract: returns 1 when a refracted ray exists (0 on total internal reflection). */
int or_refract(const double d[2], const double n[2], double n1, double n2,
               double refl[2], double refr[2], double* reflectance);
void or_reflect(const double d[2], const double n[2], double out[2]);
int or_canvas_exit(const double canvas[4], const double o[2], const double d[2],
                   double p[2]);

/* ---- lights ------------------------------------------------------------------------ */
void or_light_ray(const lg_light* light, uint64_t i, double o[2], double d[2]);
double or_initial_ior(const lg_object* objs, uint32_t n_obj, const double p[2]);

/* ---- trace ------------------------------------------------------------------------- */
/* Traces global rays i with i % ray_stride == 0 (1 = all) on n_threads worker threads in
 * chunks of chunk_size rays (rayon par_chunks, tracer.rs:290-307). Output is in the
 * reference's order. Returns NULL on allocation failure. */
or_segs* or_trace(const lg_object* objs, uint32_t n_obj, const lg_light* lights,
                  uint32_t n_lights, const lg_params* params, uint64_t ray_stride,
                  int n_threads, int chunk_size);
uint64_t or_segs_count(const or_segs* s);
void or_segs_copy(const or_segs* s, double* pos, float* color, float* color1,
                  uint64_t* key);
void or_segs_free(or_segs* s);

/* ---- accumulation ------------------------------------------------------------------ */
/* Adds n segments into accum (H*W*4, row-major RGBA). color1 may be NULL (constant
 * colour). Rows are split into bands over n_threads threads; each pixel is owned by one
 * thread and sees segments in input order, so the result is independent of n_threads. */
void or_accumulate_f64(const double* pos, const float* color0, const float* color1,
                       uint64_t n, const lg_view* view, double* accum, int n_threads);
void or_accumulate_f32(const float* pos, const float* color0, const float* color1,
                       uint64_t n, const lg_view* view, float* accum, int n_threads);
/* The same with a blend state (dst factor One, Add/ReverseSubtract, a source factor that
 * reads neither the target nor the fragment order); returns 0 and does nothing otherwise. */
int or_accumulate_blend_f64(const double* pos, const float* color0, const float* color1,
                            uint64_t n, const lg_view* view, double* accum, int n_threads,
                            const lg_blend* blend);
/* Screenshot conversion (renderer.rs:313-328): one half -> u8, and whole images. */
uint8_t or_encode_value(uint16_t h);
void or_encode_8bit(const uint16_t* rgba16f, uint64_t n_pixels, int order, uint8_t* out);
/* Fragment count of the rule: number of column/row samples. */
uint64_t or_sample_count(const double* pos, uint64_t n, const lg_view* view);
/* FP32/FP64 accumulator -> RGBA16F bits, alpha + 1. */
void or_resolve_f64(const double* accum, uint64_t n_pixels, uint16_t* out);
void or_resolve_f32(const float* accum, uint64_t n_pixels, uint16_t* out);
uint16_t or_half_from_double(double x);
double or_double_from_half(uint16_t h);

/* ---- string mod (string_mod.rs:33-158) ---------------------------------------------- */
enum { OR_SM_ADD = 0, OR_SM_MUL = 1, OR_SM_POW = 2, OR_SM_BASE = 3 };
typedef struct or_mod_rem_color {
  uint64_t modulo, rem;
  float color[4];
} or_mod_rem_color;
/* Circle curve scaled by `scale`. Writes `modulo` segments (pos 4 doubles, color0/color1 4
 * floats each). */
void or_string_mod(uint64_t modulo, uint64_t num, uint64_t turns, int mode, double scale,
                   const float color[4], const or_mod_rem_color* mrc, int n_mrc,
                   double* pos, float* color0, float* color1);

/* General StringMod (every curve and mode; lg_string_mod as the library takes it). */
uint64_t or_sm_target(const lg_string_mod* sm, uint64_t iix);          /* string_mod.rs:111-117 */
void or_sm_point(const lg_string_mod* sm, uint64_t n, double out[2]);  /* init_points, 35-83 */
/* draw_init_points (104-122) over `points` (2 doubles each; NULL = the curve's own points);
 * writes sm->modulo segments, none when n_points is 0; returns the count. */
uint64_t or_sm_draw_points(const lg_string_mod* sm, const double* points, uint64_t n_points,
                           double* pos, float* color0, float* color1);
/* line_crossings_as_points (87-102) over explicit lines {ax, ay, bx, by}; writes at most cap
 * points, returns the count. DEFINED segment crossing: see lg_oracle.c. */
uint64_t or_line_crossings(const double* lines, uint64_t n_lines, double* out, uint64_t cap);
uint64_t or_sm_crossings(const lg_string_mod* sm, double* out, uint64_t cap);
/* draw (152-158): with sm->nested, the nested StringMod over sm's crossing points. */
uint64_t or_sm_draw(const lg_string_mod* sm, double* pos, float* color0, float* color1);

/* ---- timing helper for the CPU baseline ------------------------------------------- */
double or_now_seconds(void);

#ifdef __cplusplus
}
#endif
#endif

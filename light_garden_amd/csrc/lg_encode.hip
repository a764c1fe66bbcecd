// lg_encode.hip — the image's output side: the lines pipeline's blend state (mod.rs:57-73,
// gui/settings.rs:57-133) and the screenshot's 8-bit conversion (renderer.rs:313-328).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "lg_host.h"

using lg::fail;

namespace lg {

void blend_factor_bounds(uint32_t code, float m, const float k[4], float out[4]) {
  for (int c = 0; c < 4; c++) {
    float f;
    switch ((code >> (4 * c)) & 7u) {
      case BC_ZERO: f = 0.0f; break;
      case BC_ONE: f = 1.0f; break;
      case BC_SRC:
      case BC_SRC_ALPHA: f = m; break;
      case BC_ONE_MINUS_SRC:
      case BC_ONE_MINUS_SRC_ALPHA: f = 1.0f + m; break;
      case BC_CONST: f = std::fabs(k[c]); break;
      default: f = std::fabs(1.0f - k[c]); break;
    }
    out[c] = f;
  }
}

bool blend_nonneg(uint32_t code, float m, const float k[4]) {
  for (int c = 0; c < 4; c++) {
    const uint32_t cc = (code >> (4 * c)) & 15u;
    if (cc & BC_NEGATE) return false;
    switch (cc & 7u) {
      case BC_ONE_MINUS_SRC:
      case BC_ONE_MINUS_SRC_ALPHA:
        if (!(m <= 1.0f)) return false;
        break;
      case BC_CONST:
        if (!(k[c] >= 0.0f)) return false;
        break;
      case BC_ONE_MINUS_CONST:
        if (!(k[c] <= 1.0f)) return false;
        break;
      default: break;
    }
  }
  return true;
}

}  // namespace lg

namespace {

// A wgpu blend component as BlendCode bits, or an error: supported are dst factor One with Add
// or ReverseSubtract and a source factor that reads neither the target nor fragment order.
int component_code(const lg_blend_component& b, bool alpha, uint32_t* code) {
  if (b.src_factor > LG_BF_ONE_MINUS_CONSTANT || b.dst_factor > LG_BF_ONE_MINUS_CONSTANT ||
      b.operation > LG_BO_REVERSE_SUBTRACT)
    return fail(LG_E_INVALID, "blend: factor or operation out of range");
  if (b.dst_factor != LG_BF_ONE)
    return fail(LG_E_UNSUPPORTED, "blend: only dst factor One is supported (other dst factors make the result "
                                  "depend on the order of the fragments)");
  if (b.operation != LG_BO_ADD && b.operation != LG_BO_REVERSE_SUBTRACT)
    return fail(LG_E_UNSUPPORTED, "blend: only Add and ReverseSubtract are supported (Subtract, Min and Max "
                                  "depend on the order of the fragments or need per-pixel extrema)");
  uint32_t f;
  switch (b.src_factor) {
    case LG_BF_ZERO: f = lg::BC_ZERO; break;
    case LG_BF_ONE: f = lg::BC_ONE; break;
    case LG_BF_SRC: f = alpha ? lg::BC_SRC_ALPHA : lg::BC_SRC; break;
    case LG_BF_ONE_MINUS_SRC: f = alpha ? lg::BC_ONE_MINUS_SRC_ALPHA : lg::BC_ONE_MINUS_SRC; break;
    case LG_BF_SRC_ALPHA: f = lg::BC_SRC_ALPHA; break;
    case LG_BF_ONE_MINUS_SRC_ALPHA: f = lg::BC_ONE_MINUS_SRC_ALPHA; break;
    case LG_BF_CONSTANT: f = lg::BC_CONST; break;
    case LG_BF_ONE_MINUS_CONSTANT: f = lg::BC_ONE_MINUS_CONST; break;
    default:
      return fail(LG_E_UNSUPPORTED, "blend: source factors that read the target (Dst*, SrcAlphaSaturated) are "
                                    "not supported");
  }
  if (b.operation == LG_BO_REVERSE_SUBTRACT) f |= lg::BC_NEGATE;
  *code = f;
  return LG_OK;
}

// Rust's `f as u8` for f32: NaN and values <= 0 give 0, values >= 255 give 255, else truncation.
__device__ inline uint8_t sat_u8(float v) {
  if (!(v > 0.0f)) return 0;
  if (v >= 255.0f) return 255;
  return (uint8_t)v;
}

// table[h] for every half h: f16 -> f32 (exact), f^(1/2.2) * 255 in f32, saturating cast. The
// power is the FP64 pow of the f32 operands rounded once to f32 (the exponent is 1/2.2 in f32,
// as the reference writes `f.powf(1. / 2.2)` with f: f32).
__global__ void encode_table_kernel(uint8_t* __restrict__ table) {
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= 65536u) return;
  const float f = (float)__builtin_bit_cast(_Float16, (uint16_t)h);
  const float e = 1.0f / 2.2f;
  const float p = (float)pow((double)f, (double)e);
  table[h] = sat_u8(p * 255.0f);
}

// One thread per pixel: 4 halves in, 4 bytes out through the table.
__global__ void encode_kernel(const ushort4* __restrict__ in, uint64_t n, const uint8_t* __restrict__ table,
                              uint32_t order, uchar4* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const ushort4 h = in[i];
  const uint8_t r = table[h.x], g = table[h.y], b = table[h.z], a = table[h.w];
  out[i] = order == LG_ENCODE_RGBA ? make_uchar4(r, g, b, a) : make_uchar4(b, g, r, a);
}

int ensure_table(lg_ctx* ctx) {
  if (ctx->encode_table_ready) return LG_OK;
  int rc = ctx->encode_table.ensure(65536);
  if (rc) return rc;
  hipLaunchKernelGGL(encode_table_kernel, dim3(256), dim3(256), 0, ctx->stream, ctx->encode_table.as<uint8_t>());
  LG_HIP(hipGetLastError());
  ctx->encode_table_ready = true;
  return LG_OK;
}

}  // namespace

extern "C" {

int lg_image_set_blend(lg_image* img, const lg_blend* blend) {
  if (!img || !blend) return fail(LG_E_INVALID, "image or blend is NULL");
  uint32_t cc, ca;
  int rc;
  if ((rc = component_code(blend->color, false, &cc))) return rc;
  if ((rc = component_code(blend->alpha, true, &ca))) return rc;
  for (int c = 0; c < 4; c++)
    if (!std::isfinite(blend->constant[c])) return fail(LG_E_INVALID, "blend: constant must be finite");
  img->blend = *blend;
  img->blend_code = cc | (cc << 4) | (cc << 8) | (ca << 12);
  return LG_OK;
}

int lg_image_get_blend(const lg_image* img, lg_blend* out) {
  if (!img || !out) return fail(LG_E_INVALID, "image or out is NULL");
  *out = img->blend;
  return LG_OK;
}

int lg_image_encode_8bit(lg_image* img, uint32_t order, uint8_t* host, uint64_t cap_bytes) {
  if (!img || !host) return fail(LG_E_INVALID, "image or host is NULL");
  if (order > LG_ENCODE_RGBA) return fail(LG_E_INVALID, "encode: unknown byte order");
  const uint64_t n = (uint64_t)img->view.width * img->view.height;
  if (cap_bytes < 4 * n) return fail(LG_E_CAPACITY, "encode: capacity below 4*W*H");
  lg_ctx* ctx = img->ctx;
  LG_HIP(hipSetDevice(ctx->device));
  lg::begin_call(ctx);
  int rc;
  if ((rc = ensure_table(ctx))) return rc;
  if ((rc = img->enc8.ensure(4 * n))) return rc;
  {
    lg::StageScope sc(ctx, "encode_8bit", n);
    hipLaunchKernelGGL(encode_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       img->out16.as<const ushort4>(), n, ctx->encode_table.as<const uint8_t>(), order,
                       img->enc8.as<uchar4>());
  }
  LG_HIP(hipGetLastError());
  LG_HIP(hipMemcpyAsync(host, img->enc8.p, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
  LG_HIP(hipStreamSynchronize(ctx->stream));
  lg::end_call(ctx);
  return LG_OK;
}

int lg_encode_table(lg_ctx* ctx, uint8_t* table) {
  if (!ctx || !table) return fail(LG_E_INVALID, "ctx or table is NULL");
  LG_HIP(hipSetDevice(ctx->device));
  int rc;
  if ((rc = ensure_table(ctx))) return rc;
  LG_HIP(hipMemcpyAsync(table, ctx->encode_table.p, 65536, hipMemcpyDeviceToHost, ctx->stream));
  LG_HIP(hipStreamSynchronize(ctx->stream));
  return LG_OK;
}

}  // extern "C"

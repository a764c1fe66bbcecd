// lg_accum.hip — additive line accumulation on gfx950.
//
// Replaces the reference's LineList draw with additive blend into an Rgba16Float target
// (src/sub_render_pass.rs:86-95, 188-212; src/shader.wgsl:14-28; blend
// src/light_garden/mod.rs:57-73; target src/texture_renderer.rs:5). HIP has no ROP path,
// so the raster + blend becomes three stages:
//   1. binning: every segment is enumerated against the 64x64-pixel screen tiles its
//      deposits can touch. A workgroup owns a chunk of segments and counts pairs per tile
//      in an LDS histogram; the [tile][chunk] count matrix is scanned, and a second pass
//      writes segment indices into per-tile lists at deterministic offsets (no global
//      atomics).
//   2. tiles: a persistent workgroup takes a (tile, pair range) work item, walks each of
//      its segments' column (x-major) or row (y-major) samples that fall in the tile, one
//      sample per lane, and adds the two anti-aliased deposits per sample into an FP32
//      LDS tile (4 colour planes, row stride T+1 so row- and column-walks are free of bank
//      conflicts). The tile is flushed with coalesced float4 read-add-write when one item
//      owns the tile, with FP32 atomics when a heavy tile was split across items.
//   3. resolve: FP32 accumulator -> RGBA16F, round to nearest even, alpha + 1 (the
//      target's clear alpha, renderer.rs:175).
//
// The sampling rule (shared with the FP64 oracle): world -> pixel
// X = (x - left) * sx, Y = (top - y) * sy; major axis x if |dX| >= |dY|; one sample per
// major-axis pixel whose centre k + 0.5 lies in the half-open span [M0, M1); with
// s = (k + 0.5 - M0) / dM, minor m = m0 + s dm, v = m - 0.5, j = floor(v), f = v - j,
// pixels j and j + 1 on the minor axis get weights 1 - f and f; a weight w adds w*rgb to
// RGB and w*a*a to alpha. Per-sample positions are computed in FP64 from the FP32
// segment end points so that the two tiles sharing a sample agree exactly.
#include "lg_internal.h"

namespace lg {

constexpr int kTile = 64;
constexpr int kTileStride = kTile + 1;
constexpr int kPlane = kTile * kTileStride;
constexpr int kTileBlock = 1024;

struct SegGeom {
  double M0, m0, inv_dM, slope;  // slope = dm / dM
  long klo, khi;                 // sample range on the major axis (clipped to the image)
  int xmajor;
};

// Returns false when the segment deposits nothing.
__device__ __forceinline__ bool seg_geom(const float4 p, const AccumView& v, SegGeom& g) {
  double X0 = ((double)p.x - v.left) * v.sx;
  double Y0 = (v.top - (double)p.y) * v.sy;
  double X1 = ((double)p.z - v.left) * v.sx;
  double Y1 = (v.top - (double)p.w) * v.sy;
  if (!isfinite(X0) || !isfinite(Y0) || !isfinite(X1) || !isfinite(Y1)) return false;
  double dX = X1 - X0, dY = Y1 - Y0;
  int xm = fabs(dX) >= fabs(dY);
  double M0 = xm ? X0 : Y0, M1 = xm ? X1 : Y1;
  double m0 = xm ? Y0 : X0, dM = xm ? dX : dY, dm = xm ? dY : dX;
  double lo, hi;
  if (M1 > M0) {
    lo = ceil(M0 - 0.5);
    hi = ceil(M1 - 0.5) - 1.0;
  } else if (M1 < M0) {
    lo = floor(M1 - 0.5) + 1.0;
    hi = floor(M0 - 0.5);
  } else {
    return false;
  }
  double limit = (double)(xm ? v.width : v.height);
  if (lo < 0.0) lo = 0.0;
  if (hi > limit - 1.0) hi = limit - 1.0;
  if (!(lo <= hi)) return false;
  g.M0 = M0;
  g.m0 = m0;
  g.inv_dM = 1.0 / dM;
  g.slope = dm / dM;
  g.klo = (long)lo;
  g.khi = (long)hi;
  g.xmajor = xm;
  return true;
}

__device__ __forceinline__ double seg_minor(const SegGeom& g, long k) {
  return g.m0 + ((double)k + 0.5 - g.M0) * g.slope;
}

// Samples of [ka, kb] whose deposits can reach minor pixels [lo_px, hi_px] (inclusive),
// widened by one sample each side; the caller tests every deposit exactly.
__device__ __forceinline__ void minor_window(const SegGeom& g, long lo_px, long hi_px, long& ka,
                                             long& kb) {
  if (g.slope == 0.0) {
    double m = g.m0;
    if (m < (double)lo_px - 0.5 || m >= (double)hi_px + 1.5) kb = ka - 1;
    return;
  }
  // m(k) in [lo_px - 0.5, hi_px + 1.5)
  double k1 = g.M0 - 0.5 + ((double)lo_px - 0.5 - g.m0) / g.slope;
  double k2 = g.M0 - 0.5 + ((double)hi_px + 1.5 - g.m0) / g.slope;
  double kmin = k1 < k2 ? k1 : k2, kmax = k1 < k2 ? k2 : k1;
  if (kmin > (double)kb + 2.0 || kmax < (double)ka - 2.0) {
    kb = ka - 1;
    return;
  }
  long a = (long)floor(kmin) - 1, b = (long)ceil(kmax) + 1;
  if (a > ka) ka = a;
  if (b < kb) kb = b;
}

// Calls f(tile_id) for every tile a segment's deposits can touch (conservative).
template <typename F>
__device__ __forceinline__ void for_each_tile(const SegGeom& g, const AccumView& v, F&& f) {
  const int minor_limit = g.xmajor ? v.height : v.width;
  const long t_lo = g.klo / kTile, t_hi = g.khi / kTile;
  for (long tm = t_lo; tm <= t_hi; tm++) {
    long ka = g.klo > tm * kTile ? g.klo : tm * kTile;
    long kb = g.khi < tm * kTile + kTile - 1 ? g.khi : tm * kTile + kTile - 1;
    double ma = seg_minor(g, ka), mb = seg_minor(g, kb);
    double mn = ma < mb ? ma : mb, mx = ma < mb ? mb : ma;
    long r_lo = (long)floor(mn - 0.5) - 1;
    long r_hi = (long)floor(mx - 0.5) + 2;
    if (r_lo < 0) r_lo = 0;
    if (r_hi > minor_limit - 1) r_hi = minor_limit - 1;
    if (r_lo > r_hi) continue;
    for (long tn = r_lo / kTile; tn <= r_hi / kTile; tn++) {
      int tile = g.xmajor ? (int)(tn * v.tiles_x + tm) : (int)(tm * v.tiles_x + tn);
      f(tile);
    }
  }
}

// ---- binning -----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void bin_count_kernel(const float4* __restrict__ pos, uint64_t n,
                                                        AccumView v, uint32_t chunk,
                                                        uint32_t n_chunks, uint32_t* counts) {
  extern __shared__ uint32_t s_hist[];
  const int n_tiles = v.tiles_x * v.tiles_y;
  for (int t = threadIdx.x; t < n_tiles; t += blockDim.x) s_hist[t] = 0;
  __syncthreads();
  uint64_t s0 = (uint64_t)blockIdx.x * chunk;
  uint64_t s1 = s0 + chunk < n ? s0 + chunk : n;
  for (uint64_t s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
    SegGeom g;
    if (!seg_geom(pos[s], v, g)) continue;
    for_each_tile(g, v, [&](int tile) { atomicAdd(&s_hist[tile], 1u); });
  }
  __syncthreads();
  for (int t = threadIdx.x; t < n_tiles; t += blockDim.x)
    counts[(uint64_t)t * n_chunks + blockIdx.x] = s_hist[t];
}

__global__ __launch_bounds__(256) void bin_fill_kernel(const float4* __restrict__ pos, uint64_t n,
                                                       AccumView v, uint32_t chunk,
                                                       uint32_t n_chunks,
                                                       const uint32_t* __restrict__ offsets,
                                                       uint32_t* pairs) {
  extern __shared__ uint32_t s_cur[];
  const int n_tiles = v.tiles_x * v.tiles_y;
  for (int t = threadIdx.x; t < n_tiles; t += blockDim.x)
    s_cur[t] = offsets[(uint64_t)t * n_chunks + blockIdx.x];
  __syncthreads();
  uint64_t s0 = (uint64_t)blockIdx.x * chunk;
  uint64_t s1 = s0 + chunk < n ? s0 + chunk : n;
  for (uint64_t s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
    SegGeom g;
    if (!seg_geom(pos[s], v, g)) continue;
    for_each_tile(g, v, [&](int tile) {
      uint32_t slot = atomicAdd(&s_cur[tile], 1u);
      pairs[slot] = (uint32_t)s;
    });
  }
}

// tile_start[t] = offsets[t * n_chunks], tile_start[n_tiles] = low 32 bits of the total
__global__ void tile_starts_kernel(const uint32_t* offsets, uint32_t n_chunks, int n_tiles,
                                   const unsigned long long* total64, uint32_t* tile_start) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n_tiles) tile_start[t] = offsets[(uint64_t)t * n_chunks];
  if (t == n_tiles) tile_start[t] = (uint32_t)*total64;
}

// total64 += sum(in[0..n)) in 64 bits (total64 zeroed by the caller)
__global__ __launch_bounds__(256) void sum_u64_kernel(const uint32_t* in, uint64_t n,
                                                      unsigned long long* total64) {
  __shared__ unsigned long long s_w[4];
  unsigned long long acc = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    acc += in[i];
  for (int d = 32; d > 0; d >>= 1) acc += __shfl_down(acc, d, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = s_w[0] + s_w[1] + s_w[2] + s_w[3];
    if (t) atomicAdd(total64, t);
  }
}

// ---- exclusive scan of u32 (block sums kept in u32; totals are checked on the host) ------
constexpr int kScanThreads = 256;
constexpr int kScanPer = 8;
constexpr int kScanBlock = kScanThreads * kScanPer;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

// Block-level exclusive scan of kScanBlock elements; writes block total to sums[b].
__global__ __launch_bounds__(kScanThreads) void scan_block_kernel(const uint32_t* in, uint32_t* out,
                                                                  uint64_t n, uint32_t* sums) {
  __shared__ uint32_t s_w[kScanThreads / 64];
  uint64_t base = (uint64_t)blockIdx.x * kScanBlock + (uint64_t)threadIdx.x * kScanPer;
  uint32_t v[kScanPer];
  uint32_t tot = 0;
#pragma unroll
  for (int i = 0; i < kScanPer; i++) {
    v[i] = (base + i < n) ? in[base + i] : 0u;
    uint32_t x = v[i];
    v[i] = tot;
    tot += x;
  }
  uint32_t incl = wave_incl_scan(tot);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 63) s_w[wave] = incl;
  __syncthreads();
  uint32_t wofs = 0;
  for (int w = 0; w < wave; w++) wofs += s_w[w];
  uint32_t excl = wofs + incl - tot;
#pragma unroll
  for (int i = 0; i < kScanPer; i++)
    if (base + i < n) out[base + i] = v[i] + excl;
  if (threadIdx.x == kScanThreads - 1 && sums) sums[blockIdx.x] = excl + tot;
}

__global__ __launch_bounds__(kScanThreads) void scan_add_kernel(uint32_t* out, uint64_t n,
                                                                const uint32_t* block_ofs) {
  uint64_t base = (uint64_t)blockIdx.x * kScanBlock;
  uint32_t add = block_ofs[blockIdx.x];
  for (int i = threadIdx.x; i < kScanBlock; i += kScanThreads)
    if (base + i < n) out[base + i] += add;
}

// ---- tile accumulation -----------------------------------------------------------------
// The LDS tile accumulates in 64-bit fixed point: FP32 LDS atomics (ds_add_f32) run ~30x
// slower than 64-bit integer ones (ds_add_u64) on gfx950. A sample's anti-aliasing weight f
// is quantised to wq = round(f * 2^20) and the pixel pair gets (2^20 - wq) * cq and wq * cq,
// where cq = round(c * 2^G) is the channel value in fixed point: the two deposits sum to
// exactly 2^20 * cq. G is chosen per call from the largest |colour| so that no tile can
// overflow 63 bits.
constexpr int kWBits = 20;
constexpr int kWOne = 1 << kWBits;

__device__ __forceinline__ void lds_deposit(unsigned long long* tile, int row, int col, int w,
                                            const int cq[4]) {
  const int idx = row * kTileStride + col;
  atomicAdd(&tile[idx], (unsigned long long)((long long)w * (long long)cq[0]));
  atomicAdd(&tile[kPlane + idx], (unsigned long long)((long long)w * (long long)cq[1]));
  atomicAdd(&tile[2 * kPlane + idx], (unsigned long long)((long long)w * (long long)cq[2]));
  atomicAdd(&tile[3 * kPlane + idx], (unsigned long long)((long long)w * (long long)cq[3]));
}

__device__ __forceinline__ int bcast_i(int v, int lane) {
  return __builtin_amdgcn_readlane(v, lane);
}
__device__ __forceinline__ float bcast_f(float v, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}
__device__ __forceinline__ double bcast_d(double v, int lane) {
  long long b = __builtin_bit_cast(long long, v);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ int fx(float v, float scale) { return (int)__builtin_rintf(v * scale); }

// One persistent workgroup per (tile, pair range) item. Each wave takes 64 pairs at a time:
// lane l loads pair l's segment and computes its sample range inside the tile plus the
// tile-relative FP32 parameters of its samples; the wave then walks those pairs one after
// another with one sample per lane, reading the pair's setup from scalar registers.
//
// Sample positions are anchored on the tile column (the first major-axis pixel of the
// tile), which is the same for every tile that shares a sample: m(q) = m_anchor + q * slope,
// q = k - maj0, in FP32 relative to the tile's minor origin. Two tiles sharing a sample
// differ only by the FP32 rounding of their anchors (< 1e-5 px).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void tile_accum_kernel(
    const float4* __restrict__ pos, const float4* __restrict__ col0,
    const float4* __restrict__ col1, const uint32_t* __restrict__ pairs,
    const WorkItem* __restrict__ items, uint32_t n_items, uint32_t* item_counter, AccumView v,
    float4* accum, int G) {
  __shared__ unsigned long long s_tile[4 * kPlane];
  __shared__ uint32_t s_item;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  constexpr int kWaves = BLOCK / 64;
  const float cscale = __builtin_ldexpf(1.0f, G);
  const double unscale = __builtin_ldexp(1.0, -(G + kWBits));
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(item_counter, 1u);
    __syncthreads();
    uint32_t it = s_item;
    if (it >= n_items) break;
    const WorkItem W = items[it];
    for (int k = threadIdx.x; k < 4 * kPlane; k += BLOCK) s_tile[k] = 0ull;
    __syncthreads();
    const int tx = (int)(W.tile % (uint32_t)v.tiles_x);
    const int ty = (int)(W.tile / (uint32_t)v.tiles_x);
    const int x0 = tx * kTile, y0 = ty * kTile;
    for (uint32_t base = W.begin + wave * 64; base < W.end; base += kWaves * 64) {
      // ---- per-lane setup of one pair ----
      const uint32_t pi = base + lane;
      int packed = 0;  // n | q0 << 8 | xmajor << 16 | vertex colours << 17
      float m_anchor = 0.f, slope = 0.f, s_anchor = 0.f, ds = 0.f;
      float4 ca = make_float4(0.f, 0.f, 0.f, 0.f), dc = ca;
      int cq0 = 0, cq1 = 0, cq2 = 0, cq3 = 0;
      if (pi < W.end) {
        const uint32_t s = pairs[pi];
        SegGeom g;
        if (seg_geom(pos[s], v, g)) {
          const long maj0 = g.xmajor ? x0 : y0, min0 = g.xmajor ? y0 : x0;
          long a = g.klo > maj0 ? g.klo : maj0;
          long b = g.khi < maj0 + kTile - 1 ? g.khi : maj0 + kTile - 1;
          if (a <= b) minor_window(g, min0, min0 + kTile - 1, a, b);
          if (a <= b) {
            const double u0 = ((double)maj0 + 0.5) - g.M0;
            m_anchor = (float)(g.m0 + u0 * g.slope - (double)min0);
            slope = (float)g.slope;
            ca = col0[s];
            const float4 cb = col1 ? col1[s] : ca;
            dc = make_float4(cb.x - ca.x, cb.y - ca.y, cb.z - ca.z, cb.w - ca.w);
            const int vc = (dc.x != 0.f || dc.y != 0.f || dc.z != 0.f || dc.w != 0.f) ? 1 : 0;
            s_anchor = (float)(u0 * g.inv_dM);
            ds = (float)g.inv_dM;
            cq0 = fx(ca.x, cscale);
            cq1 = fx(ca.y, cscale);
            cq2 = fx(ca.z, cscale);
            cq3 = fx(ca.w * ca.w, cscale);
            packed = (int)(b - a + 1) | ((int)(a - maj0) << 8) | (g.xmajor << 16) | (vc << 17);
          }
        }
      }
      // ---- walk the pairs that have samples ----
      uint64_t todo = __ballot(packed != 0);
      while (todo) {
        const int j = __builtin_ctzll(todo);
        todo &= todo - 1;
        const int jp = bcast_i(packed, j);
        const int jn = jp & 0xff, jq0 = (jp >> 8) & 0xff, jxm = (jp >> 16) & 1, jvc = (jp >> 17) & 1;
        const float jm = bcast_f(m_anchor, j), jsl = bcast_f(slope, j);
        const long long c20_0 = (long long)bcast_i(cq0, j) << kWBits;
        const long long c20_1 = (long long)bcast_i(cq1, j) << kWBits;
        const long long c20_2 = (long long)bcast_i(cq2, j) << kWBits;
        const long long c20_3 = (long long)bcast_i(cq3, j) << kWBits;
        const int jcq0 = bcast_i(cq0, j), jcq1 = bcast_i(cq1, j), jcq2 = bcast_i(cq2, j),
                  jcq3 = bcast_i(cq3, j);
        for (int t = lane; t < jn; t += 64) {
          const int q = jq0 + t;  // major index inside the tile
          const float m = __builtin_fmaf((float)q, jsl, jm) - 0.5f;
          const float jf = __builtin_floorf(m);
          const float f = m - jf;
          const int jl = (int)jf;  // local minor index of the first deposit pixel
          int wq = (int)__builtin_fmaf(f, (float)kWOne, 0.5f);
          wq = wq > kWOne ? kWOne : wq;
          long long hi0, hi1, hi2, hi3, lo0, lo1, lo2, lo3;
          if (!jvc) {
            hi0 = (long long)wq * jcq0;
            hi1 = (long long)wq * jcq1;
            hi2 = (long long)wq * jcq2;
            hi3 = (long long)wq * jcq3;
            lo0 = c20_0 - hi0;
            lo1 = c20_1 - hi1;
            lo2 = c20_2 - hi2;
            lo3 = c20_3 - hi3;
          } else {
            const float sp = __builtin_fmaf((float)q, bcast_f(ds, j), bcast_f(s_anchor, j));
            const float al = __builtin_fmaf(sp, bcast_f(dc.w, j), bcast_f(ca.w, j));
            const int c0 = fx(__builtin_fmaf(sp, bcast_f(dc.x, j), bcast_f(ca.x, j)), cscale);
            const int c1 = fx(__builtin_fmaf(sp, bcast_f(dc.y, j), bcast_f(ca.y, j)), cscale);
            const int c2 = fx(__builtin_fmaf(sp, bcast_f(dc.z, j), bcast_f(ca.z, j)), cscale);
            const int c3 = fx(al * al, cscale);
            hi0 = (long long)wq * c0;
            hi1 = (long long)wq * c1;
            hi2 = (long long)wq * c2;
            hi3 = (long long)wq * c3;
            lo0 = ((long long)c0 << kWBits) - hi0;
            lo1 = ((long long)c1 << kWBits) - hi1;
            lo2 = ((long long)c2 << kWBits) - hi2;
            lo3 = ((long long)c3 << kWBits) - hi3;
          }
          // deposits: (1 - f) at minor jl, f at minor jl + 1
          const int i0 = jxm ? (jl * kTileStride + q) : (q * kTileStride + jl);
          const int step = jxm ? kTileStride : 1;
          if (jl >= 0 && jl < kTile) {
            atomicAdd(&s_tile[i0], (unsigned long long)lo0);
            atomicAdd(&s_tile[kPlane + i0], (unsigned long long)lo1);
            atomicAdd(&s_tile[2 * kPlane + i0], (unsigned long long)lo2);
            atomicAdd(&s_tile[3 * kPlane + i0], (unsigned long long)lo3);
          }
          if (jl + 1 >= 0 && jl + 1 < kTile) {
            const int i1 = i0 + step;
            atomicAdd(&s_tile[i1], (unsigned long long)hi0);
            atomicAdd(&s_tile[kPlane + i1], (unsigned long long)hi1);
            atomicAdd(&s_tile[2 * kPlane + i1], (unsigned long long)hi2);
            atomicAdd(&s_tile[3 * kPlane + i1], (unsigned long long)hi3);
          }
        }
      }
    }
    __syncthreads();
    // flush: fixed point -> FP32 accumulator
    for (int k = threadIdx.x; k < kTile * kTile; k += BLOCK) {
      int r = k / kTile, cc = k % kTile;
      int gx = x0 + cc, gy = y0 + r;
      if (gx >= v.width || gy >= v.height) continue;
      int idx = r * kTileStride + cc;
      float4 d = make_float4((float)((double)(long long)s_tile[idx] * unscale),
                             (float)((double)(long long)s_tile[kPlane + idx] * unscale),
                             (float)((double)(long long)s_tile[2 * kPlane + idx] * unscale),
                             (float)((double)(long long)s_tile[3 * kPlane + idx] * unscale));
      float4* dst = accum + (size_t)gy * v.width + gx;
      if (W.split) {
        if (d.x != 0.0f) atomicAdd(&dst->x, d.x);
        if (d.y != 0.0f) atomicAdd(&dst->y, d.y);
        if (d.z != 0.0f) atomicAdd(&dst->z, d.z);
        if (d.w != 0.0f) atomicAdd(&dst->w, d.w);
      } else {
        float4 o = *dst;
        o.x += d.x;
        o.y += d.y;
        o.z += d.z;
        o.w += d.w;
        *dst = o;
      }
    }
    __syncthreads();
  }
}

// ---- resolve / clear ----------------------------------------------------------------------
__global__ void resolve_kernel(const float4* __restrict__ accum, ushort4* __restrict__ out,
                               uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 a = accum[i];
  _Float16 r = (_Float16)a.x, g = (_Float16)a.y, b = (_Float16)a.z, al = (_Float16)(a.w + 1.0f);
  ushort4 o;
  o.x = __builtin_bit_cast(unsigned short, r);
  o.y = __builtin_bit_cast(unsigned short, g);
  o.z = __builtin_bit_cast(unsigned short, b);
  o.w = __builtin_bit_cast(unsigned short, al);
  out[i] = o;
}

__global__ void fill_zero_kernel(float4* p, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// ---- host launchers ---------------------------------------------------------------------
int tile_size() { return kTile; }

hipError_t launch_bin_count(const float4* pos, uint64_t n, const AccumView& v, uint32_t chunk,
                            uint32_t n_chunks, uint32_t* counts, hipStream_t st) {
  size_t lds = (size_t)v.tiles_x * v.tiles_y * sizeof(uint32_t);
  hipLaunchKernelGGL(bin_count_kernel, dim3(n_chunks), dim3(256), lds, st, pos, n, v, chunk,
                     n_chunks, counts);
  return hipGetLastError();
}

hipError_t launch_bin_fill(const float4* pos, uint64_t n, const AccumView& v, uint32_t chunk,
                           uint32_t n_chunks, const uint32_t* offsets, uint32_t* pairs,
                           hipStream_t st) {
  size_t lds = (size_t)v.tiles_x * v.tiles_y * sizeof(uint32_t);
  hipLaunchKernelGGL(bin_fill_kernel, dim3(n_chunks), dim3(256), lds, st, pos, n, v, chunk,
                     n_chunks, offsets, pairs);
  return hipGetLastError();
}

hipError_t launch_tile_starts(const uint32_t* offsets, uint32_t n_chunks, int n_tiles,
                              const unsigned long long* total64, uint32_t* tile_start,
                              hipStream_t st) {
  int n = n_tiles + 1;
  hipLaunchKernelGGL(tile_starts_kernel, dim3((n + 255) / 256), dim3(256), 0, st, offsets,
                     n_chunks, n_tiles, total64, tile_start);
  return hipGetLastError();
}

hipError_t launch_sum_u64(const uint32_t* in, uint64_t n, unsigned long long* total64,
                          hipStream_t st) {
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(sum_u64_kernel, dim3((unsigned)blocks), dim3(256), 0, st, in, n, total64);
  return hipGetLastError();
}

uint64_t scan_scratch_elems(uint64_t n) {
  uint64_t total = 0;
  while (n > 1) {
    uint64_t nb = (n + kScanBlock - 1) / kScanBlock;
    total += nb;
    n = nb;
  }
  return total + 1;
}

// Exclusive scan in place over out (in may equal out). scratch holds
// scan_scratch_elems(n) u32.
hipError_t launch_scan_exclusive(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* scratch,
                                 hipStream_t st) {
  if (n == 0) return hipSuccess;
  uint64_t nb = (n + kScanBlock - 1) / kScanBlock;
  uint32_t* sums = scratch;
  hipLaunchKernelGGL(scan_block_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, out, n,
                     sums);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || nb == 1) return e;
  e = launch_scan_exclusive(sums, sums, nb, scratch + nb, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, st, out, n, sums);
  return hipGetLastError();
}

hipError_t launch_tile_accum(const float4* pos, const float4* col0, const float4* col1,
                             const uint32_t* pairs, const WorkItem* items, uint32_t n_items,
                             uint32_t* item_counter, const AccumView& v, float4* accum, int grid,
                             int fixed_bits, hipStream_t st) {
  hipLaunchKernelGGL((tile_accum_kernel<kTileBlock>), dim3(grid), dim3(kTileBlock), 0, st, pos, col0,
                     col1, pairs, items, n_items, item_counter, v, accum, fixed_bits);
  return hipGetLastError();
}

int tile_block() { return kTileBlock; }
int weight_bits() { return kWBits; }

hipError_t launch_resolve(const float4* accum, ushort4* out, uint64_t n, hipStream_t st) {
  hipLaunchKernelGGL(resolve_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, accum,
                     out, n);
  return hipGetLastError();
}

hipError_t launch_fill_zero(float4* p, uint64_t n, hipStream_t st) {
  hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

}  // namespace lg

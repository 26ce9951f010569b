// Depth processing kernels: bilateral filter, 2x2 downsample, vertex and
// normal maps (reference: include/kinectfusion/depth_processing.hpp:54-284,
// src/kinectfusion/depth_processing.cpp:97-163).
//
// Layout: one 32 x 8 output tile per 256-thread workgroup (four wave64 rows
// of 32 x 2).  Each workgroup stages its input tile plus halo in LDS once and
// every tap of the stencil reads LDS instead of global memory.
#include "kf_exp.hpp"
#include "kf_internal.hpp"

namespace kf {
namespace {

__constant__ double kExp2J64[64] = KF_EXP2_J64_TABLE;

constexpr int kTx = 32;
constexpr int kTy = 8;
constexpr int kMaxRadius = 8;

// Bilateral filter (BilateralFilter::operator(), depth_processing.hpp:70-114):
// clamped (not mirrored) window, samples outside [min, max] and zeros ignored,
// weight exp(d^2 * ss + dm^2 * sr) evaluated as one exponential of the float
// exponent.  The exponential is taken in FP64 (Tang's table-driven method,
// relative error below 4e-16, the table staged in LDS) and rounded to
// float, which matches a correctly rounded expf.
//
// R > 0: the window's radius fixed at compile time and each row's taps
// unrolled, so their exponentials overlap; the sums still add the taps in the
// reference's row-major order, an unusable tap (outside the image or without
// depth, both staged as -1) adding +0, which leaves them unchanged (they are
// never -0).  R = 0: the radius is a run-time value.
//
// stage (optional): each output pixel's raw depth and colour are also copied
// into the staging buffers a captured frame reads, in the same pass.
template <int R>
__global__ void __launch_bounds__(256) bilateral_kernel(const uint16_t* __restrict__ in,
                                                         uint16_t* __restrict__ out, int w, int h,
                                                         float scale, float lo, float hi, int radius_arg,
                                                         float ss, float sr, StageCopy stage) {
  const int radius = R > 0 ? R : radius_arg;
  __shared__ float meters[(kTy + 2 * kMaxRadius) * (kTx + 2 * kMaxRadius)];
  __shared__ float raws[(kTy + 2 * kMaxRadius) * (kTx + 2 * kMaxRadius)];
  __shared__ double exp_table[64];
  const int tw = kTx + 2 * radius, th = kTy + 2 * radius;
  const int x0 = blockIdx.x * kTx - radius, y0 = blockIdx.y * kTy - radius;
  const int tid = threadIdx.y * kTx + threadIdx.x;
  if (tid < 64) exp_table[tid] = kExp2J64[tid];
  for (int i = tid; i < tw * th; i += kTx * kTy) {
    const int lx = i % tw, ly = i / tw;
    const int gx = x0 + lx, gy = y0 + ly;
    float m = -1.0F, r = 0.0F;
    if (gx >= 0 && gy >= 0 && gx < w && gy < h) {
      const uint16_t raw = in[static_cast<size_t>(gy) * w + gx];
      float mm;
      if (depth_in_range(raw, scale, lo, hi, mm)) {
        m = mm;
        r = static_cast<float>(raw);
      }
    }
    meters[i] = m;  // -1 marks an unusable sample
    raws[i] = r;
  }
  __syncthreads();
  const int col = blockIdx.x * kTx + threadIdx.x, row = blockIdx.y * kTy + threadIdx.y;
  if (col >= w || row >= h) return;
  if (stage.depth != nullptr) {
    const size_t i = static_cast<size_t>(row) * w + col;
    stage.depth[i] = in[i];
    if (stage.color_in != nullptr) stage.color[i] = stage.color_in[i];
  }
  const int cl = (threadIdx.y + radius) * tw + threadIdx.x + radius;
  const float cm = meters[cl];
  uint16_t result = 0;
  if (R > 0 && cm >= 0.0F) {
    float wsum_v = 0.0F, wsum = 0.0F;
#pragma unroll 1
    for (int dy = -R; dy <= R; ++dy) {
      const int base = (threadIdx.y + R + dy) * tw + threadIdx.x + R;
#pragma unroll
      for (int dx = -R; dx <= R; ++dx) {
        const float sm = meters[base + dx];
        const float pd2 = static_cast<float>((dx * dx) + (dy * dy));
        const float dd = sm - cm;
        const float e = (pd2 * ss) + (dd * dd * sr);
        const float wt = static_cast<float>(exp_nonpositive(static_cast<double>(e), exp_table));
        const bool use = sm >= 0.0F;
        wsum_v += use ? wt * raws[base + dx] : 0.0F;
        wsum += use ? wt : 0.0F;
      }
    }
    if (wsum > 0.0F) result = static_cast<uint16_t>(static_cast<long>(roundf(wsum_v / wsum)));
  } else if (cm >= 0.0F) {
    const int xb = max(col - radius, 0), yb = max(row - radius, 0);
    const int xe = min(col + radius, w - 1), ye = min(row + radius, h - 1);
    float wsum_v = 0.0F, wsum = 0.0F;
    for (int y = yb; y <= ye; ++y) {
      const int base = (y - y0) * tw - x0;
      for (int x = xb; x <= xe; ++x) {
        const float sm = meters[base + x];
        if (sm < 0.0F) continue;
        const float pd2 =
            static_cast<float>(((x - col) * (x - col)) + ((y - row) * (y - row)));
        const float dd = sm - cm;
        const float e = (pd2 * ss) + (dd * dd * sr);
        const float wt = static_cast<float>(exp_nonpositive(static_cast<double>(e), exp_table));
        wsum_v += wt * raws[base + x];
        wsum += wt;
      }
    }
    if (wsum > 0.0F) result = static_cast<uint16_t>(static_cast<long>(roundf(wsum_v / wsum)));
  }
  out[static_cast<size_t>(row) * w + col] = result;
}

// BlockDownsample (depth_processing.hpp:152-187): the rounded mean of the
// non-zero samples of a 2 x 2 block, or 0 when they span more than max_jump.
__device__ inline uint16_t downsample_block(uint32_t a, uint32_t b, uint32_t c, uint32_t d, float scale,
                                            float max_jump) {
  const uint32_t s4[4] = {a, b, c, d};
  uint32_t sum = 0, count = 0;
  uint32_t mn = 0xFFFFU, mx = 0U;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t v = s4[i];
    if (v == 0U) continue;
    sum += v;
    ++count;
    mn = v < mn ? v : mn;
    mx = mx < v ? v : mx;
  }
  uint16_t r = 0;
  if (count != 0U) {
    const float hi_m = static_cast<float>(mx) / scale;
    const float lo_m = static_cast<float>(mn) / scale;
    if (!(hi_m - lo_m > max_jump)) r = static_cast<uint16_t>((sum + (count / 2U)) / count);
  }
  return r;
}

__global__ void __launch_bounds__(256) downsample_kernel(const uint16_t* __restrict__ in,
                                                          uint16_t* __restrict__ out, int w, int h,
                                                          float scale, float max_jump) {
  const int ow = w / 2, oh = h / 2;
  const int col = blockIdx.x * kTx + threadIdx.x, row = blockIdx.y * kTy + threadIdx.y;
  if (col >= ow || row >= oh) return;
  const uint16_t* r0 = in + static_cast<size_t>(row * 2) * w + col * 2;
  out[static_cast<size_t>(row) * ow + col] = downsample_block(r0[0], r0[1], r0[w], r0[w + 1], scale, max_jump);
}

// Vertex map (VertexProjection with the identity pose,
// depth_processing.hpp:215-226) and normal map (NormalEstimation,
// :251-279) in one pass: the workgroup back-projects its tile plus a
// one-pixel halo into LDS, then takes the 4-neighbour cross product.
__device__ inline V3 vertex_of(uint16_t raw, int col, int row, const Intr& k, float scale, float lo,
                               float hi) {
  float d;
  if (!depth_in_range(raw, scale, lo, hi, d)) return nan3();
  const V3 p = back_project(k, static_cast<float>(col), static_cast<float>(row), d);
  // identity camera pose applied as the reference does (R p + t)
  const Rigid id{mk3(1.0F, 0.0F, 0.0F), mk3(0.0F, 1.0F, 0.0F), mk3(0.0F, 0.0F, 1.0F),
                 mk3(0.0F, 0.0F, 0.0F)};
  return xform(id, p);
}

// NormalEstimation's interior rule: the cross product of the vertical and
// horizontal central differences, unless a neighbour is missing or a
// difference jumps by more than max_jump in depth.
__device__ inline V3 normal_from(V3 c, V3 l, V3 r, V3 t, V3 b, float max_jump) {
  V3 n = nan3();
  if (finite3(add3(add3(add3(add3(c, l), r), t), b)) && !(fabsf(r.z - l.z) > max_jump) &&
      !(fabsf(b.z - t.z) > max_jump)) {
    const V3 cr = cross3(sub3(b, t), sub3(r, l));
    const float len = norm3(cr);
    if (len > 0.0F) n = div3(cr, len);
  }
  return n;
}

__global__ void __launch_bounds__(256) vertex_normal_kernel(const uint16_t* __restrict__ depth,
                                                             float4* __restrict__ vert,
                                                             float4* __restrict__ norm, int w, int h,
                                                             Intr k, float scale, float lo, float hi,
                                                             float max_jump) {
  __shared__ float sx[(kTy + 2) * (kTx + 2)];
  __shared__ float sy[(kTy + 2) * (kTx + 2)];
  __shared__ float sz[(kTy + 2) * (kTx + 2)];
  const int tw = kTx + 2, th = kTy + 2;
  const int x0 = blockIdx.x * kTx - 1, y0 = blockIdx.y * kTy - 1;
  const int tid = threadIdx.y * kTx + threadIdx.x;
  for (int i = tid; i < tw * th; i += kTx * kTy) {
    const int gx = x0 + i % tw, gy = y0 + i / tw;
    V3 v = nan3();
    if (gx >= 0 && gy >= 0 && gx < w && gy < h)
      v = vertex_of(depth[static_cast<size_t>(gy) * w + gx], gx, gy, k, scale, lo, hi);
    sx[i] = v.x;
    sy[i] = v.y;
    sz[i] = v.z;
  }
  __syncthreads();
  const int col = blockIdx.x * kTx + threadIdx.x, row = blockIdx.y * kTy + threadIdx.y;
  if (col >= w || row >= h) return;
  auto at = [&](int dx, int dy) {
    const int i = (threadIdx.y + 1 + dy) * tw + threadIdx.x + 1 + dx;
    return mk3(sx[i], sy[i], sz[i]);
  };
  const V3 c = at(0, 0);
  V3 n = nan3();
  if (!(col == 0 || row == 0 || col + 1 >= w || row + 1 >= h))
    n = normal_from(c, at(-1, 0), at(1, 0), at(0, -1), at(0, 1), max_jump);
  const size_t i = static_cast<size_t>(row) * w + col;
  vert[i] = st3(c);
  norm[i] = st3(n);
}

// The three-level pyramid after the bilateral pass in one launch: level-0
// vertex and normal maps, the two 2 x 2 downsamples and their vertex and
// normal maps.  A workgroup owns an 8 x 4 tile of level 2, so 16 x 8 of
// level 1 and 32 x 16 of level 0; it stages level-0 depth with a 4-pixel
// halo in LDS, derives level 1 with a 2-pixel halo and level 2 with a
// 1-pixel halo from it (the same block rule as downsample_kernel), then
// each level's vertices with a 1-pixel halo and the normals of its tile.
// Every value is computed by the same arithmetic as the per-level kernels.
constexpr int kP2x = 8, kP2y = 4;                           // level-2 tile
constexpr int kP0w = 4 * kP2x + 8, kP0h = 4 * kP2y + 8;     // level-0 depth staged: 40 x 24
constexpr int kP1w = 2 * kP2x + 4, kP1h = 2 * kP2y + 4;     // level-1 depth: 20 x 12
constexpr int kP2w = kP2x + 2, kP2h = kP2y + 2;             // level-2 depth: 10 x 6
constexpr int kV0w = 4 * kP2x + 2, kV0h = 4 * kP2y + 2;     // level-0 vertices: 34 x 18
constexpr int kV1w = 2 * kP2x + 2, kV1h = 2 * kP2y + 2;     // level-1 vertices: 18 x 10

struct PyrLevel {
  uint16_t* depth;
  float4* vert;
  float4* norm;
  int w, h;
  Intr k;
};

__global__ void __launch_bounds__(256) pyramid3_kernel(const uint16_t* __restrict__ d0, PyrLevel l0, PyrLevel l1,
                                                        PyrLevel l2, float scale, float lo, float hi,
                                                        float ds_jump, float n_jump) {
  __shared__ uint16_t s0[kP0h * kP0w];
  __shared__ uint16_t s1[kP1h * kP1w];
  __shared__ uint16_t s2[kP2h * kP2w];
  __shared__ float v0[3][kV0h * kV0w];
  __shared__ float v1[3][kV1h * kV1w];
  __shared__ float v2[3][kP2h * kP2w];
  const int tid = threadIdx.x;
  const int X2 = blockIdx.x * kP2x, Y2 = blockIdx.y * kP2y;
  const int X1 = 2 * X2, Y1 = 2 * Y2, X0 = 4 * X2, Y0 = 4 * Y2;
  // level 0 depth, origin (X0 - 4, Y0 - 4); outside the image reads 0
  for (int i = tid; i < kP0h * kP0w; i += 256) {
    const int gx = X0 - 4 + i % kP0w, gy = Y0 - 4 + i / kP0w;
    s0[i] = (gx >= 0 && gy >= 0 && gx < l0.w && gy < l0.h) ? d0[static_cast<size_t>(gy) * l0.w + gx] : 0;
  }
  __syncthreads();
  // level 1, origin (X1 - 2, Y1 - 2): its 2 x 2 level-0 block starts at s0 (2 lx, 2 ly)
  for (int i = tid; i < kP1h * kP1w; i += 256) {
    const int lx = i % kP1w, ly = i / kP1w, gx = X1 - 2 + lx, gy = Y1 - 2 + ly;
    uint16_t r = 0;
    if (gx >= 0 && gy >= 0 && gx < l1.w && gy < l1.h) {
      const uint16_t* b = s0 + (2 * ly) * kP0w + 2 * lx;
      r = downsample_block(b[0], b[1], b[kP0w], b[kP0w + 1], scale, ds_jump);
      if (lx >= 2 && ly >= 2 && lx < 2 + 2 * kP2x && ly < 2 + 2 * kP2y) l1.depth[static_cast<size_t>(gy) * l1.w + gx] = r;
    }
    s1[i] = r;
  }
  __syncthreads();
  // level 2, origin (X2 - 1, Y2 - 1): its block starts at s1 (2 lx, 2 ly)
  for (int i = tid; i < kP2h * kP2w; i += 256) {
    const int lx = i % kP2w, ly = i / kP2w, gx = X2 - 1 + lx, gy = Y2 - 1 + ly;
    uint16_t r = 0;
    if (gx >= 0 && gy >= 0 && gx < l2.w && gy < l2.h) {
      const uint16_t* b = s1 + (2 * ly) * kP1w + 2 * lx;
      r = downsample_block(b[0], b[1], b[kP1w], b[kP1w + 1], scale, ds_jump);
      if (lx >= 1 && ly >= 1 && lx < 1 + kP2x && ly < 1 + kP2y) l2.depth[static_cast<size_t>(gy) * l2.w + gx] = r;
    }
    s2[i] = r;
  }
  __syncthreads();
  // vertices with a one-pixel halo (NaN outside the image)
  auto vertex_at = [&](const uint16_t* sd, int sw, int sx, int sy, int gx, int gy, const PyrLevel& L) {
    if (gx < 0 || gy < 0 || gx >= L.w || gy >= L.h) return nan3();
    return vertex_of(sd[sy * sw + sx], gx, gy, L.k, scale, lo, hi);
  };
  for (int i = tid; i < kV0h * kV0w; i += 256) {
    const int lx = i % kV0w, ly = i / kV0w;
    const V3 v = vertex_at(s0, kP0w, lx + 3, ly + 3, X0 - 1 + lx, Y0 - 1 + ly, l0);
    v0[0][i] = v.x;
    v0[1][i] = v.y;
    v0[2][i] = v.z;
  }
  for (int i = tid; i < kV1h * kV1w; i += 256) {
    const int lx = i % kV1w, ly = i / kV1w;
    const V3 v = vertex_at(s1, kP1w, lx + 1, ly + 1, X1 - 1 + lx, Y1 - 1 + ly, l1);
    v1[0][i] = v.x;
    v1[1][i] = v.y;
    v1[2][i] = v.z;
  }
  for (int i = tid; i < kP2h * kP2w; i += 256) {
    const int lx = i % kP2w, ly = i / kP2w;
    const V3 v = vertex_at(s2, kP2w, lx, ly, X2 - 1 + lx, Y2 - 1 + ly, l2);
    v2[0][i] = v.x;
    v2[1][i] = v.y;
    v2[2][i] = v.z;
  }
  __syncthreads();
  // each level's tile: vertex and normal
  auto tile = [&](float* vx, float* vy, float* vz, int vw, int tw, int th, int X, int Y, const PyrLevel& L) {
    for (int i = tid; i < tw * th; i += 256) {
      const int tx = i % tw, ty = i / tw, col = X + tx, row = Y + ty;
      if (col >= L.w || row >= L.h) continue;
      auto at = [&](int dx, int dy) {
        const int j = (ty + 1 + dy) * vw + tx + 1 + dx;
        return mk3(vx[j], vy[j], vz[j]);
      };
      const V3 c = at(0, 0);
      V3 n = nan3();
      if (!(col == 0 || row == 0 || col + 1 >= L.w || row + 1 >= L.h))
        n = normal_from(c, at(-1, 0), at(1, 0), at(0, -1), at(0, 1), n_jump);
      const size_t o = static_cast<size_t>(row) * L.w + col;
      L.vert[o] = st3(c);
      L.norm[o] = st3(n);
    }
  };
  tile(v0[0], v0[1], v0[2], kV0w, 4 * kP2x, 4 * kP2y, X0, Y0, l0);
  tile(v1[0], v1[1], v1[2], kV1w, 2 * kP2x, 2 * kP2y, X1, Y1, l1);
  tile(v2[0], v2[1], v2[2], kP2w, kP2x, kP2y, X2, Y2, l2);
}

}  // namespace

bool bilateral_active(const kf_config& c) { return c.bilateral_filter != 0 && c.bilateral_radius != 0; }

void launch_bilateral(kf_context* ctx, const uint16_t* raw, StageCopy stage) {
  const kf_config& c = ctx->cfg;
  LevelBuffers& L = ctx->lv[0];
  if (!bilateral_active(c)) {
    if (stage.depth != nullptr) fail(KF_ERR_STATE, "launch_bilateral: staging needs the filter pass");
    KF_HIP(hipMemcpyAsync(L.depth, raw, sizeof(uint16_t) * L.w * L.h, hipMemcpyDeviceToDevice,
                          ctx->stream));
    return;
  }
  const dim3 block(kTx, kTy), grid((L.w + kTx - 1) / kTx, (L.h + kTy - 1) / kTy);
  const float ss = -0.5F / (c.bilateral_spatial_sigma * c.bilateral_spatial_sigma);
  const float sr = -0.5F / (c.bilateral_depth_sigma * c.bilateral_depth_sigma);
  if (c.bilateral_radius == 3)
    hipLaunchKernelGGL(bilateral_kernel<3>, grid, block, 0, ctx->stream, raw, L.depth, L.w, L.h,
                     c.depth_scale, c.min_depth, c.max_depth, c.bilateral_radius, ss, sr, stage);
  else
    hipLaunchKernelGGL(bilateral_kernel<0>, grid, block, 0, ctx->stream, raw, L.depth, L.w, L.h,
                     c.depth_scale, c.min_depth, c.max_depth, c.bilateral_radius, ss, sr, stage);
  KF_HIP(hipGetLastError());
}

bool launch_pyramid3(kf_context* ctx) {
  const kf_config& c = ctx->cfg;
  if (c.levels != 3 || c.width % 4 != 0 || c.height % 4 != 0) return false;
  auto level = [&](int l) {
    LevelBuffers& L = ctx->lv[l];
    return PyrLevel{L.depth, L.vert, L.norm, L.w, L.h, L.k};
  };
  const PyrLevel l2 = level(2);
  const dim3 grid((l2.w + kP2x - 1) / kP2x, (l2.h + kP2y - 1) / kP2y);
  hipLaunchKernelGGL(pyramid3_kernel, grid, dim3(256), 0, ctx->stream, ctx->lv[0].depth, level(0), level(1), l2,
                     c.depth_scale, c.min_depth, c.max_depth, c.max_downsample_depth_jump,
                     c.max_normal_depth_jump);
  KF_HIP(hipGetLastError());
  return true;
}

void launch_level_maps(kf_context* ctx, int level) {
  const kf_config& c = ctx->cfg;
  LevelBuffers& L = ctx->lv[level];
  if (level > 0) {
    const LevelBuffers& P = ctx->lv[level - 1];
    const dim3 block(kTx, kTy), grid((L.w + kTx - 1) / kTx, (L.h + kTy - 1) / kTy);
    hipLaunchKernelGGL(downsample_kernel, grid, block, 0, ctx->stream, P.depth, L.depth, P.w, P.h,
                       c.depth_scale, c.max_downsample_depth_jump);
    KF_HIP(hipGetLastError());
  }
  const dim3 block(kTx, kTy), grid((L.w + kTx - 1) / kTx, (L.h + kTy - 1) / kTy);
  hipLaunchKernelGGL(vertex_normal_kernel, grid, block, 0, ctx->stream, L.depth, L.vert, L.norm, L.w,
                     L.h, L.k, c.depth_scale, c.min_depth, c.max_depth, c.max_normal_depth_jump);
  KF_HIP(hipGetLastError());
}

}  // namespace kf

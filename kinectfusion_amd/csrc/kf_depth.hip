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
__global__ void __launch_bounds__(256) bilateral_kernel(const uint16_t* __restrict__ in,
                                                         uint16_t* __restrict__ out, int w, int h,
                                                         float scale, float lo, float hi, int radius,
                                                         float ss, float sr) {
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
  const int cl = (threadIdx.y + radius) * tw + threadIdx.x + radius;
  const float cm = meters[cl];
  uint16_t result = 0;
  if (cm >= 0.0F) {
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

// BlockDownsample (depth_processing.hpp:152-187).
__global__ void __launch_bounds__(256) downsample_kernel(const uint16_t* __restrict__ in,
                                                          uint16_t* __restrict__ out, int w, int h,
                                                          float scale, float max_jump) {
  const int ow = w / 2, oh = h / 2;
  const int col = blockIdx.x * kTx + threadIdx.x, row = blockIdx.y * kTy + threadIdx.y;
  if (col >= ow || row >= oh) return;
  uint32_t sum = 0, count = 0;
  uint32_t mn = 0xFFFFU, mx = 0U;
  for (int dy = 0; dy < 2; ++dy)
    for (int dx = 0; dx < 2; ++dx) {
      const uint32_t s = in[static_cast<size_t>(row * 2 + dy) * w + (col * 2 + dx)];
      if (s == 0U) continue;
      sum += s;
      ++count;
      mn = s < mn ? s : mn;
      mx = mx < s ? s : mx;
    }
  uint16_t r = 0;
  if (count != 0U) {
    const float hi_m = static_cast<float>(mx) / scale;
    const float lo_m = static_cast<float>(mn) / scale;
    if (!(hi_m - lo_m > max_jump)) r = static_cast<uint16_t>((sum + (count / 2U)) / count);
  }
  out[static_cast<size_t>(row) * ow + col] = r;
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
  if (!(col == 0 || row == 0 || col + 1 >= w || row + 1 >= h)) {
    const V3 l = at(-1, 0), r = at(1, 0), t = at(0, -1), b = at(0, 1);
    if (finite3(add3(add3(add3(add3(c, l), r), t), b)) && !(fabsf(r.z - l.z) > max_jump) &&
        !(fabsf(b.z - t.z) > max_jump)) {
      const V3 cr = cross3(sub3(b, t), sub3(r, l));
      const float len = norm3(cr);
      if (len > 0.0F) n = div3(cr, len);
    }
  }
  const size_t i = static_cast<size_t>(row) * w + col;
  vert[i] = st3(c);
  norm[i] = st3(n);
}

}  // namespace

void launch_bilateral(kf_context* ctx, const uint16_t* raw) {
  const kf_config& c = ctx->cfg;
  LevelBuffers& L = ctx->lv[0];
  if (!c.bilateral_filter || c.bilateral_radius == 0) {
    KF_HIP(hipMemcpyAsync(L.depth, raw, sizeof(uint16_t) * L.w * L.h, hipMemcpyDeviceToDevice,
                          ctx->stream));
    return;
  }
  const dim3 block(kTx, kTy), grid((L.w + kTx - 1) / kTx, (L.h + kTy - 1) / kTy);
  const float ss = -0.5F / (c.bilateral_spatial_sigma * c.bilateral_spatial_sigma);
  const float sr = -0.5F / (c.bilateral_depth_sigma * c.bilateral_depth_sigma);
  hipLaunchKernelGGL(bilateral_kernel, grid, block, 0, ctx->stream, raw, L.depth, L.w, L.h,
                     c.depth_scale, c.min_depth, c.max_depth, c.bilateral_radius, ss, sr);
  KF_HIP(hipGetLastError());
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

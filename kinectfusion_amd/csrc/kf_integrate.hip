// Dense TSDF integration (reference: IntegrationContext::observe/fuse and the
// Classic / AngleWeighted rules, include/kinectfusion/tsdf_integration.hpp:
// 183-342; full sweep :356-370; BasicVoxel::fused, volume.hpp:137-146).
//
// Three kernels per frame:
//  1. pixel_prep  -- everything observe() derives from the pixel alone:
//     measured depth, |K^-1 [u v 1]| (lambda), the rule's observation weight
//     (angle rule: cos(theta)/depth from the level-0 normal), the truncation
//     band; plus the maximum usable depth of every 16 x 16 pixel tile.
//  2. brick_cull  -- one thread per 64 x 4 x 4 voxel brick: a brick is kept
//     unless no voxel of it can project into the image in front of the
//     camera, or its nearest voxel lies farther than the deepest depth of
//     the tiles it covers plus the truncation band.  The bounds carry
//     margins so that no voxel the full sweep would fuse is ever culled: the
//     result is bit-identical to the full sweep.  Kept bricks are compacted
//     into a list with one atomic per wave.
//  3. integrate   -- persistent workgroups walk the list; a wave covers one
//     64-voxel x-row of a brick, so voxel loads and stores are contiguous
//     (512 B per wave for 8 B voxels, 256 B for 4 B voxels).
#include "kf_internal.hpp"

namespace kf {
namespace {

struct FrameParams {
  Rigid w2c;
  Intr k;
  int w, h;
  float scale, lo, hi;
  float obs_weight, max_weight;
  int projective;
  int dist_scaled;
  float trunc_scale;
  int angle_rule;
  const IcpState* st;  // when set: w2c from the running device pose; skip on failure
};

// world_to_camera of this frame, computed once per workgroup.
__device__ inline bool block_w2c(const FrameParams& f, Rigid& out) {
  __shared__ float xf[12];
  __shared__ int skip;
  if (f.st == nullptr) {
    out = f.w2c;
    return true;
  }
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    skip = f.st->frame_failed;
    double inv[16];
    rigid_inverse64(f.st->pose, inv);
    for (int i = 0; i < 12; ++i) xf[i] = static_cast<float>(inv[i]);
  }
  __syncthreads();
  out = rigid_from12(xf);
  return skip == 0;
}

// 1. per-pixel record {depth m (or -1), lambda, weight, truncation}.
__global__ void __launch_bounds__(256) pixel_prep_kernel(const uint16_t* __restrict__ depth,
                                                          const float4* __restrict__ normals,
                                                          float4* __restrict__ pix,
                                                          float* __restrict__ tile_max,
                                                          FrameParams f, float base_trunc,
                                                          uint32_t* __restrict__ counters,
                                                          unsigned long long* __restrict__ fused) {
  __shared__ float red[256];
  // the brick and fused counters start the pass at zero (read by later launches only)
  if (blockIdx.x == 0 && blockIdx.y == 0) {
    if (threadIdx.y == 0 && threadIdx.x < 4) counters[threadIdx.x] = 0U;
    if (threadIdx.y == 0 && threadIdx.x < 2) fused[threadIdx.x] = 0ULL;
  }
  const int col = blockIdx.x * kTile + threadIdx.x, row = blockIdx.y * kTile + threadIdx.y;
  const int tid = threadIdx.y * kTile + threadIdx.x;
  float dmax = 0.0F;
  if (col < f.w && row < f.h) {
    const size_t i = static_cast<size_t>(row) * f.w + col;
    float d;
    float4 rec = make_float4(-1.0F, 0.0F, 0.0F, 0.0F);
    if (depth_in_range(depth[i], f.scale, f.lo, f.hi, d)) {
      const V3 ray = back_project(f.k, static_cast<float>(col), static_cast<float>(row), 1.0F);
      const float lambda = norm3(ray);
      if (finite3(ray) && lambda != 0.0F) {
        float w = f.obs_weight;
        if (f.angle_rule && normals != nullptr) {
          const V3 n = ld3(normals[i]);
          if (finite3(n)) {
            const V3 nr = normalize3(ray);
            const V3 view = mk3(-nr.x, -nr.y, -nr.z);
            const float cos_theta = fmax_ref(0.0F, dot3(normalize3(n), view));
            w *= cos_theta / d;
          }
        }
        const float trunc = f.dist_scaled ? base_trunc + (f.trunc_scale * d) : base_trunc;
        rec = make_float4(d, lambda, w, trunc);
        dmax = d;
      }
    }
    pix[i] = rec;
  }
  red[tid] = dmax;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) red[tid] = fmaxf(red[tid], red[tid + s]);
    __syncthreads();
  }
  if (tid == 0) tile_max[blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// Coarser depth-maximum levels for the cull: 4 x 4 tiles of 16 px (64 px)
// and 16 x 16 tiles (256 px), plus the frame maximum.  One workgroup.
__global__ void __launch_bounds__(256) tile_pyramid_kernel(float* __restrict__ tile_max, int tiles_x, int tiles_y) {
  float* l1 = tile_max + tiles_x * tiles_y;
  const int t1x = (tiles_x + 3) / 4, t1y = (tiles_y + 3) / 4;
  float* l2 = l1 + t1x * t1y;
  const int t2x = (t1x + 3) / 4, t2y = (t1y + 3) / 4;
  float* all = l2 + t2x * t2y;
  for (int i = threadIdx.x; i < t1x * t1y; i += 256) {
    const int x = i % t1x, y = i / t1x;
    float m = 0.0F;
    for (int yy = y * 4; yy < min(y * 4 + 4, tiles_y); ++yy)
      for (int xx = x * 4; xx < min(x * 4 + 4, tiles_x); ++xx) m = fmaxf(m, tile_max[yy * tiles_x + xx]);
    l1[i] = m;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < t2x * t2y; i += 256) {
    const int x = i % t2x, y = i / t2x;
    float m = 0.0F;
    for (int yy = y * 4; yy < min(y * 4 + 4, t1y); ++yy)
      for (int xx = x * 4; xx < min(x * 4 + 4, t1x); ++xx) m = fmaxf(m, l1[yy * t1x + xx]);
    l2[i] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = 0.0F;
    for (int i = 0; i < t2x * t2y; ++i) m = fmaxf(m, l2[i]);
    *all = m;
  }
}

// Maximum usable depth over the 16 px tiles [tx0, tx1] x [ty0, ty1], read at
// the coarsest level that still needs at most 4 x 4 lookups (a coarser tile
// covers more pixels, so the bound stays conservative).
__device__ inline float region_max(const float* __restrict__ tile_max, int tiles_x, int tiles_y, int tx0, int tx1,
                                   int ty0, int ty1) {
  const int t1x = (tiles_x + 3) / 4, t1y = (tiles_y + 3) / 4;
  const int t2x = (t1x + 3) / 4, t2y = (t1y + 3) / 4;
  const float* l1 = tile_max + tiles_x * tiles_y;
  const float* l2 = l1 + t1x * t1y;
  const float* all = l2 + t2x * t2y;
  const float* lv = tile_max;
  int w = tiles_x;
  for (int level = 0; level < 3; ++level) {
    if ((tx1 - tx0 + 1) * (ty1 - ty0 + 1) <= 16) {
      float m = 0.0F;
      for (int y = ty0; y <= ty1; ++y)
        for (int x = tx0; x <= tx1; ++x) m = fmaxf(m, lv[y * w + x]);
      return m;
    }
    tx0 >>= 2;
    tx1 >>= 2;
    ty0 >>= 2;
    ty1 >>= 2;
    lv = level == 0 ? l1 : l2;
    w = level == 0 ? t1x : t2x;
  }
  (void)t2y;
  return *all;
}

// 2. conservative brick cull.
__global__ void __launch_bounds__(256) brick_cull_kernel(Grid g, FrameParams f,
                                                          const float* __restrict__ tile_max,
                                                          int tiles_x, int tiles_y,
                                                          uint32_t* __restrict__ list,
                                                          uint32_t* __restrict__ count) {
  const int bnx = (g.nx + kBrickX - 1) / kBrickX, bny = (g.ny + kBrickY - 1) / kBrickY;
  const int bnz = (g.nz + kBrickZ - 1) / kBrickZ;
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  Rigid w2c;
  if (!block_w2c(f, w2c)) return;
  bool keep = false;
  if (b < static_cast<uint32_t>(bnx * bny * bnz)) {
    const int bx = b % bnx, by = (b / bnx) % bny, bz = b / (bnx * bny);
    const int xs[2] = {bx * kBrickX, min(bx * kBrickX + kBrickX - 1, g.nx - 1)};
    const int ys[2] = {by * kBrickY, min(by * kBrickY + kBrickY - 1, g.ny - 1)};
    const int zs[2] = {bz * kBrickZ, min(bz * kBrickZ + kBrickZ - 1, g.nz - 1)};
    float zmin = 3.0e38F, zmax = -3.0e38F, umin = 3.0e38F, umax = -3.0e38F, vmin = 3.0e38F,
          vmax = -3.0e38F;
    V3 cam[8];
    for (int c = 0; c < 8; ++c) {
      cam[c] = xform(w2c, g.center(xs[c & 1], ys[(c >> 1) & 1], zs[c >> 2]));
      zmin = fminf(zmin, cam[c].z);
      zmax = fmaxf(zmax, cam[c].z);
    }
    keep = zmax > -1.0e-4F;
    int tx0 = 0, tx1 = tiles_x - 1, ty0 = 0, ty1 = tiles_y - 1;
    if (keep && zmin > 1.0e-3F) {
      for (int c = 0; c < 8; ++c) {
        float u, v;
        project(f.k, cam[c], u, v);
        umin = fminf(umin, u);
        umax = fmaxf(umax, u);
        vmin = fminf(vmin, v);
        vmax = fmaxf(vmax, v);
      }
      // a voxel lands on pixel round(u): usable columns are [0, w); keep a
      // two-pixel margin against rounding of the hull vertices.
      umin -= 2.0F;
      vmin -= 2.0F;
      umax += 2.0F;
      vmax += 2.0F;
      if (umax < -0.5F || vmax < -0.5F || umin > static_cast<float>(f.w) - 0.5F ||
          vmin > static_cast<float>(f.h) - 0.5F) {
        keep = false;
      } else {
        tx0 = max(0, static_cast<int>(floorf((fmaxf(umin, -1.0F) + 0.5F) / kTile)));
        ty0 = max(0, static_cast<int>(floorf((fmaxf(vmin, -1.0F) + 0.5F) / kTile)));
        tx1 = min(tiles_x - 1, static_cast<int>(floorf((fminf(umax, 1.0e6F) + 0.5F) / kTile)));
        ty1 = min(tiles_y - 1, static_cast<int>(floorf((fminf(vmax, 1.0e6F) + 0.5F) / kTile)));
      }
    }
    if (keep && (tx0 > tx1 || ty0 > ty1)) keep = false;
    if (keep) {
      const float dmax = region_max(tile_max, tiles_x, tiles_y, tx0, tx1, ty0, ty1);
      if (!(dmax > 0.0F)) {
        keep = false;
      } else if (zmin > 1.0e-3F) {
        // |p| / lambda >= z (1 - 0.71 / fx) for a voxel rounded to its pixel;
        // fusion needs |p| / lambda <= d + truncation.
        const float trunc = f.dist_scaled ? g.trunc + f.trunc_scale * dmax : g.trunc;
        const float shrink = 1.0F - 1.0F / fmaxf(f.k.fx, 1.0F);
        if (zmin * shrink > dmax + trunc + 1.0e-3F) keep = false;
      }
    }
  }
  // wave-aggregated append
  const unsigned long long mask = __ballot(keep);
  const int lane = threadIdx.x & 63;
  uint32_t base = 0;
  if (lane == 0 && mask != 0ULL) base = atomicAdd(count, static_cast<uint32_t>(__popcll(mask)));
  base = __shfl(base, 0);
  if (keep) {
    const unsigned long long below = mask & ((1ULL << lane) - 1ULL);
    list[base + __popcll(below)] = pack_brick(static_cast<int>(b % bnx), static_cast<int>((b / bnx) % bny),
                                              static_cast<int>(b / (bnx * bny)));
  }
}

// 3. integrate the listed bricks.
template <typename Raw>
__device__ inline bool same_raw(Raw a, Raw b) {
  if constexpr (sizeof(Raw) == 8) {
    return __builtin_bit_cast(unsigned long long, a) == __builtin_bit_cast(unsigned long long, b);
  } else {
    return a == b;
  }
}

// observe + fuse (tsdf_integration.hpp:184-252) of N voxels of one thread,
// in phases so that the N pixel-record loads, then the N voxel loads, are in
// flight together: (1) project, (2) read the pixel records, (3) decide,
// (4) read the voxels, (5) update.  any: some voxel fused; ones_ok: every
// fused voxel reads tsdf exactly 1 afterwards; complete: every voxel was
// fused, or sits at an in-image pixel without depth and already holds an
// observed 1.  A free-space part (every voxel observed and reading exactly
// 1) is known when ones_ok holds for all its voxels and either all are
// complete now or the part was free before (an unfused voxel keeps its
// value).
//
// Most voxels are decided without IEEE divisions: the projection and the
// signed distance are first computed with the hardware reciprocal and square
// root (relative error below 2^-21), and the exact IEEE computation runs only
// when an approximate value lies within a safety margin of a decision
// boundary (a half-integer pixel coordinate, -truncation, +truncation).  Away
// from the boundaries the approximate value decides the same way as the exact
// one, so the outcome is bit-identical:
//  * the pixel: lround(u) = floor(u' + 0.5) when u' is farther than
//    1e-3 + 2e-6 |q'| from a half-integer (|u'| < 4096, q' = fx x / z);
//  * sd >= truncation (free space): tsdf = clamp(sd / trunc) = 1 exactly and
//    no colour; the update ((t0 w0) + (1 w)) / (w0 + w) is 1 exactly when
//    t0 = 1 (numerator and denominator are the same rounded sum), so the
//    division runs only for t0 != 1;
//  * sd < -truncation: rejected.
// A voxel whose stored bits do not change (saturated free space) is not
// written back.
template <int Kind, bool Color, bool Count, int N, bool Column = false>
__device__ inline void fuse_batch(const Grid& g, const FrameParams& f, const Rigid& w2c,
                                  const float4* __restrict__ pix, const uint32_t* __restrict__ color_img,
                                  typename VoxelIO<Kind>::Raw* __restrict__ vox, float4* __restrict__ colors,
                                  const int (&xs)[N], const int (&ys)[N], const int (&zs)[N],
                                  const size_t (&vis)[N], const bool (&inside)[N], unsigned int& my_fused,
                                  unsigned int& my_color, bool& any, bool& ones_ok, bool& complete) {
  using IO = VoxelIO<Kind>;
  int pi[N];
  V3 cp[N];
  // A column (every voxel at the same x, y): the voxel centre's x and y and
  // each row's first two products and their sum are the same for all of
  // them, so they are formed once; the rest of xform(w2c, centre) is the
  // same float sequence: ((r.x cx + r.y cy) + r.z cz) + t.
  float a0 = 0.0F, a1 = 0.0F, a2 = 0.0F;
  if constexpr (Column) {
    const float cx = g.ox + ((static_cast<float>(xs[0]) + 0.5F) * g.voxel);
    const float cy = g.oy + ((static_cast<float>(ys[0]) + 0.5F) * g.voxel);
    a0 = (w2c.r0.x * cx) + (w2c.r0.y * cy);
    a1 = (w2c.r1.x * cx) + (w2c.r1.y * cy);
    a2 = (w2c.r2.x * cx) + (w2c.r2.y * cy);
  }
#pragma unroll
  for (int k = 0; k < N; ++k) {
    pi[k] = -1;
    cp[k] = mk3(0.0F, 0.0F, 0.0F);
    if (!inside[k]) continue;
    if constexpr (Column) {
      const float cz = g.oz + ((static_cast<float>(zs[k]) + 0.5F) * g.voxel);
      cp[k] = mk3((a0 + (w2c.r0.z * cz)) + w2c.t.x, (a1 + (w2c.r1.z * cz)) + w2c.t.y, (a2 + (w2c.r2.z * cz)) + w2c.t.z);
    } else {
      cp[k] = xform(w2c, g.center(xs[k], ys[k], zs[k]));
    }
    if (cp[k].z <= 0.0F) continue;
    const float rz = __builtin_amdgcn_rcpf(cp[k].z);
    const float qx = (f.k.fx * cp[k].x) * rz, qy = (f.k.fy * cp[k].y) * rz;
    const float ua = qx + f.k.cx, va = qy + f.k.cy;
    bool exact = !(fabsf(ua) < 4096.0F && fabsf(va) < 4096.0F);
    if (!exact) {
      const float eu = ua - floorf(ua), ev = va - floorf(va);
      exact = fabsf(eu - 0.5F) <= 1.0e-3F + 2.0e-6F * fabsf(qx) || fabsf(ev - 0.5F) <= 1.0e-3F + 2.0e-6F * fabsf(qy);
    }
    int col, row;
    if (exact) {
      float u, v;
      project(f.k, cp[k], u, v);
      if (!round_index(u, f.w, col) || !round_index(v, f.h, row)) continue;
    } else {
      col = static_cast<int>(floorf(ua + 0.5F));
      row = static_cast<int>(floorf(va + 0.5F));
      if (col < 0 || col >= f.w || row < 0 || row >= f.h) continue;
    }
    pi[k] = row * f.w + col;
  }
  float4 p[N];
#pragma unroll
  for (int k = 0; k < N; ++k) p[k] = pi[k] >= 0 ? pix[pi[k]] : make_float4(-1.0F, 1.0F, 0.0F, 1.0F);
  // mode: 0 no fuse, 1 free space (tsdf 1, no colour), 2 in the band (exact
  // sd), 3 no fuse at an in-image pixel without depth: the voxel is read
  // only to learn whether it already holds an observed 1 (a depth dropout
  // inside free space then does not break the part's completeness).
  // The decisions are selects; only voxels whose approximate distance lies
  // within a margin of the band take the exact (IEEE) path.
  int mode[N];
  float sd[N];
  bool exact[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const bool has_depth = !(p[k].x < 0.0F || p[k].z <= 0.0F);  // no depth, or w <= 0 drops the fuse
    int m = (!has_depth && pi[k] >= 0) ? 3 : 0;
    float s = 0.0F;
    bool ex = false;
    if (f.projective) {
      const float sa = __builtin_amdgcn_sqrtf(dot3(cp[k], cp[k])) * __builtin_amdgcn_rcpf(p[k].y);
      const float sda = p[k].x - sa;
      const float margin = 1.0e-6F + 4.0e-6F * sa;
      const bool fr = sda >= p[k].w + margin;
      const bool rj = sda < -p[k].w - margin;
      m = has_depth ? (fr ? 1 : (rj ? 0 : 2)) : m;
      ex = has_depth && !fr && !rj;
    } else {
      s = p[k].x - cp[k].z;
      m = has_depth ? (s >= p[k].w ? 1 : (s < -p[k].w ? 0 : 2)) : m;
    }
    mode[k] = m;
    sd[k] = s;
    exact[k] = ex;
  }
#pragma unroll
  for (int k = 0; k < N; ++k) {
    if (!exact[k]) continue;
    sd[k] = p[k].x - (norm3(cp[k]) / p[k].y);
    if (sd[k] < -p[k].w) mode[k] = 0;
  }
  typename IO::Raw raw[N];
#pragma unroll
  for (int k = 0; k < N; ++k) raw[k] = mode[k] != 0 ? vox[vis[k]] : IO::fresh();
  ones_ok = true;
  complete = true;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    if (mode[k] == 0) {
      complete = false;
      continue;
    }
    const float w = p[k].z;
    float t0, w0;
    bool obs, fin;
    IO::unpack(raw[k], t0, w0, obs, fin);
    if (mode[k] == 3) {
      complete = complete && obs && t0 == 1.0F;
      continue;
    }
    float avg;
    if (mode[k] == 1) {
      avg = t0 == 1.0F ? 1.0F : ((t0 * w0) + (1.0F * w)) / (w0 + w);
    } else {
      const float tsdf = fclamp_ref(sd[k] / p[k].w, -1.0F, 1.0F);
      avg = ((t0 * w0) + (tsdf * w)) / (w0 + w);
    }
    const float tw = fmin_ref(w0 + w, f.max_weight);
    const typename IO::Raw nv = IO::pack(avg, tw);
    if (!same_raw(nv, raw[k])) vox[vis[k]] = nv;
    IO::unpack(nv, t0, w0, obs, fin);
    ones_ok = ones_ok && t0 == 1.0F;
    any = true;
    if (Count) ++my_fused;
    if (Color && mode[k] == 2) {
      if (color_img != nullptr && sd[k] <= p[k].w * 0.5F) {
        const uint32_t px = color_img[pi[k]];
        const V3 c = mk3(static_cast<float>(px & 0xFFU), static_cast<float>((px >> 8U) & 0xFFU),
                         static_cast<float>((px >> 16U) & 0xFFU));
        const float4 old = colors[vis[k]];
        const float cw = old.w;
        const float den = cw + w;
        colors[vis[k]] = make_float4(((old.x * cw) + (c.x * w)) / den, ((old.y * cw) + (c.y * w)) / den,
                                     ((old.z * cw) + (c.z * w)) / den, fmin_ref(cw + w, f.max_weight));
        if (Count) ++my_color;
      }
    }
  }
}

// The voxel updates of one z-part (4 voxels from z0) of an (x, y) column.
template <int Kind, bool Color, bool Count>
__device__ inline bool integrate_column(const Grid& g, const FrameParams& f, const Rigid& w2c,
                                        const float4* __restrict__ pix, const uint32_t* __restrict__ color_img,
                                        typename VoxelIO<Kind>::Raw* __restrict__ vox,
                                        float4* __restrict__ colors, int x, int y, int z0,
                                        unsigned int& my_fused, unsigned int& my_color, bool& ones_ok,
                                        bool& complete) {
  constexpr int kHalf = kBrickZ / 2;  // one z-part of a block
  int xs[kHalf], ys[kHalf], zs[kHalf];
  size_t vis[kHalf];
  bool inside[kHalf];
  const size_t v0 = g.flat(x, y, z0), slice = static_cast<size_t>(g.nx) * static_cast<size_t>(g.ny);
#pragma unroll
  for (int k = 0; k < kHalf; ++k) {
    xs[k] = x;
    ys[k] = y;
    zs[k] = z0 + k;
    inside[k] = zs[k] < g.nz;
    vis[k] = inside[k] ? v0 + k * slice : 0;
  }
  bool any = false;
  fuse_batch<Kind, Color, Count, kHalf, true>(g, f, w2c, pix, color_img, vox, colors, xs, ys, zs, vis, inside,
                                              my_fused, my_color, any, ones_ok, complete);
  return any;
}

// Part bits of block fb: bit `part` (y-half * 2 + z-half) of its nibble.
__device__ inline void set_part(uint32_t* part_bits, int fb, int part, bool on) {
  const uint32_t bit = 1U << ((fb & 7) * 4 + part);
  if (on) atomicOr(&part_bits[fb >> 3], bit);
  else atomicAnd(&part_bits[fb >> 3], ~bit);
}

template <int Kind, bool Color, bool Count>
__global__ void __launch_bounds__(256) integrate_kernel(Grid g, FrameParams f,
                                                         const float4* __restrict__ pix,
                                                         const uint32_t* __restrict__ color_img,
                                                         typename VoxelIO<Kind>::Raw* __restrict__ vox,
                                                         float4* __restrict__ colors,
                                                         const uint32_t* __restrict__ list,
                                                         const uint32_t* __restrict__ count,
                                                         unsigned long long* __restrict__ fused,
                                                         uint32_t* __restrict__ obs_bits, BlockMap bm,
                                                         uint32_t* __restrict__ part_bits) {
  // A brick is 8 blocks along x of one y-half; a thread's column of 8
  // voxels is both z-halves of its block.  parts_*[h] bit j: block j's part
  // (y-half, z-half h) had every voxel ok / complete in every row.
  __shared__ uint32_t parts_ok[2], parts_complete[2];
  Rigid w2c;
  if (!block_w2c(f, w2c)) return;
  const uint32_t n = *count;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  unsigned int my_fused = 0, my_color = 0;
  for (uint32_t li = blockIdx.x; li < n; li += gridDim.x) {
    const uint32_t b = list[li];
    const int bx = static_cast<int>(b & 0xFFU), by = static_cast<int>((b >> 8U) & 0xFFFU),
              bz = static_cast<int>(b >> 20U);
    const int x = bx * kBrickX + tx, y = by * kBrickY + ty, z0 = bz * kBrickZ;
    if (threadIdx.x < 2) {
      parts_ok[threadIdx.x] = 0xFFU;
      parts_complete[threadIdx.x] = 0xFFU;
    }
    __syncthreads();
    bool any0 = false, any1 = false, ok0 = false, ok1 = false, cpl0 = false, cpl1 = false;
    if (x < g.nx && y < g.ny) {
      any0 = integrate_column<Kind, Color, Count>(g, f, w2c, pix, color_img, vox, colors, x, y, z0, my_fused,
                                                  my_color, ok0, cpl0);
      any1 = integrate_column<Kind, Color, Count>(g, f, w2c, pix, color_img, vox, colors, x, y, z0 + kBrickZ / 2,
                                                  my_fused, my_color, ok1, cpl1);
    }
    // Lanes 8j..8j+7 of a wave are block j's row: one ballot per flag; lane
    // j < 8 folds its block's byte of each.
    const unsigned long long om0 = __ballot(ok0), cm0 = __ballot(cpl0), om1 = __ballot(ok1),
                             cm1 = __ballot(cpl1), am = __ballot(any0 || any1);
    if (tx < 8) {
      const int j = tx;
      auto full = [&](unsigned long long m) { return ((m >> (8 * j)) & 0xFFULL) == 0xFFULL; };
      const uint32_t bit = 1U << j;
      if (!full(om0)) atomicAnd(&parts_ok[0], ~bit);
      if (!full(cm0)) atomicAnd(&parts_complete[0], ~bit);
      if (!full(om1)) atomicAnd(&parts_ok[1], ~bit);
      if (!full(cm1)) atomicAnd(&parts_complete[1], ~bit);
      const int blk_x = bx * 8 + j;
      if (((am >> (8 * j)) & 0xFFULL) != 0ULL && blk_x < bm.bx) {
        const int fb = bm.flat(blk_x, (by * kBrickY) >> 3, z0 >> 3);
        atomicOr(&obs_bits[fb >> 5], 1U << (fb & 31));
      }
    }
    __syncthreads();
    if (threadIdx.x < 16) {
      // part (y-half, z-half h) of block j
      const int j = threadIdx.x & 7, h = threadIdx.x >> 3;
      const int blk_x = bx * 8 + j;
      if (blk_x < bm.bx) {
        const int part = (((by * kBrickY) >> 2) & 1) * 2 + h;
        const bool ok = ((parts_ok[h] >> j) & 1U) != 0U, cpl = ((parts_complete[h] >> j) & 1U) != 0U;
        if (!ok || cpl) set_part(part_bits, bm.flat(blk_x, (by * kBrickY) >> 3, z0 >> 3), part, ok);
      }
    }
    __syncthreads();
  }
  if (Count) {
    for (int off = 32; off > 0; off >>= 1) {
      my_fused += __shfl_xor(my_fused, off);
      my_color += __shfl_xor(my_color, off);
    }
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&fused[0], static_cast<unsigned long long>(my_fused));
      atomicAdd(&fused[1], static_cast<unsigned long long>(my_color));
    }
  }
}



// running totals f[2..3] += per-call counts f[0..1]
__global__ void accumulate_counts_kernel(unsigned long long* f) {
  if (threadIdx.x < 2) f[2 + threadIdx.x] += f[threadIdx.x];
}

// skip bit of a block = the block and its 26 neighbours hold no observed
// voxel.  One thread per block; a wave's 64 consecutive blocks form two words.
__device__ inline bool block_free(const uint32_t* __restrict__ parts, int fb) {
  return ((parts[fb >> 3] >> ((fb & 7) * 4)) & 0xFU) == 0xFU;
}

// free bit of a block = the block and its 26 neighbours (all inside the
// grid) hold only fused voxels reading tsdf exactly 1.
__global__ void __launch_bounds__(256) skip_map_kernel(const uint32_t* __restrict__ obs, uint32_t* __restrict__ skip,
                                                        const uint32_t* __restrict__ parts,
                                                        uint32_t* __restrict__ free_bits, BlockMap bm, int nblocks,
                                                        IcpState* end_state, double* trajectory) {
  // end_state: this launch also closes the frame (one thread): the
  // trajectory slot is the device frame counter, so the same launch serves
  // every frame of a captured graph
  if (end_state != nullptr && blockIdx.x == 0 && threadIdx.x == 0) {
    const long long slot = end_state->frames_done;
    for (int i = 0; i < 16; ++i) {
      trajectory[slot * 16 + i] = end_state->pose[i];
      end_state->prev2_pose[i] = end_state->prev_pose[i];
      end_state->prev_pose[i] = end_state->pose[i];
    }
    end_state->frames_done += 1;
  }
  const int b = blockIdx.x * 256 + threadIdx.x;
  bool empty = false, free_nb = false;
  if (b < nblocks) {
    const int x = b % bm.bx, y = (b / bm.bx) % bm.by, z = b / (bm.bx * bm.by);
    free_nb = x > 0 && y > 0 && z > 0 && x + 1 < bm.bx && y + 1 < bm.by && z + 1 < bm.bz;
    for (int dz = -1; dz <= 1 && free_nb; ++dz)
      for (int dy = -1; dy <= 1 && free_nb; ++dy)
        for (int dx = -1; dx <= 1; ++dx)
          if (!block_free(parts, bm.flat(x + dx, y + dy, z + dz))) {
            free_nb = false;
            break;
          }
  }
  {
    const unsigned long long m = __ballot(free_nb);
    const int lane = threadIdx.x & 63;
    const int word0 = (blockIdx.x * 256 + (threadIdx.x & ~63)) >> 5;
    const int nwords = (nblocks + 31) >> 5;
    if (lane == 0 && word0 < nwords) free_bits[word0] = static_cast<uint32_t>(m);
    if (lane == 32 && word0 + 1 < nwords) free_bits[word0 + 1] = static_cast<uint32_t>(m >> 32);
  }
  if (b < nblocks) {
    const int x = b % bm.bx, y = (b / bm.bx) % bm.by, z = b / (bm.bx * bm.by);
    empty = true;
    for (int dz = -1; dz <= 1 && empty; ++dz)
      for (int dy = -1; dy <= 1 && empty; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int xx = x + dx, yy = y + dy, zz = z + dz;
          if (xx < 0 || yy < 0 || zz < 0 || xx >= bm.bx || yy >= bm.by || zz >= bm.bz) continue;
          if (bm.bit(obs, bm.flat(xx, yy, zz))) {
            empty = false;
            break;
          }
        }
  }
  const unsigned long long m = __ballot(empty);
  const int lane = threadIdx.x & 63;
  const int word0 = (blockIdx.x * 256 + (threadIdx.x & ~63)) >> 5;
  const int nwords = (nblocks + 31) >> 5;
  if (lane == 0 && word0 < nwords) skip[word0] = static_cast<uint32_t>(m);
  if (lane == 32 && word0 + 1 < nwords) skip[word0 + 1] = static_cast<uint32_t>(m >> 32);
}

// observed bits from the voxel array (after an upload): one thread per block.
template <int Kind>
__global__ void __launch_bounds__(256) rebuild_observed_kernel(const typename VoxelIO<Kind>::Raw* __restrict__ v,
                                                                Grid g, uint32_t* __restrict__ obs, BlockMap bm,
                                                                int nblocks) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  bool any = false;
  if (b < nblocks) {
    const int bx = b % bm.bx, by = (b / bm.bx) % bm.by, bz = b / (bm.bx * bm.by);
    for (int z = bz * 8; z < min(bz * 8 + 8, g.nz) && !any; ++z)
      for (int y = by * 8; y < min(by * 8 + 8, g.ny) && !any; ++y)
        for (int x = bx * 8; x < min(bx * 8 + 8, g.nx); ++x) {
          float t, w;
          bool o, fin;
          VoxelIO<Kind>::unpack(v[g.flat(x, y, z)], t, w, o, fin);
          if (o) {
            any = true;
            break;
          }
        }
  }
  const unsigned long long m = __ballot(any);
  const int lane = threadIdx.x & 63;
  const int word0 = (blockIdx.x * 256 + (threadIdx.x & ~63)) >> 5;
  const int nwords = (nblocks + 31) >> 5;
  if (lane == 0 && word0 < nwords) obs[word0] = static_cast<uint32_t>(m);
  if (lane == 32 && word0 + 1 < nwords) obs[word0 + 1] = static_cast<uint32_t>(m >> 32);
}

template <int Kind>
__global__ void fill_voxels_kernel(typename VoxelIO<Kind>::Raw* v, float4* colors, size_t n) {
  const typename VoxelIO<Kind>::Raw fresh = VoxelIO<Kind>::fresh();
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    v[i] = fresh;
    if (colors != nullptr) colors[i] = make_float4(0.0F, 0.0F, 0.0F, 0.0F);
  }
}

// ---- hashed volume: TruncationBandWalk + BlockRep::integrate
// (tsdf_integration.hpp:378-428, block_rep.hpp:244-307) ----------------------
__global__ void sparse_begin_kernel(uint32_t* sp) {
  if (threadIdx.x == 0) {
    sp[3] += 1U;  // frame id
    sp[1] = 0U;   // work count
  }
}

// RigidTransform inverse {R^T, -R^T t} in float (vector.hpp:214-224).
__device__ inline Rigid rigid_inverse_f(const Rigid& m) {
  Rigid r;
  r.r0 = mk3(m.r0.x, m.r1.x, m.r2.x);
  r.r1 = mk3(m.r0.y, m.r1.y, m.r2.y);
  r.r2 = mk3(m.r0.z, m.r1.z, m.r2.z);
  const V3 t = rot(r, m.t);
  r.t = mk3(-t.x, -t.y, -t.z);
  return r;
}

// Finds or inserts `key`; its bucket, or kNoSlot when the table or the pool
// is exhausted (the request is counted as overflow).  Entries only go from
// empty to (key, pending) here, so a stale read can only see "empty", which
// the compare-and-swap corrects.
__device__ inline uint32_t hash_insert(const HashView& hv, uint32_t key, uint32_t* sp, uint32_t entry_limit,
                                       uint32_t capacity) {
  uint32_t h = hv.bucket(key);
  for (uint32_t probe = 0; probe <= hv.mask; ++probe) {
    unsigned long long e = hv.table[h];
    if (e == kHashEmpty) {
      if (sp[2] >= entry_limit || sp[0] >= capacity) break;
      const unsigned long long mine = (static_cast<unsigned long long>(key) << 32) | kPendingSlot;
      const unsigned long long old = atomicCAS(&hv.table[h], kHashEmpty, mine);
      if (old == kHashEmpty) {
        atomicAdd(&sp[2], 1U);
        return h;
      }
      e = old;
    }
    if (static_cast<uint32_t>(e >> 32) == key) return h;
    h = (h + 1) & hv.mask;
  }
  atomicAdd(&sp[4], 1U);
  return kNoSlot;
}

// Marks the blocks of every pixel's truncation band (visit_pixel): samples
// from d - band to d + band every `step`, skipping a block equal to the
// pixel's previous visit.
__global__ void __launch_bounds__(256) band_walk_kernel(Grid g, FrameParams f, const uint16_t* __restrict__ depth,
                                                         HashView hv, uint32_t* __restrict__ stamp,
                                                         uint32_t* __restrict__ sp, uint32_t entry_limit,
                                                         uint32_t capacity, float step) {
  Rigid w2c;
  if (!block_w2c(f, w2c)) return;
  const Rigid c2w = rigid_inverse_f(w2c);
  const int col = blockIdx.x * kTile + threadIdx.x, row = blockIdx.y * kTile + threadIdx.y;
  if (col >= f.w || row >= f.h) return;
  float measured;
  if (!depth_in_range(depth[static_cast<size_t>(row) * f.w + col], f.scale, f.lo, f.hi, measured)) return;
  const float band = f.dist_scaled ? g.trunc + (f.trunc_scale * measured) : g.trunc;
  const int samples = static_cast<int>((2.0F * band) / step) + 1;
  const float inv_voxel = 1.0F / g.voxel;
  const uint32_t frame = sp[3];
  uint32_t previous = kNoSlot;
  for (int s = 0; s <= samples; ++s) {
    const float d = (measured - band) + (static_cast<float>(s) * step);
    const V3 world = xform(c2w, back_project(f.k, static_cast<float>(col), static_cast<float>(row), d));
    const V3 gc = scale3(sub3(world, mk3(g.ox, g.oy, g.oz)), inv_voxel);
    if (!(gc.x >= 0.0F && gc.y >= 0.0F && gc.z >= 0.0F && gc.x < static_cast<float>(g.nx) &&
          gc.y < static_cast<float>(g.ny) && gc.z < static_cast<float>(g.nz)))
      continue;
    const uint32_t key = hv.key_of(static_cast<int>(gc.x), static_cast<int>(gc.y), static_cast<int>(gc.z));
    if (key == previous) continue;
    previous = key;
    const uint32_t b = hash_insert(hv, key, sp, entry_limit, capacity);
    if (b != kNoSlot) stamp[b] = frame;
  }
}

// Pool slots for this frame's new blocks and the frame's work list (each
// visited block once).  Appends are aggregated per wave.
__global__ void __launch_bounds__(256) assign_slots_kernel(HashView hv, const uint32_t* __restrict__ stamp,
                                                            uint32_t* __restrict__ sp,
                                                            uint32_t* __restrict__ block_of_slot,
                                                            uint32_t* __restrict__ work, uint32_t capacity,
                                                            uint32_t buckets) {
  const uint32_t frame = sp[3];
  const int lane = threadIdx.x & 63;
  const uint32_t stride = gridDim.x * 256U;
  const uint32_t rounds = (buckets + stride - 1) / stride;
  for (uint32_t r = 0; r < rounds; ++r) {
    const uint32_t b = r * stride + blockIdx.x * 256U + threadIdx.x;
    unsigned long long e = kHashEmpty;
    bool visited = false;
    if (b < buckets) {
      e = hv.table[b];
      visited = e != kHashEmpty && stamp[b] == frame;
    }
    uint32_t slot = visited ? static_cast<uint32_t>(e) : kNoSlot;
    const bool fresh = visited && slot == kPendingSlot;
    const unsigned long long fm = __ballot(fresh);
    uint32_t base = 0;
    if (lane == 0 && fm != 0ULL) base = atomicAdd(&sp[0], static_cast<uint32_t>(__popcll(fm)));
    base = __shfl(base, 0);
    if (fresh) {
      const uint32_t s = base + static_cast<uint32_t>(__popcll(fm & ((1ULL << lane) - 1ULL)));
      const uint32_t key = static_cast<uint32_t>(e >> 32);
      if (s < capacity) {
        slot = s;
        block_of_slot[s] = key;
      } else {
        slot = kNoSlot;
        atomicAdd(&sp[4], 1U);
      }
      hv.table[b] = (static_cast<unsigned long long>(key) << 32) | slot;
    }
    const bool listed = visited && slot != kNoSlot;
    const unsigned long long lm = __ballot(listed);
    uint32_t wbase = 0;
    if (lane == 0 && lm != 0ULL) wbase = atomicAdd(&sp[1], static_cast<uint32_t>(__popcll(lm)));
    wbase = __shfl(wbase, 0);
    if (listed) work[wbase + __popcll(lm & ((1ULL << lane) - 1ULL))] = slot;
  }
}

// Every voxel of the listed blocks (BlockVoxels, clipped to the grid): one
// workgroup per block, x fastest, so pool accesses are contiguous.
template <int Kind, bool Color, bool Count>
__global__ void __launch_bounds__(256) integrate_blocks_kernel(Grid g, FrameParams f, const float4* __restrict__ pix,
                                                                const uint32_t* __restrict__ color_img,
                                                                typename VoxelIO<Kind>::Raw* __restrict__ pool,
                                                                float4* __restrict__ cpool,
                                                                const uint32_t* __restrict__ block_of_slot,
                                                                const uint32_t* __restrict__ work,
                                                                const uint32_t* __restrict__ sp,
                                                                unsigned long long* __restrict__ fused,
                                                                uint32_t* __restrict__ obs_bits, BlockMap bm,
                                                                uint32_t* __restrict__ part_bits) {
  __shared__ int any_fused;
  __shared__ int all_ok, all_complete;
  Rigid w2c;
  if (!block_w2c(f, w2c)) return;
  const uint32_t n = sp[1];
  unsigned int my_fused = 0, my_color = 0;
  for (uint32_t li = blockIdx.x; li < n; li += gridDim.x) {
    const uint32_t slot = work[li];
    const uint32_t key = block_of_slot[slot];
    const int bx = static_cast<int>(key % static_cast<uint32_t>(bm.bx));
    const int by = static_cast<int>((key / static_cast<uint32_t>(bm.bx)) % static_cast<uint32_t>(bm.by));
    const int bz = static_cast<int>(key / static_cast<uint32_t>(bm.bx * bm.by));
    if (threadIdx.x == 0) {
      any_fused = 0;
      all_ok = 1;
      all_complete = 1;
    }
    __syncthreads();
    int xs[2], ys[2], zs[2];
    size_t vis[2];
    bool inside[2];
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
      const int i = static_cast<int>(threadIdx.x) + pass * 256;
      xs[pass] = bx * 8 + (i & 7);
      ys[pass] = by * 8 + ((i >> 3) & 7);
      zs[pass] = bz * 8 + (i >> 6);
      inside[pass] = xs[pass] < g.nx && ys[pass] < g.ny && zs[pass] < g.nz;
      vis[pass] = static_cast<size_t>(slot) * kBlockVoxels + i;
    }
    bool any = false, ones_ok = true, complete = true;
    fuse_batch<Kind, Color, Count, 2>(g, f, w2c, pix, color_img, pool, cpool, xs, ys, zs, vis, inside, my_fused,
                                      my_color, any, ones_ok, complete);
    if (any) any_fused = 1;
    if (!ones_ok) all_ok = 0;
    if (!complete) all_complete = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
      const int fb = bm.flat(bx, by, bz);
      if (any_fused != 0) atomicOr(&obs_bits[fb >> 5], 1U << (fb & 31));
      const uint32_t nib = 0xFU << ((fb & 7) * 4);
      if (all_ok == 0) atomicAnd(&part_bits[fb >> 3], ~nib);
      else if (all_complete != 0) atomicOr(&part_bits[fb >> 3], nib);  // else unchanged
    }
    __syncthreads();
  }
  if (Count) {
    for (int off = 32; off > 0; off >>= 1) {
      my_fused += __shfl_xor(my_fused, off);
      my_color += __shfl_xor(my_color, off);
    }
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&fused[0], static_cast<unsigned long long>(my_fused));
      atomicAdd(&fused[1], static_cast<unsigned long long>(my_color));
    }
  }
}

__global__ void fill_table_kernel(unsigned long long* table, uint32_t* stamp, size_t buckets, uint32_t* sp) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < buckets;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    table[i] = kHashEmpty;
    stamp[i] = 0U;
  }
  if (blockIdx.x == 0 && threadIdx.x < 8) sp[threadIdx.x] = 0U;
}

template <int Kind, bool Color>
void launch_blocks_t(kf_context* ctx, const FrameParams& f, const uint32_t* color_img, bool count) {
  using Raw = typename VoxelIO<Kind>::Raw;
  const int grid = 2048;
  if (count) {
    hipLaunchKernelGGL((integrate_blocks_kernel<Kind, Color, true>), dim3(grid), dim3(256), 0, ctx->stream,
                       ctx->grid, f, ctx->pix, color_img, static_cast<Raw*>(ctx->voxels), ctx->colors,
                       ctx->block_of_slot, ctx->work, ctx->sparse, ctx->fused, ctx->obs_bits, block_map(ctx->grid),
                       ctx->part_bits);
  } else {
    hipLaunchKernelGGL((integrate_blocks_kernel<Kind, Color, false>), dim3(grid), dim3(256), 0, ctx->stream,
                       ctx->grid, f, ctx->pix, color_img, static_cast<Raw*>(ctx->voxels), ctx->colors,
                       ctx->block_of_slot, ctx->work, ctx->sparse, ctx->fused, ctx->obs_bits, block_map(ctx->grid),
                       ctx->part_bits);
  }
}

// Dense expansion of a hashed volume (SparseVolumeSnapshot + host_dense_view,
// block_rep.hpp:170-182,254-272): unallocated voxels read as fresh.
template <int Kind>
__global__ void expand_hashed_kernel(Grid g, HashView hv, const typename VoxelIO<Kind>::Raw* __restrict__ pool,
                                     const float4* __restrict__ cpool, typename VoxelIO<Kind>::Raw* __restrict__ out,
                                     float4* __restrict__ cout_) {
  const size_t n = static_cast<size_t>(g.nx) * g.ny * g.nz;
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int x = static_cast<int>(i % g.nx), y = static_cast<int>((i / g.nx) % g.ny);
    const int z = static_cast<int>(i / (static_cast<size_t>(g.nx) * g.ny));
    const uint32_t slot = hv.find(hv.key_of(x, y, z));
    const size_t pi = static_cast<size_t>(slot) * kBlockVoxels + intra_of(x, y, z);
    out[i] = slot == kNoSlot ? VoxelIO<Kind>::fresh() : pool[pi];
    if (cout_ != nullptr) cout_[i] = slot == kNoSlot ? make_float4(0.0F, 0.0F, 0.0F, 0.0F) : cpool[pi];
  }
}

template <int Kind, bool Color>
void launch_integrate_t(kf_context* ctx, const FrameParams& f, const uint32_t* color_img, bool count) {
  using Raw = typename VoxelIO<Kind>::Raw;
  const int grid = 2048;
  if (count) {
    hipLaunchKernelGGL((integrate_kernel<Kind, Color, true>), dim3(grid), dim3(256), 0, ctx->stream,
                       ctx->grid, f, ctx->pix, color_img, static_cast<Raw*>(ctx->voxels),
                       ctx->colors, ctx->bricks, ctx->counters, ctx->fused, ctx->obs_bits,
                       block_map(ctx->grid), ctx->part_bits);
  } else {
    hipLaunchKernelGGL((integrate_kernel<Kind, Color, false>), dim3(grid), dim3(256), 0, ctx->stream,
                       ctx->grid, f, ctx->pix, color_img, static_cast<Raw*>(ctx->voxels),
                       ctx->colors, ctx->bricks, ctx->counters, ctx->fused, ctx->obs_bits,
                       block_map(ctx->grid), ctx->part_bits);
  }
}

}  // namespace

bool hashed(const kf_context* ctx) { return ctx->cfg.storage == KF_STORAGE_HASHED; }

HashView hash_view(const kf_context* ctx) {
  const BlockMap bm = block_map(ctx->grid);
  return HashView{ctx->hash_table, static_cast<uint32_t>(ctx->buckets - 1), bm.bx, bm.by, bm.bz};
}

void launch_expand_hashed(kf_context* ctx, void* voxels, float4* colors) {
  const HashView hv = hash_view(ctx);
  switch (ctx->cfg.voxel_store) {
    case KF_VOXEL_QUANTIZED:
      hipLaunchKernelGGL(expand_hashed_kernel<KF_VOXEL_QUANTIZED>, dim3(4096), dim3(256), 0, ctx->stream, ctx->grid,
                         hv, static_cast<const uint32_t*>(ctx->voxels), ctx->colors, static_cast<uint32_t*>(voxels),
                         colors);
      break;
    case KF_VOXEL_BF16:
      hipLaunchKernelGGL(expand_hashed_kernel<KF_VOXEL_BF16>, dim3(4096), dim3(256), 0, ctx->stream, ctx->grid, hv,
                         static_cast<const uint32_t*>(ctx->voxels), ctx->colors, static_cast<uint32_t*>(voxels),
                         colors);
      break;
    default:
      hipLaunchKernelGGL(expand_hashed_kernel<KF_VOXEL_FLOAT>, dim3(4096), dim3(256), 0, ctx->stream, ctx->grid, hv,
                         static_cast<const float2*>(ctx->voxels), ctx->colors, static_cast<float2*>(voxels), colors);
      break;
  }
  KF_HIP(hipGetLastError());
}

void launch_fill_pool(kf_context* ctx) {
  hipLaunchKernelGGL(fill_table_kernel, dim3(2048), dim3(256), 0, ctx->stream, ctx->hash_table, ctx->hash_stamp,
                     ctx->buckets, ctx->sparse);
  KF_HIP(hipGetLastError());
}

void launch_fill_voxels(kf_context* ctx) {
  const int grid = 4096;
  switch (ctx->cfg.voxel_store) {
    case KF_VOXEL_QUANTIZED:
      hipLaunchKernelGGL(fill_voxels_kernel<KF_VOXEL_QUANTIZED>, dim3(grid), dim3(256), 0,
                         ctx->stream, static_cast<uint32_t*>(ctx->voxels), ctx->colors, ctx->n_voxels);
      break;
    case KF_VOXEL_BF16:
      hipLaunchKernelGGL(fill_voxels_kernel<KF_VOXEL_BF16>, dim3(grid), dim3(256), 0, ctx->stream,
                         static_cast<uint32_t*>(ctx->voxels), ctx->colors, ctx->n_voxels);
      break;
    default:
      hipLaunchKernelGGL(fill_voxels_kernel<KF_VOXEL_FLOAT>, dim3(grid), dim3(256), 0, ctx->stream,
                         static_cast<float2*>(ctx->voxels), ctx->colors, ctx->n_voxels);
      break;
  }
  KF_HIP(hipGetLastError());
}

void launch_skip_map(kf_context* ctx, bool end_frame) {
  const BlockMap bm = block_map(ctx->grid);
  const int nblocks = bm.bx * bm.by * bm.bz;
  hipLaunchKernelGGL(skip_map_kernel, dim3((nblocks + 255) / 256), dim3(256), 0, ctx->stream, ctx->obs_bits,
                     ctx->skip_bits, ctx->part_bits, ctx->free_bits, bm, nblocks, end_frame ? ctx->icp : nullptr,
                     ctx->trajectory);
  KF_HIP(hipGetLastError());
}

void launch_rebuild_observed(kf_context* ctx) {
  const BlockMap bm = block_map(ctx->grid);
  const int nblocks = bm.bx * bm.by * bm.bz;
  const dim3 grid((nblocks + 255) / 256);
  switch (ctx->cfg.voxel_store) {
    case KF_VOXEL_QUANTIZED:
      hipLaunchKernelGGL(rebuild_observed_kernel<KF_VOXEL_QUANTIZED>, grid, dim3(256), 0, ctx->stream,
                         static_cast<const uint32_t*>(ctx->voxels), ctx->grid, ctx->obs_bits, bm, nblocks);
      break;
    case KF_VOXEL_BF16:
      hipLaunchKernelGGL(rebuild_observed_kernel<KF_VOXEL_BF16>, grid, dim3(256), 0, ctx->stream,
                         static_cast<const uint32_t*>(ctx->voxels), ctx->grid, ctx->obs_bits, bm, nblocks);
      break;
    default:
      hipLaunchKernelGGL(rebuild_observed_kernel<KF_VOXEL_FLOAT>, grid, dim3(256), 0, ctx->stream,
                         static_cast<const float2*>(ctx->voxels), ctx->grid, ctx->obs_bits, bm, nblocks);
      break;
  }
  KF_HIP(hipGetLastError());
  launch_skip_map(ctx);
}

void launch_integrate(kf_context* ctx, const kf_depth_frame& fr, bool count, bool from_state, bool end_frame,
                      bool accumulate) {
  const kf_config& c = ctx->cfg;
  FrameParams f{};
  f.w2c = rigid_from12(fr.world_to_camera);
  f.k = Intr{fr.fx, fr.fy, fr.cx, fr.cy};
  f.w = fr.width;
  f.h = fr.height;
  f.scale = c.depth_scale;
  f.lo = c.min_depth;
  f.hi = c.max_depth;
  f.obs_weight = c.observation_weight;
  f.max_weight = c.max_weight;
  f.projective = c.projective_distance;
  f.dist_scaled = c.distance_scaled_truncation;
  f.trunc_scale = c.truncation_distance_scale;
  f.angle_rule = c.tsdf_rule == KF_RULE_ANGLE_WEIGHTED;
  f.st = from_state ? ctx->icp : nullptr;
  const int tiles_x = (f.w + kTile - 1) / kTile, tiles_y = (f.h + kTile - 1) / kTile;
  // per-pixel and per-tile scratch is sized for the configured camera
  if (f.w > c.width || f.h > c.height)
    fail(KF_ERR_INVALID_ARGUMENT, "integrate: frame larger than the configured camera");
  hipLaunchKernelGGL(pixel_prep_kernel, dim3(tiles_x, tiles_y), dim3(kTile, kTile), 0, ctx->stream,
                     fr.depth, reinterpret_cast<const float4*>(fr.normals), ctx->pix, ctx->tile_max,
                     f, ctx->grid.trunc, ctx->counters, ctx->fused);
  const bool color = ctx->colors != nullptr;
  const uint32_t* cimg = fr.color;
  const bool timing = c.stage_timing != 0 && from_state;
  if (hashed(ctx)) {
    // band walk -> slots and work list -> listed blocks
    const float step = 0.5F * 8.0F * ctx->grid.voxel;  // TruncationBandWalk::half_block_step
    const uint32_t limit = static_cast<uint32_t>(ctx->buckets - ctx->buckets / 8);
    hipLaunchKernelGGL(sparse_begin_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->sparse);
    hipLaunchKernelGGL(band_walk_kernel, dim3(tiles_x, tiles_y), dim3(kTile, kTile), 0, ctx->stream, ctx->grid, f,
                       fr.depth, hash_view(ctx), ctx->hash_stamp, ctx->sparse, limit,
                       static_cast<uint32_t>(ctx->capacity), step);
    hipLaunchKernelGGL(assign_slots_kernel, dim3(1024), dim3(256), 0, ctx->stream, hash_view(ctx), ctx->hash_stamp,
                       ctx->sparse, ctx->block_of_slot, ctx->work, static_cast<uint32_t>(ctx->capacity),
                       static_cast<uint32_t>(ctx->buckets));
    if (timing) KF_HIP(hipEventRecord(ctx->ev[12], ctx->stream));
    switch (c.voxel_store) {
      case KF_VOXEL_QUANTIZED:
        color ? launch_blocks_t<KF_VOXEL_QUANTIZED, true>(ctx, f, cimg, count)
              : launch_blocks_t<KF_VOXEL_QUANTIZED, false>(ctx, f, cimg, count);
        break;
      case KF_VOXEL_BF16:
        color ? launch_blocks_t<KF_VOXEL_BF16, true>(ctx, f, cimg, count)
              : launch_blocks_t<KF_VOXEL_BF16, false>(ctx, f, cimg, count);
        break;
      default:
        color ? launch_blocks_t<KF_VOXEL_FLOAT, true>(ctx, f, cimg, count)
              : launch_blocks_t<KF_VOXEL_FLOAT, false>(ctx, f, cimg, count);
        break;
    }
    if (timing) KF_HIP(hipEventRecord(ctx->ev[13], ctx->stream));
    launch_skip_map(ctx, end_frame);
    if (count && accumulate)
      hipLaunchKernelGGL(accumulate_counts_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->fused);
    KF_HIP(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(tile_pyramid_kernel, dim3(1), dim3(256), 0, ctx->stream, ctx->tile_max, tiles_x, tiles_y);
  const Grid& g = ctx->grid;
  const int nb = ((g.nx + kBrickX - 1) / kBrickX) * ((g.ny + kBrickY - 1) / kBrickY) *
                 ((g.nz + kBrickZ - 1) / kBrickZ);
  hipLaunchKernelGGL(brick_cull_kernel, dim3((nb + 255) / 256), dim3(256), 0, ctx->stream, g, f,
                     ctx->tile_max, tiles_x, tiles_y, ctx->bricks, ctx->counters);
  if (timing) KF_HIP(hipEventRecord(ctx->ev[12], ctx->stream));
  switch (c.voxel_store) {
    case KF_VOXEL_QUANTIZED:
      color ? launch_integrate_t<KF_VOXEL_QUANTIZED, true>(ctx, f, cimg, count)
            : launch_integrate_t<KF_VOXEL_QUANTIZED, false>(ctx, f, cimg, count);
      break;
    case KF_VOXEL_BF16:
      color ? launch_integrate_t<KF_VOXEL_BF16, true>(ctx, f, cimg, count)
            : launch_integrate_t<KF_VOXEL_BF16, false>(ctx, f, cimg, count);
      break;
    default:
      color ? launch_integrate_t<KF_VOXEL_FLOAT, true>(ctx, f, cimg, count)
            : launch_integrate_t<KF_VOXEL_FLOAT, false>(ctx, f, cimg, count);
      break;
  }
  if (timing) KF_HIP(hipEventRecord(ctx->ev[13], ctx->stream));
  launch_skip_map(ctx, end_frame);
  if (count && accumulate)
    hipLaunchKernelGGL(accumulate_counts_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->fused);
  KF_HIP(hipGetLastError());
}

}  // namespace kf

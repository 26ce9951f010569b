// Dense TSDF integration (reference: IntegrationContext::observe/fuse and the
// Classic / AngleWeighted rules, include/kinectfusion/tsdf_integration.hpp:
// 183-342; full sweep :356-370; BasicVoxel::fused, volume.hpp:137-146).
//
// Three kernels per frame:
//  1. pixel_prep  -- everything observe() derives from the pixel alone:
//     measured depth, |K^-1 [u v 1]| (lambda), the rule's observation weight
//     (angle rule: cos(theta)/depth from the level-0 normal), the truncation
//     band; plus the maximum usable depth of every 16 x 16 pixel tile.
//  2. brick_cull  -- one thread per 16 x 4 x 8 voxel brick: a brick is kept
//     unless no voxel of it can project into the image in front of the
//     camera, or its nearest voxel lies farther than the deepest depth of
//     the tiles it covers plus the truncation band.  The bounds carry
//     margins so that no voxel the full sweep would fuse is ever culled: the
//     result is bit-identical to the full sweep.  Kept bricks are compacted
//     into a list in brick order, one atomic per workgroup.
//  3. integrate   -- persistent waves walk the list, one brick per wave: a
//     lane owns an (x, y) column of 8 voxels, and each z-slice of the brick
//     is 4 rows of 16 contiguous voxels.
#include <type_traits>

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
  // When st is set: w2c comes from the running device pose, and the pass is
  // skipped in a failed frame.  The pixel pass (the frame's first launch of
  // the pass) writes that world_to_camera (12 floats) and the skip flag (1.0
  // to skip) to frame_w2c once; the later launches read them.
  const IcpState* st;
  float* frame_w2c;
};

// world_to_camera of this frame (false: the frame failed, skip the pass).
__device__ inline bool block_w2c(const FrameParams& f, Rigid& out) {
  if (f.st == nullptr) {
    out = f.w2c;
    return true;
  }
  out = rigid_from12(f.frame_w2c);
  return f.frame_w2c[12] == 0.0F;
}

// 1. per-pixel record {depth m (or -1), lambda, weight, truncation}.
__global__ void __launch_bounds__(256) pixel_prep_kernel(const uint16_t* __restrict__ depth,
                                                          const float4* __restrict__ normals,
                                                          float4* __restrict__ pix,
                                                          float* __restrict__ tile_max,
                                                          FrameParams f, float base_trunc,
                                                          uint32_t* __restrict__ counters,
                                                          unsigned long long* __restrict__ fused,
                                                          IcpState* learned_st) {
  __shared__ float red[4];
  // the brick and fused counters start the pass at zero (read by later
  // launches only); learned_st: the learned warm start's end of the frame
  if (blockIdx.x == 0 && blockIdx.y == 0) {
    if (threadIdx.y == 0 && threadIdx.x < 4) counters[threadIdx.x] = 0U;
    if (threadIdx.y == 0 && threadIdx.x < 2) fused[threadIdx.x] = 0ULL;
    if (threadIdx.y == 0 && threadIdx.x == 0) {
      if (learned_st != nullptr) learned_end(learned_st);
      if (f.st != nullptr) {
        double inv[16];
        rigid_inverse64(f.st->pose, inv);
        for (int i = 0; i < 12; ++i) f.frame_w2c[i] = static_cast<float>(inv[i]);
        f.frame_w2c[12] = f.st->frame_failed != 0 ? 1.0F : 0.0F;
      }
    }
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
  // tile maximum: a butterfly in each wave, then the 4 wave maxima
  float m = dmax;
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if ((tid & 63) == 0) red[tid >> 6] = m;
  __syncthreads();
  if (tid == 0) tile_max[blockIdx.y * gridDim.x + blockIdx.x] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// Depth maxima the cull reads: the 16 px tiles (written by the pixel pass)
// and, built by each cull workgroup in LDS, 4 x 4 groups of tiles (64 px),
// 4 x 4 groups of those (256 px) and the frame maximum.  The group levels
// hold bit patterns (for floats >= 0 the unsigned order is the float order)
// so that LDS atomic maxima build them; l1 is null when they do not fit.
struct TilePyramid {
  const float* l0;
  const uint32_t* l1;
  const uint32_t* l2;
  int tx, ty, t1x, t2x;
  float frame_max;
};

constexpr int kPyramidLdsBytes = 32768;  // LDS for the group levels

// Maximum of the entries [x0, x1] x [y0, y1] (at most 4 x 4) of a level w wide.
template <typename T>
__device__ inline float gather4x4(const T* lv, int w, int x0, int x1, int y0, int y1) {
  float m = 0.0F;
#pragma unroll
  for (int dy = 0; dy < 4; ++dy) {
    const int y = min(y0 + dy, y1);
#pragma unroll
    for (int dx = 0; dx < 4; ++dx) {
      const T e = lv[y * w + min(x0 + dx, x1)];
      if constexpr (sizeof(T) == 4 && !std::is_same_v<T, float>) m = fmaxf(m, __uint_as_float(e));
      else m = fmaxf(m, e);
    }
  }
  return m;
}

// Maximum usable depth over the 16 px tiles [tx0, tx1] x [ty0, ty1], read at
// the finest level where the range spans at most 4 x 4 entries (a coarser
// entry covers more pixels, so the bound stays conservative).  The 16 reads
// are issued together, clamped into the range (a repeated entry does not
// change the maximum).
__device__ inline float region_max(const TilePyramid& p, int tx0, int tx1, int ty0, int ty1) {
  if (tx1 - tx0 < 4 && ty1 - ty0 < 4) return gather4x4(p.l0, p.tx, tx0, tx1, ty0, ty1);
  if (p.l1 == nullptr) return p.frame_max;
  tx0 >>= 2;
  tx1 >>= 2;
  ty0 >>= 2;
  ty1 >>= 2;
  if (tx1 - tx0 < 4 && ty1 - ty0 < 4) return gather4x4(p.l1, p.t1x, tx0, tx1, ty0, ty1);
  tx0 >>= 2;
  tx1 >>= 2;
  ty0 >>= 2;
  ty1 >>= 2;
  if (tx1 - tx0 < 4 && ty1 - ty0 < 4) return gather4x4(p.l2, p.t2x, tx0, tx1, ty0, ty1);
  return p.frame_max;
}

// 2. conservative brick cull: whether any voxel of brick (bx, by, bz) can be
// fused by this frame.  The brick's 8 corner centres bound it; in front of
// the camera, its image lies in the hull of their projections (a convex
// box projects to the hull of its corners), widened by a margin; every
// voxel behind the deepest depth of that image region by more than the
// truncation is rejected by the fusion test.
__device__ inline bool brick_may_fuse(const Grid& g, const FrameParams& f, const Rigid& w2c, const TilePyramid& tp,
                                      int bx, int by, int bz) {
  const int tiles_x = tp.tx, tiles_y = tp.ty;
  bool keep;
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
      // hardware reciprocal: its error (2^-21 relative) is far inside the margin below
      const float rz = __builtin_amdgcn_rcpf(cam[c].z);
      const float u = ((f.k.fx * cam[c].x) * rz) + f.k.cx, v = ((f.k.fy * cam[c].y) * rz) + f.k.cy;
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
    const float dmax = region_max(tp, tx0, tx1, ty0, ty1);
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
  return keep;
}

// One thread per brick for kCullBricks bricks in a row; the workgroup's kept
// bricks take one range of the list (one atomic per workgroup), in brick
// order.
constexpr int kCullBricks = 2;

__global__ void __launch_bounds__(256) brick_cull_kernel(Grid g, FrameParams f,
                                                          const float* __restrict__ tile_max,
                                                          int tiles_x, int tiles_y,
                                                          uint32_t* __restrict__ list,
                                                          uint32_t* __restrict__ count, int group_levels) {
  extern __shared__ uint32_t levels[];  // l1 | l2 when group_levels
  __shared__ uint32_t run_start[kCullBricks * 4];
  __shared__ uint32_t wg_base;
  __shared__ float wave_max[4];
  const int bnx = (g.nx + kBrickX - 1) / kBrickX, bny = (g.ny + kBrickY - 1) / kBrickY;
  const int bnz = (g.nz + kBrickZ - 1) / kBrickZ;
  const int nb = bnx * bny * bnz;
  Rigid w2c;
  if (!block_w2c(f, w2c)) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // the workgroup's depth pyramid
  const int t1x = (tiles_x + 3) / 4, t1y = (tiles_y + 3) / 4, t2x = (t1x + 3) / 4, t2y = (t1y + 3) / 4;
  const int n1 = t1x * t1y, n_tiles = tiles_x * tiles_y;
  if (group_levels != 0)
    for (int i = static_cast<int>(threadIdx.x); i < n1 + t2x * t2y; i += 256) levels[i] = 0U;
  __syncthreads();
  float fmax_t = 0.0F;
  for (int i0 = 0; i0 < n_tiles; i0 += 8 * 256) {
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = tile_max[min(i0 + k * 256 + static_cast<int>(threadIdx.x), n_tiles - 1)];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int i = i0 + k * 256 + static_cast<int>(threadIdx.x);
      fmax_t = fmaxf(fmax_t, v[k]);
      if (group_levels != 0 && i < n_tiles && v[k] > 0.0F) {
        const int x = i % tiles_x, y = i / tiles_x;
        atomicMax(&levels[(y >> 2) * t1x + (x >> 2)], __float_as_uint(v[k]));
        atomicMax(&levels[n1 + (y >> 4) * t2x + (x >> 4)], __float_as_uint(v[k]));
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) fmax_t = fmaxf(fmax_t, __shfl_xor(fmax_t, off));
  if (lane == 0) wave_max[wave] = fmax_t;
  __syncthreads();
  TilePyramid tp{tile_max, group_levels != 0 ? levels : nullptr, levels + n1, tiles_x, tiles_y, t1x, t2x,
                 fmaxf(fmaxf(wave_max[0], wave_max[1]), fmaxf(wave_max[2], wave_max[3]))};
  unsigned long long kept[kCullBricks];
#pragma unroll
  for (int k = 0; k < kCullBricks; ++k) {
    const int b = (static_cast<int>(blockIdx.x) * kCullBricks + k) * 256 + static_cast<int>(threadIdx.x);
    const bool keep = b < nb && brick_may_fuse(g, f, w2c, tp, b % bnx, (b / bnx) % bny, b / (bnx * bny));
    kept[k] = __ballot(keep);
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < kCullBricks; ++k) run_start[k * 4 + wave] = static_cast<uint32_t>(__popcll(kept[k]));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t total = 0;
    for (int i = 0; i < kCullBricks * 4; ++i) {
      const uint32_t c = run_start[i];
      run_start[i] = total;
      total += c;
    }
    wg_base = total != 0U ? atomicAdd(count, total) : 0U;
  }
  __syncthreads();
  const unsigned long long below = (1ULL << lane) - 1ULL;
#pragma unroll
  for (int k = 0; k < kCullBricks; ++k) {
    if (((kept[k] >> lane) & 1ULL) != 0ULL) {
      const int b = (static_cast<int>(blockIdx.x) * kCullBricks + k) * 256 + static_cast<int>(threadIdx.x);
      list[wg_base + run_start[k * 4 + wave] + static_cast<uint32_t>(__popcll(kept[k] & below))] =
          pack_brick(b % bnx, (b / bnx) % bny, b / (bnx * bny));
    }
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
//    3e-4 + 2e-6 |q'| from a half-integer (q' = fx x / z): u' differs from
//    the exact float u by less than 2e-6 |q'| (the reciprocal and two
//    roundings) plus the float rounding of both (below 1.3e-4 wherever the
//    pixel can be in the image, |u| < 1024); farther out both land outside;
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
    // the nearest pixel; u' lies within the margin of a half-integer when
    // |u' - n| >= 0.5 - margin (a NaN projection fails the test too)
    const float nu = floorf(ua + 0.5F), nv = floorf(va + 0.5F);
    const bool fast = fabsf(ua - nu) < (0.5F - 3.0e-4F) - 2.0e-6F * fabsf(qx) &&
                      fabsf(va - nv) < (0.5F - 3.0e-4F) - 2.0e-6F * fabsf(qy);
    int col, row;
    if (!fast) {
      float u, v;
      project(f.k, cp[k], u, v);
      if (!round_index(u, f.w, col) || !round_index(v, f.h, row)) continue;
    } else {
      col = static_cast<int>(nu);
      row = static_cast<int>(nv);
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
  // One wave per listed brick, the waves independent of each other: lane
  // (x = lane % 16, y = lane / 16) owns the brick's z-column of 8 voxels,
  // both z-halves of its block.  Lanes 8j .. 8j + 7 of each 16 are block j's
  // row (j = 0, 1), so kBlockRows selects block j's 32 lanes of a ballot.
  constexpr unsigned long long kBlockRows = 0x00FF00FF00FF00FFULL;
  Rigid w2c;
  if (!block_w2c(f, w2c)) return;
  const uint32_t n = *count;
  const int lane = threadIdx.x & 63;
  const int tx = lane & 15, ty = lane >> 4;
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  unsigned int my_fused = 0, my_color = 0;
  for (uint32_t li = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); li < n; li += waves) {
    const uint32_t b = list[li];
    const int bx = static_cast<int>(b & 0xFFU), by = static_cast<int>((b >> 8U) & 0x3FFU),
              bz = static_cast<int>(b >> 18U);
    const int x = bx * kBrickX + tx, y = by * kBrickY + ty, z0 = bz * kBrickZ;
    bool any0 = false, any1 = false, ok0 = false, ok1 = false, cpl0 = false, cpl1 = false;
    if (x < g.nx && y < g.ny) {
      any0 = integrate_column<Kind, Color, Count>(g, f, w2c, pix, color_img, vox, colors, x, y, z0, my_fused,
                                                  my_color, ok0, cpl0);
      any1 = integrate_column<Kind, Color, Count>(g, f, w2c, pix, color_img, vox, colors, x, y, z0 + kBrickZ / 2,
                                                  my_fused, my_color, ok1, cpl1);
    }
    const unsigned long long om0 = __ballot(ok0), cm0 = __ballot(cpl0), om1 = __ballot(ok1),
                             cm1 = __ballot(cpl1), am = __ballot(any0 || any1);
    if (lane < 4) {
      // lane = z-half h * 2 + block j: part (y-half, h) of block j, and j's observed bit
      const int j = lane & 1, h = lane >> 1;
      const int blk_x = bx * 2 + j;
      if (blk_x < bm.bx) {
        const unsigned long long rows = kBlockRows << (8 * j);
        const int fb = bm.flat(blk_x, (by * kBrickY) >> 3, z0 >> 3);
        const bool ok = ((h != 0 ? om1 : om0) & rows) == rows, cpl = ((h != 0 ? cm1 : cm0) & rows) == rows;
        const int part = (((by * kBrickY) >> 2) & 1) * 2 + h;
        if (!ok || cpl) set_part(part_bits, fb, part, ok);
        if (h == 0 && (am & rows) != 0ULL) atomicOr(&obs_bits[fb >> 5], 1U << (fb & 31));
      }
    }
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

// A copy of n 16-bit words made by a whole grid on the side of its own work
// (n = 0: none).
struct CopyU16 {
  const uint16_t* src;
  uint16_t* dst;
  int n;
  __device__ inline void run() const {
    const int stride = static_cast<int>(gridDim.x * blockDim.x);
    for (int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += stride) dst[i] = src[i];
  }
};

// The end of a frame: its pose goes to the trajectory slot of the device
// frame counter and becomes the previous pose.
__device__ inline void close_frame(IcpState* st, double* trajectory) {
  const long long slot = st->frames_done;
  for (int i = 0; i < 16; ++i) {
    trajectory[slot * 16 + i] = st->pose[i];
    st->prev2_pose[i] = st->prev_pose[i];
    st->prev_pose[i] = st->pose[i];
  }
  st->frames_done += 1;
}

// free bit of a block = the block and its 26 neighbours (all inside the
// grid) hold only fused voxels reading tsdf exactly 1.
// Bit j of x at bit 2j (j < 16).
__device__ inline uint32_t spread16(uint32_t x) {
  x &= 0xFFFFU;
  x = (x | (x << 8)) & 0x00FF00FFU;
  x = (x | (x << 4)) & 0x0F0F0F0FU;
  x = (x | (x << 2)) & 0x33333333U;
  x = (x | (x << 1)) & 0x55555555U;
  return x;
}
// The class words of 16 blocks from their skip and free bits.
__device__ inline uint32_t class_word(uint32_t skip16, uint32_t free16) {
  return spread16(skip16) | (spread16(free16) << 1);
}

__global__ void __launch_bounds__(256) skip_map_kernel(const uint32_t* __restrict__ obs, uint32_t* __restrict__ skip,
                                                        const uint32_t* __restrict__ parts,
                                                        uint32_t* __restrict__ free_bits,
                                                        uint32_t* __restrict__ class_bits, BlockMap bm, int nblocks,
                                                        IcpState* end_state, double* trajectory, CopyU16 keep) {
  keep.run();
  // end_state: this launch also closes the frame (one thread): the
  // trajectory slot is the device frame counter, so the same launch serves
  // every frame of a captured graph
  if (end_state != nullptr && blockIdx.x == 0 && threadIdx.x == 0) close_frame(end_state, trajectory);
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
  const unsigned long long m_free = __ballot(free_nb);
  {
    const int lane = threadIdx.x & 63;
    const int word0 = (blockIdx.x * 256 + (threadIdx.x & ~63)) >> 5;
    const int nwords = (nblocks + 31) >> 5;
    if (lane == 0 && word0 < nwords) free_bits[word0] = static_cast<uint32_t>(m_free);
    if (lane == 32 && word0 + 1 < nwords) free_bits[word0 + 1] = static_cast<uint32_t>(m_free >> 32);
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
  // lanes 0, 16, 32, 48: the class word of their 16 blocks
  const int cwords = (nblocks + 15) >> 4;
  const int cw = 2 * word0 + (lane >> 4);
  if ((lane & 15) == 0 && cw < cwords)
    class_bits[cw] = class_word(static_cast<uint32_t>(m >> lane), static_cast<uint32_t>(m_free >> lane));
}

// Bit i of the result: block i of the 32 blocks whose part nibbles are the
// four words p has every part set.
__device__ inline uint32_t full_blocks(uint4 p) {
  uint32_t m = 0;
  const uint32_t w[4] = {p.x, p.y, p.z, p.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    uint32_t x = w[q] & (w[q] >> 1);
    x = x & (x >> 2) & 0x11111111U;  // bit 4k: nibble k is 0xF
    x = (x | (x >> 3)) & 0x03030303U;
    x = (x | (x >> 6)) & 0x000F000FU;
    x = (x | (x >> 12)) & 0x000000FFU;
    m |= x << (8 * q);
  }
  return m;
}

// The same maps a word at a time, when every row of blocks (fixed y, z)
// fills whole 32-bit words (bm.bx a multiple of 32): one thread per word.
// skip = no observed block among the 3 x 3 x 3 around (the rows around are
// OR-ed, then spread by one block along x); free = every block of the
// 3 x 3 x 3 full (the rows around are AND-ed, then shrunk by one block along
// x; blocks at the grid's faces are never free).
__global__ void __launch_bounds__(64) skip_map_rows_kernel(const uint32_t* __restrict__ obs,
                                                             uint32_t* __restrict__ skip,
                                                             const uint32_t* __restrict__ parts,
                                                             uint32_t* __restrict__ free_bits,
                                                             uint32_t* __restrict__ class_bits, BlockMap bm,
                                                             IcpState* end_state, double* trajectory, CopyU16 keep) {
  if (end_state != nullptr && blockIdx.x == 0 && threadIdx.x == 0) close_frame(end_state, trajectory);
  keep.run();
  const int wpr = bm.bx >> 5;  // words per row
  const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= wpr * bm.by * bm.bz) return;
  const int wx = t % wpr, row = t / wpr, y = row % bm.by, z = row / bm.by;
  const bool interior = y > 0 && z > 0 && y + 1 < bm.by && z + 1 < bm.bz;
  // every read of the 9 rows around is issued at once: a row outside the
  // grid, or a neighbouring word outside the row, reads the thread's own
  // word and is masked out
  uint32_t o[9], o_lo[9], o_hi[9], p_lo[9], p_hi[9];
  uint4 p[9];
  bool in[9];
  const auto* parts4 = reinterpret_cast<const uint4*>(parts);
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int yy = y + k % 3 - 1, zz = z + k / 3 - 1;
    in[k] = yy >= 0 && zz >= 0 && yy < bm.by && zz < bm.bz;
    const int w = in[k] ? (zz * bm.by + yy) * wpr + wx : t;
    o[k] = obs[w];
    o_lo[k] = obs[wx > 0 ? w - 1 : w];
    o_hi[k] = obs[wx + 1 < wpr ? w + 1 : w];
    p[k] = parts4[w];
    p_lo[k] = parts[wx > 0 ? 4 * w - 1 : 4 * w];
    p_hi[k] = parts[wx + 1 < wpr ? 4 * w + 4 : 4 * w];
  }
  uint32_t seen = 0U, seen_lo = 0U, seen_hi = 0U;  // lo / hi: the neighbouring words' edge blocks
  uint32_t full = 0xFFFFFFFFU, full_lo = wx > 0 ? 1U : 0U, full_hi = wx + 1 < wpr ? 1U : 0U;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    if (!in[k]) continue;
    seen |= o[k];
    if (wx > 0) seen_lo |= o_lo[k] >> 31;
    if (wx + 1 < wpr) seen_hi |= o_hi[k] & 1U;
    full &= full_blocks(p[k]);
    full_lo &= (p_lo[k] >> 28) == 0xFU ? 1U : 0U;
    full_hi &= (p_hi[k] & 0xFU) == 0xFU ? 1U : 0U;
  }
  const uint32_t sk = ~(seen | (seen << 1) | (seen >> 1) | seen_lo | (seen_hi << 31));
  const uint32_t fr = interior ? full & ((full << 1) | full_lo) & ((full >> 1) | (full_hi << 31)) : 0U;
  skip[t] = sk;
  free_bits[t] = fr;
  class_bits[2 * t] = class_word(sk, fr);
  class_bits[2 * t + 1] = class_word(sk >> 16, fr >> 16);
}

// coarse bit of a cell of 4^3 blocks = every block of it (inside the grid)
// is a skip block.  One thread per cell; a wave's 64 consecutive cells form
// two words.
__global__ void __launch_bounds__(256) coarse_skip_kernel(const uint32_t* __restrict__ skip,
                                                           uint32_t* __restrict__ coarse, BlockMap bm, BlockMap cm,
                                                           int ncells) {
  const int c = static_cast<int>(blockIdx.x * 256 + threadIdx.x);
  bool all = false;
  if (c < ncells) {
    const int cx = c % cm.bx, cy = (c / cm.bx) % cm.by, cz = c / (cm.bx * cm.by);
    all = true;
    for (int z = cz * 4; z < min(cz * 4 + 4, bm.bz) && all; ++z)
      for (int y = cy * 4; y < min(cy * 4 + 4, bm.by) && all; ++y)
        for (int x = cx * 4; x < min(cx * 4 + 4, bm.bx); ++x)
          if (!bm.bit(skip, bm.flat(x, y, z))) {
            all = false;
            break;
          }
  }
  const unsigned long long m = __ballot(all);
  const int lane = threadIdx.x & 63;
  const int word0 = (blockIdx.x * 256 + (threadIdx.x & ~63)) >> 5;
  const int nwords = (ncells + 31) >> 5;
  if (lane == 0 && word0 < nwords) coarse[word0] = static_cast<uint32_t>(m);
  if (lane == 32 && word0 + 1 < nwords) coarse[word0 + 1] = static_cast<uint32_t>(m >> 32);
}

// Neighbour masks of the non-skip blocks: bit (dz + 1) * 9 + (dy + 1) * 3 +
// (dx + 1) = the block at that offset lies in the grid and holds an observed
// voxel.  A skip block keeps its old mask: the raycaster uses a mask only for
// a non-skip block.  One thread per block.
__global__ void __launch_bounds__(256) neighbour_mask_kernel(const uint32_t* __restrict__ obs,
                                                              const uint32_t* __restrict__ skip,
                                                              uint32_t* __restrict__ nbr, BlockMap bm, int nblocks) {
  const int b = static_cast<int>(blockIdx.x * 256 + threadIdx.x);
  if (b >= nblocks || bm.bit(skip, b)) return;
  const int x = b % bm.bx, y = (b / bm.bx) % bm.by, z = b / (bm.bx * bm.by);
  uint32_t m = 0U;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int xx = x + dx, yy = y + dy, zz = z + dz;
        if (xx < 0 || yy < 0 || zz < 0 || xx >= bm.bx || yy >= bm.by || zz >= bm.bz) continue;
        if (bm.bit(obs, bm.flat(xx, yy, zz))) m |= 1U << ((dz + 1) * 9 + (dy + 1) * 3 + (dx + 1));
      }
  nbr[b] = m;
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

// Exclusive prefix sum of v over the 256 threads of a workgroup, and the
// workgroup's total (`sums`: 4 words of LDS).
__device__ inline uint32_t workgroup_scan(uint32_t v, uint32_t* sums, uint32_t& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;  // inclusive scan within the wave
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  __syncthreads();  // the words may still hold an earlier scan's totals
  if (lane == 63) sums[wave] = x;
  __syncthreads();
  uint32_t before = 0;
  for (int w = 0; w < wave; ++w) before += sums[w];
  total = sums[0] + sums[1] + sums[2] + sums[3];
  return before + x - v;
}

// Pool slots for this frame's new blocks and the frame's work list (each
// visited block once).  A workgroup scans kAssignRun consecutive buckets
// per thread and takes its share of the slot and list counters with one
// atomic each: with one atomic per wave, the waves' appends to the same two
// words serialised (171 us per frame for 2^20 buckets in the 2048^3 hall).
constexpr int kAssignRun = 16;
__global__ void __launch_bounds__(256) assign_slots_kernel(HashView hv, const uint32_t* __restrict__ stamp,
                                                            uint32_t* __restrict__ sp,
                                                            uint32_t* __restrict__ block_of_slot,
                                                            uint32_t* __restrict__ work, uint32_t capacity,
                                                            uint32_t buckets) {
  __shared__ uint32_t sums[4];
  __shared__ uint32_t base;
  const uint32_t frame = sp[3];
  const uint32_t first = blockIdx.x * (256U * kAssignRun) + threadIdx.x;
  unsigned long long e[kAssignRun];
  uint32_t visited = 0U, fresh = 0U;  // bit r: bucket first + 256 r
#pragma unroll
  for (int r = 0; r < kAssignRun; ++r) {
    const uint32_t b = first + 256U * r;
    e[r] = kHashEmpty;
    if (b < buckets) {
      e[r] = hv.table[b];
      if (e[r] != kHashEmpty && stamp[b] == frame) {
        visited |= 1U << r;
        if (static_cast<uint32_t>(e[r]) == kPendingSlot) fresh |= 1U << r;
      }
    }
  }
  uint32_t total;
  uint32_t s = workgroup_scan(static_cast<uint32_t>(__popc(fresh)), sums, total);
  if (threadIdx.x == 0) base = total != 0U ? atomicAdd(&sp[0], total) : 0U;
  __syncthreads();
  s += base;
  uint32_t listed = 0U;  // bit r: the bucket's block goes to the work list
#pragma unroll
  for (int r = 0; r < kAssignRun; ++r) {
    uint32_t slot = static_cast<uint32_t>(e[r]);
    if (((fresh >> r) & 1U) != 0U) {
      const uint32_t key = static_cast<uint32_t>(e[r] >> 32);
      if (s < capacity) {
        slot = s;
        block_of_slot[s] = key;
      } else {
        slot = kNoSlot;
        atomicAdd(&sp[4], 1U);
      }
      ++s;
      hv.table[first + 256U * r] = (static_cast<unsigned long long>(key) << 32) | slot;
      e[r] = (static_cast<unsigned long long>(key) << 32) | slot;
    }
    if (((visited >> r) & 1U) != 0U && slot != kNoSlot) listed |= 1U << r;
  }
  uint32_t w = workgroup_scan(static_cast<uint32_t>(__popc(listed)), sums, total);
  __syncthreads();  // every thread has read `base`
  if (threadIdx.x == 0) base = total != 0U ? atomicAdd(&sp[1], total) : 0U;
  __syncthreads();
  w += base;
#pragma unroll
  for (int r = 0; r < kAssignRun; ++r)
    if (((listed >> r) & 1U) != 0U) work[w++] = static_cast<uint32_t>(e[r]);
}

// Every voxel of the listed blocks (BlockVoxels, clipped to the grid): one
// wave per block, each lane the z-column of one (x, y) of it (two runs of 4
// voxels with the column's shared transform), so for every z the wave's
// voxels are contiguous.  The block's flags are wave ballots: no workgroup
// barrier (one workgroup per block with three barriers per block was
// latency-bound at 100 us per hall frame).  Hashed (Dense = false): the list
// holds pool slots, block_of_slot their keys, and a block's voxels are its
// pool slot's 512.  Dense band: the list holds block keys and voxels sit at
// their dense indices.  n: the list's length.
template <int Kind, bool Color, bool Count, bool Dense>
__global__ void __launch_bounds__(256) integrate_blocks_kernel(Grid g, FrameParams f, const float4* __restrict__ pix,
                                                                const uint32_t* __restrict__ color_img,
                                                                typename VoxelIO<Kind>::Raw* __restrict__ pool,
                                                                float4* __restrict__ cpool,
                                                                const uint32_t* __restrict__ block_of_slot,
                                                                const uint32_t* __restrict__ work,
                                                                const uint32_t* __restrict__ count,
                                                                unsigned long long* __restrict__ fused,
                                                                uint32_t* __restrict__ obs_bits, BlockMap bm,
                                                                uint32_t* __restrict__ part_bits) {
  Rigid w2c;
  if (!block_w2c(f, w2c)) return;
  const uint32_t n = *count;
  const int lane = static_cast<int>(threadIdx.x & 63);
  const uint32_t waves = gridDim.x * 4U;
  unsigned int my_fused = 0, my_color = 0;
  for (uint32_t li = blockIdx.x * 4U + (threadIdx.x >> 6); li < n; li += waves) {
    const uint32_t slot = work[li];
    const uint32_t key = Dense ? slot : block_of_slot[slot];
    const int bx = static_cast<int>(key % static_cast<uint32_t>(bm.bx));
    const int by = static_cast<int>((key / static_cast<uint32_t>(bm.bx)) % static_cast<uint32_t>(bm.by));
    const int bz = static_cast<int>(key / static_cast<uint32_t>(bm.bx * bm.by));
    const int x = bx * 8 + (lane & 7), y = by * 8 + (lane >> 3);
    bool any = false, ones_ok = true, complete = true;
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      int xs[4], ys[4], zs[4];
      size_t vis[4];
      bool inside[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int zi = half * 4 + k;
        xs[k] = x;
        ys[k] = y;
        zs[k] = bz * 8 + zi;
        inside[k] = x < g.nx && y < g.ny && zs[k] < g.nz;
        if constexpr (Dense) vis[k] = inside[k] ? g.flat(x, y, zs[k]) : 0;
        else vis[k] = static_cast<size_t>(slot) * kBlockVoxels + static_cast<size_t>(zi * 64 + lane);
      }
      bool a = false, o = true, c = true;
      fuse_batch<Kind, Color, Count, 4, true>(g, f, w2c, pix, color_img, pool, cpool, xs, ys, zs, vis, inside,
                                              my_fused, my_color, a, o, c);
      any = any || a;
      ones_ok = ones_ok && o;
      complete = complete && c;
    }
    const bool any_fused = __ballot(any) != 0ULL, all_ok = __ballot(!ones_ok) == 0ULL,
               all_complete = __ballot(!complete) == 0ULL;
    if (lane == 0) {
      const int fb = bm.flat(bx, by, bz);
      if (any_fused) atomicOr(&obs_bits[fb >> 5], 1U << (fb & 31));
      const uint32_t nib = 0xFU << ((fb & 7) * 4);
      if (!all_ok) atomicAnd(&part_bits[fb >> 3], ~nib);
      else if (all_complete) atomicOr(&part_bits[fb >> 3], nib);  // else unchanged
    }
  }
  if (Count) {
    for (int off = 32; off > 0; off >>= 1) {
      my_fused += __shfl_xor(my_fused, off);
      my_color += __shfl_xor(my_color, off);
    }
    if (lane == 0) {
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

template <int Kind, bool Color, bool Dense>
void launch_blocks_kernel(kf_context* ctx, const FrameParams& f, const uint32_t* color_img, bool count,
                          const uint32_t* work, const uint32_t* n) {
  using Raw = typename VoxelIO<Kind>::Raw;
  const int grid = 2048;
  if (count) {
    hipLaunchKernelGGL((integrate_blocks_kernel<Kind, Color, true, Dense>), dim3(grid), dim3(256), 0, ctx->stream,
                       ctx->grid, f, ctx->pix, color_img, static_cast<Raw*>(ctx->voxels), ctx->colors,
                       ctx->block_of_slot, work, n, ctx->fused, ctx->obs_bits, block_map(ctx->grid),
                       ctx->part_bits);
  } else {
    hipLaunchKernelGGL((integrate_blocks_kernel<Kind, Color, false, Dense>), dim3(grid), dim3(256), 0, ctx->stream,
                       ctx->grid, f, ctx->pix, color_img, static_cast<Raw*>(ctx->voxels), ctx->colors,
                       ctx->block_of_slot, work, n, ctx->fused, ctx->obs_bits, block_map(ctx->grid),
                       ctx->part_bits);
  }
}

template <int Kind, bool Color>
void launch_blocks_t(kf_context* ctx, const FrameParams& f, const uint32_t* color_img, bool count) {
  if (hashed(ctx)) launch_blocks_kernel<Kind, Color, false>(ctx, f, color_img, count, ctx->work, ctx->sparse + 1);
  else launch_blocks_kernel<Kind, Color, true>(ctx, f, color_img, count, ctx->band_list, ctx->counters);
}

// ---- dense band integration: BandIntegrationSweep<kHost>
// (tsdf_integration.hpp:444-497) -------------------------------------------
// Marks the blocks of every pixel's truncation band in a bitmap
// (TruncationBandWalk::visit_pixel at a one-voxel step, the host sweep's
// step, tsdf_integration.hpp:462-463), skipping a block equal to the pixel's
// previous visit.
__global__ void __launch_bounds__(256) band_mark_kernel(Grid g, FrameParams f, const uint16_t* __restrict__ depth,
                                                         uint32_t* __restrict__ bits, BlockMap bm, float step) {
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
  int previous = -1;
  for (int s = 0; s <= samples; ++s) {
    const float d = (measured - band) + (static_cast<float>(s) * step);
    const V3 world = xform(c2w, back_project(f.k, static_cast<float>(col), static_cast<float>(row), d));
    const V3 gc = scale3(sub3(world, mk3(g.ox, g.oy, g.oz)), inv_voxel);
    if (!(gc.x >= 0.0F && gc.y >= 0.0F && gc.z >= 0.0F && gc.x < static_cast<float>(g.nx) &&
          gc.y < static_cast<float>(g.ny) && gc.z < static_cast<float>(g.nz)))
      continue;
    const int b = bm.flat(static_cast<int>(gc.x) >> 3, static_cast<int>(gc.y) >> 3, static_cast<int>(gc.z) >> 3);
    if (b == previous) continue;
    previous = b;
    if (!bm.bit(bits, b)) atomicOr(&bits[b >> 5], 1U << (b & 31));
  }
}

// The marked blocks as a work list (each once), one thread per bitmap word;
// the word is cleared as it is read, so the bitmap starts the next frame
// empty.  A wave's appends take one range of the list (an inclusive scan of
// the per-lane counts).  The list order does not matter: blocks are
// disjoint.
__global__ void __launch_bounds__(256) band_compact_kernel(uint32_t* __restrict__ bits, int words,
                                                            uint32_t* __restrict__ list, uint32_t* __restrict__ count) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 256 + static_cast<int>(threadIdx.x);
  uint32_t w = 0U;
  if (i < words) {
    w = bits[i];
    if (w != 0U) bits[i] = 0U;
  }
  const int c = __popc(w);
  int incl = c;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int v = __shfl_up(incl, off);
    if (lane >= off) incl += v;
  }
  const int total = __shfl(incl, 63);
  uint32_t base = 0U;
  if (lane == 63 && total > 0) base = atomicAdd(count, static_cast<uint32_t>(total));
  base = __shfl(base, 63);
  uint32_t at = base + static_cast<uint32_t>(incl - c);
  while (w != 0U) {
    const int b = __ffs(static_cast<int>(w)) - 1;
    w &= w - 1U;
    list[at++] = static_cast<uint32_t>(i) * 32U + static_cast<uint32_t>(b);
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
  // 24 K waves (the launch sees no brick count: the list is built on the
  // device): up to that many listed bricks (about 21 K in a 512^3 room
  // at 640x480) each get a wave of their own, and the dispatcher starts the
  // waves as slots free up, which balances bricks of unequal cost.  Measured
  // faster than a grid of 2048 workgroups or than one sized to the resident
  // waves, whose waves stride over several bricks each.  The counting
  // variant adds two counter atomics per wave, so it keeps 2048 workgroups.
  const int grid = count ? 2048 : 6144;
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
  // the frame's end keeps its coarse depth for the next frame's proposal
  CopyU16 keep{nullptr, nullptr, 0};
  if (end_frame && ctx->cfg.init_mode == KF_INIT_LEARNED && ctx->prev_coarse != nullptr) {
    const LevelBuffers& L = ctx->lv[ctx->cfg.levels - 1];
    keep = CopyU16{L.depth, ctx->prev_coarse, L.w * L.h};
  }
  if (bm.bx % 32 == 0) {
    const int words = nblocks / 32;
    hipLaunchKernelGGL(skip_map_rows_kernel, dim3((words + 63) / 64), dim3(64), 0, ctx->stream, ctx->obs_bits,
                       ctx->skip_bits, ctx->part_bits, ctx->free_bits, ctx->class_bits, bm,
                       end_frame ? ctx->icp : nullptr,
                       ctx->trajectory, keep);
  } else {
    hipLaunchKernelGGL(skip_map_kernel, dim3((nblocks + 255) / 256), dim3(256), 0, ctx->stream, ctx->obs_bits,
                       ctx->skip_bits, ctx->part_bits, ctx->free_bits, ctx->class_bits, bm, nblocks,
                       end_frame ? ctx->icp : nullptr,
                       ctx->trajectory, keep);
  }
  if (ctx->coarse_bits != nullptr) {
    const BlockMap cm = coarse_map(ctx->grid);
    const int ncells = cm.bx * cm.by * cm.bz;
    hipLaunchKernelGGL(coarse_skip_kernel, dim3((ncells + 255) / 256), dim3(256), 0, ctx->stream, ctx->skip_bits,
                       ctx->coarse_bits, bm, cm, ncells);
  }
  if (ctx->nbr_obs != nullptr)
    hipLaunchKernelGGL(neighbour_mask_kernel, dim3((nblocks + 255) / 256), dim3(256), 0, ctx->stream, ctx->obs_bits,
                       ctx->skip_bits, ctx->nbr_obs, bm, nblocks);
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
                      bool accumulate, bool end_learned) {
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
  f.frame_w2c = ctx->frame_w2c;
  const int tiles_x = (f.w + kTile - 1) / kTile, tiles_y = (f.h + kTile - 1) / kTile;
  // per-pixel and per-tile scratch is sized for the configured camera
  if (f.w > c.width || f.h > c.height)
    fail(KF_ERR_INVALID_ARGUMENT, "integrate: frame larger than the configured camera");
  hipLaunchKernelGGL(pixel_prep_kernel, dim3(tiles_x, tiles_y), dim3(kTile, kTile), 0, ctx->stream,
                     fr.depth, reinterpret_cast<const float4*>(fr.normals), ctx->pix, ctx->tile_max,
                     f, ctx->grid.trunc, ctx->counters, ctx->fused, end_learned ? ctx->icp : nullptr);
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
    const uint32_t per_group = 256U * kAssignRun;
    hipLaunchKernelGGL(assign_slots_kernel, dim3((static_cast<uint32_t>(ctx->buckets) + per_group - 1) / per_group),
                       dim3(256), 0, ctx->stream, hash_view(ctx), ctx->hash_stamp,
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
  if (c.integration_mode == KF_INTEGRATION_BAND) {
    // band walk -> block list -> listed blocks; the list count is the brick
    // counter, which the pixel pass zeroed
    const BlockMap bm = block_map(ctx->grid);
    const int words = static_cast<int>(ctx->block_words);
    hipLaunchKernelGGL(band_mark_kernel, dim3(tiles_x, tiles_y), dim3(kTile, kTile), 0, ctx->stream, ctx->grid, f,
                       fr.depth, ctx->band_bits, bm, ctx->grid.voxel);
    hipLaunchKernelGGL(band_compact_kernel, dim3((words + 255) / 256), dim3(256), 0, ctx->stream, ctx->band_bits,
                       words, ctx->band_list, ctx->counters);
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
  const Grid& g = ctx->grid;
  const int nb = ((g.nx + kBrickX - 1) / kBrickX) * ((g.ny + kBrickY - 1) / kBrickY) *
                 ((g.nz + kBrickZ - 1) / kBrickZ);
  const int t1x = (tiles_x + 3) / 4, t1y = (tiles_y + 3) / 4;
  const size_t level_bytes = (static_cast<size_t>(t1x) * t1y + ((t1x + 3) / 4) * ((t1y + 3) / 4)) * sizeof(uint32_t);
  const bool group_levels = level_bytes <= static_cast<size_t>(kPyramidLdsBytes);
  hipLaunchKernelGGL(brick_cull_kernel, dim3((nb + 256 * kCullBricks - 1) / (256 * kCullBricks)), dim3(256),
                     group_levels ? level_bytes : 0, ctx->stream, g, f, ctx->tile_max, tiles_x, tiles_y, ctx->bricks,
                     ctx->counters, group_levels ? 1 : 0);
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

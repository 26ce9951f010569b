// The raycast of one pixel (reference: SurfaceRaycast::render_pixel /
// find_zero_crossing / write_surface_sample, include/kinectfusion/
// raycasting.hpp:116-245).  Corner blends visit the 2^3 stencil in
// z -> y -> x order with the reference's weight products, so every sample is
// bit-identical to the CPU sampler.
#pragma once

#include "kf_sampler.hpp"
#include "kf_steps.hpp"

namespace kf {

struct RayParams {
  Intr k;
  int w, h;
  float min_depth, max_depth, step_scale;
  int require_all;
  int cell_normals;
  int seed_previous;
};

__device__ inline uint32_t pixel_from_color(V3 c) {
  auto to_byte = [](float v) { return static_cast<uint32_t>(fclamp_ref(v, 0.0F, 255.0F)); };
  return to_byte(c.x) | (to_byte(c.y) << 8U) | (to_byte(c.z) << 16U) | (0xFFU << 24U);
}

// One axis of a box's safe interior: the box covers grid coordinates
// [first, min(first + size, n)); the ray's coordinate o + d * len lies at
// least `safe` inside it for len in [enter, exit), intersected into the
// running interval.  A ray (almost) parallel to the axis keeps its coordinate.
__device__ inline void box_span(int first, int size, float o, float d, int n, float safe, float& enter,
                                float& exit) {
  const float lo = static_cast<float>(first) + safe, hi = static_cast<float>(min(first + size, n)) - safe;
  if (fabsf(d) <= 1.0e-12F) {
    if (o < lo || o > hi) exit = -3.0e38F;
    return;
  }
  const float r = __builtin_amdgcn_rcpf(d);  // 1 ulp: far inside the margin
  const float t0 = (lo - o) * r, t1 = (hi - o) * r;
  enter = fmaxf(enter, fminf(t0, t1));
  exit = fminf(exit, fmaxf(t0, t1));
}

// Pixel (x, y) of the maps (pts, nrm, col) rendered from `pose`.
template <int Kind, typename Store>
__device__ inline void raycast_pixel(const Sampler<Kind, Store>& vol, const RayParams& rp, const Rigid& pose, int x,
                                     int y, float4* __restrict__ pts, float4* __restrict__ nrm,
                                     uint32_t* __restrict__ col, const uint32_t* __restrict__ class_bits,
                                     const uint32_t* __restrict__ nbr_obs, const BlockMap& bm,
                                     const uint32_t* __restrict__ coarse_bits, const BlockMap& cm) {
  const size_t i = static_cast<size_t>(y) * rp.w + x;
  // RaycastOptions::seed_from_previous (raycasting.hpp:122-131): start just
  // in front of the surface this pixel found in the previous raycast into
  // this map (a zero point is the maps' initial fill, not a surface)
  float seed = rp.min_depth;
  if (rp.seed_previous != 0) {
    const V3 previous = ld3(pts[i]);
    if (finite3(previous) && dot3(previous, previous) > 0.0F) {
      const float previous_depth = norm3(sub3(previous, pose.t));
      seed = fmax_ref(rp.min_depth, previous_depth - (2.0F * vol.g.trunc));
    }
  }
  V3 P = nan3(), N = nan3();
  uint32_t C = 0U;
  const V3 translated = back_project(rp.k, static_cast<float>(x), static_cast<float>(y), 1.0F);
  const V3 direction = rot(pose, translated);
  const float dn = norm3(direction);
  if (dn > 0.0F) {
    const V3 dir = div3(direction, dn);
    const float base_step = vol.g.voxel * rp.step_scale;
    bool have_prev = false, hit = false;
    float prev_t = 0.0F;
    V3 prev_p = mk3(0.0F, 0.0F, 0.0F), surface = prev_p;
    float len = seed;
    // Samples whose outcome is known are stepped over without touching the
    // volume, reproducing the reference's float sequence of ray lengths:
    // before the ray enters the volume box, after it has left it (it cannot
    // come back), inside 8^3 blocks whose 3x3x3 block neighbourhood holds no
    // observed voxel (no TSDF), and inside blocks whose neighbourhood holds
    // only observed voxels reading exactly 1 (the trilinear blend of eight
    // corners of 1 is exactly 1: the weighted sum and the weight sum are the
    // same float sum).  The output is identical to marching every sample.
    const double inv = 1.0 / static_cast<double>(vol.g.voxel);
    const double ga[3] = {(static_cast<double>(pose.t.x) - vol.g.ox) * inv,
                          (static_cast<double>(pose.t.y) - vol.g.oy) * inv,
                          (static_cast<double>(pose.t.z) - vol.g.oz) * inv};
    const double gb[3] = {static_cast<double>(dir.x) * inv, static_cast<double>(dir.y) * inv,
                          static_cast<double>(dir.z) * inv};
    const double dims[3] = {static_cast<double>(vol.g.nx), static_cast<double>(vol.g.ny),
                            static_cast<double>(vol.g.nz)};
    double t_in = -1.0e300, t_out = 1.0e300, slope_in = 1.0, slope_out = 1.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (fabs(gb[k]) <= 1.0e-12) {
        if (ga[k] < 0.0 || ga[k] > dims[k]) t_out = -1.0e300;  // parallel and outside
        continue;
      }
      double t0 = -ga[k] / gb[k], t1 = (dims[k] - ga[k]) / gb[k];
      if (t0 > t1) {
        const double tmp = t0;
        t0 = t1;
        t1 = tmp;
      }
      if (t0 > t_in) {
        t_in = t0;
        slope_in = fabs(gb[k]);
      }
      if (t1 < t_out) {
        t_out = t1;
        slope_out = fabs(gb[k]);
      }
    }
    // margins of 0.01 voxel along the deciding axis absorb the rounding of
    // the float grid coordinates the reference compares
    const float len_stop = t_out < t_in ? -1.0F : static_cast<float>(fmin(t_out + 0.01 / slope_out, 1.0e30));
    const double len_enter = t_in - 0.01 / slope_in;
    while (static_cast<double>(len) < len_enter && len <= rp.max_depth) len += base_step;
    const float free_step = fmax_ref(base_step, (1.0F * vol.g.trunc) * rp.step_scale);
    int cur_blk = -1;
    bool cur_skip = false, cur_free = false, cur_empty = false;
    uint32_t cur_nbr = 0U;  // observed blocks among the current block's 3^3 (hashed store)
    SlotCache slots;  // the last block a sample probed (hashed store)
    // The samples of a skip or free block that lie deeper than kSafe voxels
    // inside it on every axis are stepped in one run (the float additions of
    // the reference's loop, bit for bit): [enter, exit) is that stretch of
    // ray lengths, from the
    // ray in grid coordinates, go + gd * len.  The grid coordinates the
    // march computes for a sample differ from the exact ones by a few float
    // ulps of the magnitudes involved (< 4e-3 voxel for points within 2^14
    // voxels of the grid origin), and the interval's own rounding adds as
    // little, so every sample in it falls in the block as the per-sample
    // path would classify it.
    //
    // A block that holds no observed voxel but has an observed neighbour
    // ("empty") is no skip block, yet a sample whose 2^3 stencil lies inside
    // it reads no observed corner either: the samples deeper than 0.5 + kSafe
    // voxels inside it (the stencil base floor(g - 0.5) and its +1 corner
    // then stay in the block) are stepped the same way, each a sample
    // without TSDF.  nbr_obs holds, per block, which of its 3^3 blocks hold
    // an observed voxel; a sample of an empty block whose stencil reaches
    // only unobserved blocks is a sample without TSDF as well, decided
    // without a read.
    //
    // coarse_bits (optional) marks cells of 4^3 blocks that are all skip
    // blocks: a skip block in such a cell is stepped over the whole cell's
    // safe interior at once, so a ray crossing unobserved space reads one
    // class per 32^3-voxel cell instead of one per block.
    constexpr float kSafe = 0.05F;
    // len <= max_depth is len < depth_lim: a run of steps lasts while len < min(exit, depth_lim)
    const float depth_lim = nextafterf(rp.max_depth, __builtin_inff());
    float unused_last;
    // A run of steps: a direct store's runs are short and a plain loop is
    // faster there (the closed form cost 1.5 % of the headline frame rate);
    // a hashed store's rays cross long unobserved runs, where advance_to()
    // gained 5.6 % in the 2048^3 hall.
    auto run_to = [](float l, float step, float lim, float& last) {
      if constexpr (Store::kDirect) {
        while (l < lim) {
          last = l;
          l += step;
        }
        return l;
      } else {
        return advance_to(l, step, lim, last);
      }
    };
    // Empty blocks pay where a sample costs a hash probe before its corner
    // reads; a direct store reads the corners at once, and the extra class
    // word per block was slower there (-1.4 % headline frame rate).
    constexpr bool kEmptyClass = !Store::kDirect;
    float enter = 0.0F, exit = -1.0F;
    while (len <= rp.max_depth) {
      if (len > len_stop) break;
      const V3 point = add3(pose.t, scale3(dir, len));
      const V3 gc = vol.grid_coords(point);
      if (vol.contains(gc)) {
        // the block of this sample, classified exactly as the sampler sees it
        const int gx = static_cast<int>(gc.x) >> 3, gy = static_cast<int>(gc.y) >> 3, gz = static_cast<int>(gc.z) >> 3;
        const int fb = bm.flat(gx, gy, gz);
        if (fb != cur_blk) {
          // the block's skip and free bits: one word, one read
          cur_blk = fb;
          const uint32_t cls = class_bits[fb >> 4] >> (2U * static_cast<uint32_t>(fb & 15));
          // a skip block holds no observed voxel, a free block only
          // observed ones: the three bits give four exclusive classes
          const uint32_t skip = cls & 1U, fre = (cls >> 1) & 1U;
          if constexpr (kEmptyClass) cur_nbr = nbr_obs[fb];
          const uint32_t seen = kEmptyClass ? (cur_nbr >> 13) & 1U : 1U;
          bool coarse = false;
          if (coarse_bits != nullptr) {
            const int cf = cm.flat(gx >> 2, gy >> 2, gz >> 2);
            coarse = ((coarse_bits[cf >> 5] >> (cf & 31)) & 1U) != 0U;
          }
          cur_skip = skip != 0U;
          cur_free = fre != 0U;
          cur_empty = (skip | seen) == 0U;
          if (cur_skip || cur_free || cur_empty) {
            const V3 go = vol.grid_coords(pose.t), gd = scale3(dir, vol.inv_voxel);
            const float margin = cur_empty ? 0.5F + kSafe : kSafe;
            const int size = coarse ? 32 : 8, sh3 = coarse ? 5 : 3;
            enter = -3.0e38F;
            exit = 3.0e38F;
            box_span((gx << 3) >> sh3 << sh3, size, go.x, gd.x, vol.g.nx, margin, enter, exit);
            box_span((gy << 3) >> sh3 << sh3, size, go.y, gd.y, vol.g.ny, margin, enter, exit);
            box_span((gz << 3) >> sh3 << sh3, size, go.z, gd.z, vol.g.nz, margin, enter, exit);
          }
        }
        if (cur_skip) {  // no corner is observed: no TSDF
          have_prev = false;
          len += base_step;
          if (len >= enter) len = run_to(len, base_step, fminf(exit, depth_lim), unused_last);
          continue;
        }
        if (cur_free) {  // every corner reads exactly 1: no read, free steps
          if (have_prev && prev_t < 0.0F) break;
          have_prev = true;
          prev_t = 1.0F;
          float last = len;
          len += free_step;
          if (len >= enter) len = run_to(len, free_step, fminf(exit, depth_lim), last);
          prev_p = add3(pose.t, scale3(dir, last));  // the last free sample's point
          continue;
        }
        if (cur_empty && len >= enter && len < exit) {  // a stencil inside an empty block: no TSDF
          have_prev = false;
          len += base_step;
          len = run_to(len, base_step, fminf(exit, depth_lim), unused_last);
          continue;
        }
        if (cur_empty) {
          // the stencil's corners c0 and c0 + 1 per axis (c0 = floor(g - 0.5),
          // as the sampler computes it) lie in the blocks -1, 0 or +1 from
          // this one; bit (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1) of `reach`
          // marks each block the stencil reaches
          auto axis = [](float g, int b) {
            const int c0 = static_cast<int>(floorf(g - 0.5F));
            return (1U << ((c0 >> 3) - b + 1)) | (1U << (((c0 + 1) >> 3) - b + 1));
          };
          const uint32_t xs = axis(gc.x, gx), ys = axis(gc.y, gy), zs = axis(gc.z, gz);
          const uint32_t yx = ((ys & 1U) != 0U ? xs : 0U) | ((ys & 2U) != 0U ? xs << 3 : 0U) |
                              ((ys & 4U) != 0U ? xs << 6 : 0U);
          const uint32_t reach = ((zs & 1U) != 0U ? yx : 0U) | ((zs & 2U) != 0U ? yx << 9 : 0U) |
                                 ((zs & 4U) != 0U ? yx << 18 : 0U);
          if ((reach & cur_nbr) == 0U) {  // no observed corner: no TSDF
            have_prev = false;
            len += base_step;
            continue;
          }
        }
      }
      float t;
      if (!vol.tsdf(point, rp.require_all != 0, t, slots)) {
        have_prev = false;
        len += base_step;
        continue;
      }
      if (have_prev && prev_t > 0.0F && t <= 0.0F) {
        const float gap = prev_t - t;
        const float alpha = gap != 0.0F ? prev_t / gap : 0.0F;
        surface = add3(prev_p, scale3(sub3(point, prev_p), alpha));
        hit = true;
        break;
      }
      if (have_prev && prev_t < 0.0F && t > 0.0F) break;
      have_prev = true;
      prev_t = t;
      prev_p = point;
      const float pre = t * vol.g.trunc * rp.step_scale;
      const float step = fmax_ref(base_step, pre);
      len += t > 0.0F ? step : base_step;
    }
    if (hit) {
      V3 n;
      const bool ok = rp.cell_normals ? vol.cell_normal(surface, n)
                                      : vol.normal(surface, rp.require_all != 0, n);
      if (ok) {
        P = surface;
        N = n;
        if (vol.s.has_colors()) {
          V3 c;
          if (vol.color(surface, c)) C = pixel_from_color(c);
        } else {
          const V3 view = normalize3(sub3(pose.t, surface));
          const float inten = 255.0F * fmax_ref(0.0F, dot3(n, view));
          C = pixel_from_color(mk3(inten, inten, inten));
        }
      }
    }
  }
  pts[i] = st3(P);
  nrm[i] = st3(N);
  col[i] = C;
}

// Pixel of thread t (0..255) of 16 x 16 tile `tile` in a grid `tiles_x`
// tiles wide: each wave takes an 8 x 8 quadrant (neighbouring rays march
// alike, so a squarer wave diverges less than a 16 x 4 strip).
__device__ inline void tile_pixel(int tile, int tiles_x, int t, int& x, int& y) {
  x = (tile % tiles_x) * 16 + ((t >> 6) & 1) * 8 + (t & 7);
  y = (tile / tiles_x) * 16 + (t >> 7) * 8 + ((t >> 3) & 7);
}

// The raycast options of the context's level `level` (its camera and size).
inline RayParams ray_params(const kf_context* ctx, int level) {
  const kf_config& c = ctx->cfg;
  const LevelBuffers& L = ctx->lv[level];
  RayParams rp{};
  rp.k = L.k;
  rp.w = L.w;
  rp.h = L.h;
  rp.min_depth = c.ray_min_depth;
  rp.max_depth = c.ray_max_depth;
  rp.step_scale = c.step_scale;
  rp.require_all = c.require_all_corners;
  rp.cell_normals = c.cell_normals;
  rp.seed_previous = c.seed_from_previous;
  return rp;
}

}  // namespace kf

// Trilinear TSDF sampler over the dense voxel array or the hashed block pool
// (reference: TrilinearSampler + DenseCornerAccess / SparseCornerAccess, include/kinectfusion/
// volume_sampler.hpp:36-255; blend_tsdf / blend_color / cell_gradient /
// central_normal, trilinear.hpp:85-269).  Shared by the raycaster and the
// mesher so both read the volume with the same float operations as the CPU
// sampler: the 2^3 stencil is visited z -> y -> x and each corner weight is
// ((wx * wy) * wz).
#pragma once

#include "kf_internal.hpp"

namespace kf {

// Corner access (DenseCornerAccess / SparseCornerAccess,
// volume_sampler.hpp:36-109 and block_rep.hpp:113-160).  A voxel of an
// unallocated block reads as a fresh voxel and a colour voxel as zero: the
// blends treat both exactly like the reference's missing (null) corners.
// The last block a hashed store looked up (key, pool slot).
struct SlotCache {
  uint32_t key = kNoSlot;
  uint32_t slot = kNoSlot;
};

// Two x-neighbouring voxels as one vector load (x-fastest layout: they are
// adjacent in memory; the pair is only as aligned as one voxel).
typedef uint32_t u32x2_a4 __attribute__((ext_vector_type(2), aligned(4)));
typedef float f32x4_a8 __attribute__((ext_vector_type(4), aligned(8)));

template <typename Raw>
__device__ inline void load_pair(const Raw* p, Raw& a, Raw& b);
template <>
__device__ inline void load_pair<uint32_t>(const uint32_t* p, uint32_t& a, uint32_t& b) {
  const u32x2_a4 v = *reinterpret_cast<const u32x2_a4*>(p);
  a = v.x;
  b = v.y;
}
template <>
__device__ inline void load_pair<float2>(const float2* p, float2& a, float2& b) {
  const f32x4_a8 v = *reinterpret_cast<const f32x4_a8*>(p);
  a = make_float2(v.x, v.y);
  b = make_float2(v.z, v.w);
}

template <int Kind>
struct DenseStore {
  using Raw = typename VoxelIO<Kind>::Raw;
  static constexpr bool kDirect = true;  // a voxel is one load
  const Raw* __restrict__ vox;
  const float4* __restrict__ colors;
  __device__ inline Raw voxel(const Grid& g, int x, int y, int z) const { return vox[g.flat(x, y, z)]; }
  // voxels (x, y, z) and (x + 1, y, z), both inside the grid
  __device__ inline void pair(const Grid& g, int x, int y, int z, Raw& a, Raw& b) const {
    load_pair<Raw>(vox + g.flat(x, y, z), a, b);
  }
  // The eight corners of the cell at (x, y, z) (z -> y -> x order), all
  // inside the grid: one 64-bit index, then four x-pair loads at fixed
  // row and slice offsets.
  __device__ inline void cell(const Grid& g, int x, int y, int z, Raw (&raw)[8]) const {
    const Raw* p = vox + g.flat(x, y, z);
    const size_t row = static_cast<size_t>(g.nx), slice = static_cast<size_t>(g.nx) * static_cast<size_t>(g.ny);
    load_pair<Raw>(p, raw[0], raw[1]);
    load_pair<Raw>(p + row, raw[2], raw[3]);
    load_pair<Raw>(p + slice, raw[4], raw[5]);
    load_pair<Raw>(p + slice + row, raw[6], raw[7]);
  }
  __device__ inline Raw voxel_cached(const Grid& g, int x, int y, int z, SlotCache&) const {
    return vox[g.flat(x, y, z)];
  }
  __device__ inline float4 color(const Grid& g, int x, int y, int z) const { return colors[g.flat(x, y, z)]; }
  __device__ inline float4 color_cached(const Grid& g, int x, int y, int z, SlotCache&) const {
    return colors[g.flat(x, y, z)];
  }
  __device__ inline bool has_colors() const { return colors != nullptr; }
};

template <int Kind>
struct HashedStore {
  using Raw = typename VoxelIO<Kind>::Raw;
  static constexpr bool kDirect = false;  // a voxel needs a hash probe first
  HashView hv;
  const Raw* __restrict__ pool;
  const float4* __restrict__ cpool;
  __device__ inline uint32_t slot_of(int x, int y, int z, SlotCache& cache) const {
    const uint32_t key = hv.key_of(x, y, z);
    if (key != cache.key) {
      cache.key = key;
      cache.slot = hv.find(key);
    }
    return cache.slot;
  }
  __device__ inline Raw voxel_cached(const Grid&, int x, int y, int z, SlotCache& cache) const {
    const uint32_t slot = slot_of(x, y, z, cache);
    return slot == kNoSlot ? VoxelIO<Kind>::fresh() : pool[static_cast<size_t>(slot) * kBlockVoxels + intra_of(x, y, z)];
  }
  __device__ inline Raw voxel(const Grid& g, int x, int y, int z) const {
    SlotCache cache;
    return voxel_cached(g, x, y, z, cache);
  }
  __device__ inline float4 color_cached(const Grid&, int x, int y, int z, SlotCache& cache) const {
    const uint32_t slot = slot_of(x, y, z, cache);
    return slot == kNoSlot ? make_float4(0.0F, 0.0F, 0.0F, 0.0F)
                           : cpool[static_cast<size_t>(slot) * kBlockVoxels + intra_of(x, y, z)];
  }
  __device__ inline float4 color(const Grid& g, int x, int y, int z) const {
    SlotCache cache;
    return color_cached(g, x, y, z, cache);
  }
  __device__ inline bool has_colors() const { return cpool != nullptr; }
  // The pool slots of the eight corners (x fastest) of the 2^3 stencil with
  // base corner (x, y, z); bit c of `in` marks a corner inside the grid.
  // The first probe of every corner outside the cached block is issued
  // before any is examined, so a stencil across block faces costs one round
  // of table reads, not one per block; a probe that meets another key walks
  // the chain.  The cache then holds the block of corner `own`.
  __device__ inline void slots8(int x, int y, int z, unsigned in, int own, uint32_t (&slot)[8],
                                SlotCache& cache) const {
    uint32_t key[8];
    uint2 e[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      key[c] = hv.key_of(x + (c & 1), y + ((c >> 1) & 1), z + (c >> 2));
      e[c] = make_uint2(0xFFFFFFFFU, 0xFFFFFFFFU);  // kHashEmpty
      if (((in >> c) & 1U) != 0U && key[c] != cache.key)
        e[c] = reinterpret_cast<const uint2*>(hv.table)[hv.bucket(key[c])];
    }
    unsigned walk = 0U;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      uint32_t sl = kNoSlot;
      if (((in >> c) & 1U) != 0U) {
        if (key[c] == cache.key) {
          sl = cache.slot;
        } else if (e[c].y == key[c]) {
          sl = e[c].x == kPendingSlot ? kNoSlot : e[c].x;
        } else if (e[c].x != 0xFFFFFFFFU || e[c].y != 0xFFFFFFFFU) {
          walk |= 1U << c;  // the home bucket holds another key
        }
      }
      slot[c] = sl;
    }
    if (walk != 0U) {
#pragma unroll
      for (int c = 0; c < 8; ++c)
        if (((walk >> c) & 1U) != 0U) slot[c] = hv.find(key[c]);
    }
    if (((in >> own) & 1U) != 0U) {
      uint32_t k = 0U, sl = 0U;
#pragma unroll
      for (int c = 0; c < 8; ++c)
        if (c == own) {
          k = key[c];
          sl = slot[c];
        }
      cache.key = k;
      cache.slot = sl;
    }
  }
};

template <int Kind, typename Store = DenseStore<Kind>>
struct Sampler {
  using IO = VoxelIO<Kind>;
  Store s;
  Grid g;
  float inv_voxel;

  __device__ inline bool contains(V3 p) const {
    return p.x >= 0.0F && p.y >= 0.0F && p.z >= 0.0F && p.x < static_cast<float>(g.nx) &&
           p.y < static_cast<float>(g.ny) && p.z < static_cast<float>(g.nz);
  }
  __device__ inline V3 grid_coords(V3 p) const {
    return scale3(sub3(p, mk3(g.ox, g.oy, g.oz)), inv_voxel);
  }
  __device__ static inline float axis_w(float f, int o) { return o == 0 ? 1.0F - f : f; }

  // The 2^3 stencil of a point: base corner, fractions, in-volume flag.
  struct Stencil {
    int bx, by, bz;
    V3 fr;
    bool ok;        // the point lies inside the grid
    bool interior;  // all eight corners lie inside the grid
  };
  __device__ inline Stencil stencil(V3 p) const { return stencil_gc(grid_coords(p)); }
  // the same from grid coordinates already computed
  __device__ inline Stencil stencil_gc(V3 gc) const {
    Stencil st;
    st.ok = contains(gc);
    const V3 gg = sub3(gc, mk3(0.5F, 0.5F, 0.5F));
    const V3 fl = mk3(floorf(gg.x), floorf(gg.y), floorf(gg.z));
    st.bx = static_cast<int>(fl.x);
    st.by = static_cast<int>(fl.y);
    st.bz = static_cast<int>(fl.z);
    st.fr = sub3(gg, fl);
    st.interior = st.bx >= 0 && st.by >= 0 && st.bz >= 0 && st.bx + 1 < g.nx && st.by + 1 < g.ny &&
                  st.bz + 1 < g.nz;
    return st;
  }
  // Loads of the eight corners (z -> y -> x order); bit c of `inb` marks a
  // corner inside the grid.
  __device__ inline void load8(const Stencil& st, typename IO::Raw (&raw)[8], unsigned& inb, SlotCache& cache) const {
    if constexpr (Store::kDirect) {
      if (st.ok && st.interior) {  // four x-pairs
        s.cell(g, st.bx, st.by, st.bz, raw);
        inb = 0xFFU;
        return;
      }
    } else {
      // the eight slots first (their probes together), then the voxels
      inb = 0U;
      if (st.ok) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const int x = st.bx + (c & 1), y = st.by + ((c >> 1) & 1), z = st.bz + (c >> 2);
          if (st.interior || (x >= 0 && y >= 0 && z >= 0 && x < g.nx && y < g.ny && z < g.nz)) inb |= 1U << c;
        }
      }
      // the corner holding the stencil's point: base + 1 where the fraction is >= 0.5
      const int own = (st.fr.x >= 0.5F ? 1 : 0) | (st.fr.y >= 0.5F ? 2 : 0) | (st.fr.z >= 0.5F ? 4 : 0);
      uint32_t slot[8];
      s.slots8(st.bx, st.by, st.bz, inb, own, slot, cache);
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int x = st.bx + (c & 1), y = st.by + ((c >> 1) & 1), z = st.bz + (c >> 2);
        raw[c] = slot[c] == kNoSlot ? IO::fresh()
                                    : s.pool[static_cast<size_t>(slot[c]) * kBlockVoxels + intra_of(x, y, z)];
      }
      return;
    }
    inb = 0U;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int x = st.bx + (c & 1), y = st.by + ((c >> 1) & 1), z = st.bz + (c >> 2);
      raw[c] = IO::fresh();
      if (!st.ok) continue;
      if (!st.interior && (x < 0 || y < 0 || z < 0 || x >= g.nx || y >= g.ny || z >= g.nz)) continue;
      raw[c] = s.voxel_cached(g, x, y, z, cache);
      inb |= 1U << c;
    }
  }
  // blend_tsdf over loaded corners (the same float operations as tsdf()).
  __device__ inline bool blend(const Stencil& st, const typename IO::Raw (&raw)[8], unsigned inb, bool require_all,
                               float& out) const {
    if (!st.ok) return false;
    float acc = 0.0F, wsum = 0.0F;
    int valid = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      if (((inb >> c) & 1U) == 0U) continue;
      float t, w;
      bool obs, fin;
      IO::unpack(raw[c], t, w, obs, fin);
      if (!obs || !fin) continue;
      const float cw = axis_w(st.fr.x, c & 1) * axis_w(st.fr.y, (c >> 1) & 1) * axis_w(st.fr.z, c >> 2);
      acc += cw * t;
      wsum += cw;
      ++valid;
    }
    if (require_all && valid < 8) return false;
    if (wsum <= 1.0e-6F) return false;
    const float d = acc / wsum;
    if (!isfinite(d)) return false;
    out = d;
    return true;
  }

  // blend() of a cell whose eight corners all lie inside the grid, without
  // branches: a corner that is unobserved or not finite adds +0 to both sums,
  // which leaves them unchanged (neither sum is ever -0: they start at +0 and
  // a round-to-nearest sum is -0 only when both terms are).
  __device__ inline bool blend_all(const Stencil& st, const typename IO::Raw (&raw)[8], bool require_all,
                                   float& out) const {
    float acc = 0.0F, wsum = 0.0F;
    int valid = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      float t, w;
      bool obs, fin;
      IO::unpack(raw[c], t, w, obs, fin);
      const bool use = obs && fin;
      const float cw = axis_w(st.fr.x, c & 1) * axis_w(st.fr.y, (c >> 1) & 1) * axis_w(st.fr.z, c >> 2);
      acc += use ? cw * t : 0.0F;
      wsum += use ? cw : 0.0F;
      valid += use ? 1 : 0;
    }
    if (require_all && valid < 8) return false;
    if (wsum <= 1.0e-6F) return false;
    const float d = acc / wsum;
    if (!isfinite(d)) return false;
    out = d;
    return true;
  }

  // TrilinearSampler::tsdf + blend_tsdf.
  __device__ bool tsdf(V3 p, bool require_all, float& out) const {
    SlotCache cache;
    return tsdf(p, require_all, out, cache);
  }
  // The same with a block-slot cache that the caller keeps across samples: a
  // hashed store probes a block once while consecutive samples stay in it
  // (the table does not change during a launch).
  __device__ bool tsdf(V3 p, bool require_all, float& out, SlotCache& cache) const {
    const Stencil st = stencil(p);
    if (!st.ok) return false;
    typename IO::Raw raw[8];
    if constexpr (Store::kDirect) {
      // the eight corners as four x-pair loads; blend_all() and blend() are
      // the same float sequence
      if (st.interior) {
        s.cell(g, st.bx, st.by, st.bz, raw);
        return blend_all(st, raw, require_all, out);
      }
    }
    unsigned inb;
    load8(st, raw, inb, cache);
    return blend(st, raw, inb, require_all, out);
  }

  // blend_color.
  __device__ bool color(V3 p, V3& out) const {
    if (!s.has_colors()) return false;
    const V3 gc = grid_coords(p);
    if (!contains(gc)) return false;
    const V3 gg = sub3(gc, mk3(0.5F, 0.5F, 0.5F));
    const V3 fl = mk3(floorf(gg.x), floorf(gg.y), floorf(gg.z));
    const int bx = static_cast<int>(fl.x), by = static_cast<int>(fl.y), bz = static_cast<int>(fl.z);
    const V3 fr = sub3(gg, fl);
    V3 acc = mk3(0.0F, 0.0F, 0.0F);
    float wsum = 0.0F;
    SlotCache cache;
#pragma unroll
    for (int oz = 0; oz <= 1; ++oz)
#pragma unroll
      for (int oy = 0; oy <= 1; ++oy)
#pragma unroll
        for (int ox = 0; ox <= 1; ++ox) {
          const int x = bx + ox, y = by + oy, z = bz + oz;
          if (x < 0 || y < 0 || z < 0 || x >= g.nx || y >= g.ny || z >= g.nz) continue;
          const float4 c = s.color_cached(g, x, y, z, cache);
          if (c.w <= 0.0F) continue;
          const float cw = axis_w(fr.x, ox) * axis_w(fr.y, oy) * axis_w(fr.z, oz);
          acc = add3(acc, scale3(mk3(c.x, c.y, c.z), cw));
          wsum += cw;
        }
    if (wsum <= 1.0e-6F) return false;
    out = div3(acc, wsum);
    return true;
  }

  // central_normal (trilinear.hpp:246-269).  The corner loads of the two
  // samples of an axis are issued together; the blends run in the
  // reference's order.  Two samples at a time (not all six) bound the
  // register use of the raycaster, where the normal's peak would otherwise
  // set the register allocation, and so the occupancy, of the whole march.
  __device__ bool normal(V3 p, bool require_all, V3& out) const {
    // the six samples kBatch at a time: a direct store reads two at once; a
    // hashed store's slot probes take too many registers for two
    constexpr int kBatch = Store::kDirect ? 2 : 1;
    float v[6];
#pragma unroll
    for (int a0 = 0; a0 < 6; a0 += kBatch) {
      Stencil st[kBatch];
      typename IO::Raw raw[kBatch][8];
      unsigned inb[kBatch];
      SlotCache cache;
#pragma unroll
      for (int i = 0; i < kBatch; ++i) {
        const int axis = (a0 + i) >> 1;
        const V3 off = scale3(mk3(axis == 0 ? 1.0F : 0.0F, axis == 1 ? 1.0F : 0.0F, axis == 2 ? 1.0F : 0.0F), g.voxel);
        st[i] = stencil(((a0 + i) & 1) == 0 ? add3(p, off) : sub3(p, off));
        load8(st[i], raw[i], inb[i], cache);
      }
#pragma unroll
      for (int i = 0; i < kBatch; ++i)
        if (!blend(st[i], raw[i], inb[i], require_all, v[a0 + i])) return false;
    }
    const V3 dir = mk3(v[0] - v[1], v[2] - v[3], v[4] - v[5]);
    const float len = norm3(dir);
    if (len <= 0.0F) return false;
    out = div3(dir, len);
    return true;
  }

  // cell_normal + cell_gradient (volume_sampler.hpp:189-239).
  __device__ bool cell_normal(V3 p, V3& out) const {
    const V3 gc = grid_coords(p);
    if (!contains(gc)) return false;
    const V3 gg = sub3(gc, mk3(0.5F, 0.5F, 0.5F));
    const V3 fl = mk3(floorf(gg.x), floorf(gg.y), floorf(gg.z));
    const int bx = static_cast<int>(fl.x), by = static_cast<int>(fl.y), bz = static_cast<int>(fl.z);
    const V3 f = sub3(gg, fl);
    float c[2][2][2];
    bool complete = true;
    SlotCache cache;
#pragma unroll
    for (int oz = 0; oz <= 1; ++oz)
#pragma unroll
      for (int oy = 0; oy <= 1; ++oy)
#pragma unroll
        for (int ox = 0; ox <= 1; ++ox) {
          const int x = bx + ox, y = by + oy, z = bz + oz;
          c[oz][oy][ox] = 0.0F;
          if (x < 0 || y < 0 || z < 0 || x >= g.nx || y >= g.ny || z >= g.nz) {
            complete = false;
            continue;
          }
          float t, w;
          bool obs, fin;
          IO::unpack(s.voxel_cached(g, x, y, z, cache), t, w, obs, fin);
          if (!obs || !fin) complete = false;
          c[oz][oy][ox] = t;
        }
    if (!complete) return normal(p, false, out);
    auto lerp2 = [](float lo, float hi, float m) { return lo + (m * (hi - lo)); };
    const float gx = lerp2(lerp2(c[0][0][1] - c[0][0][0], c[0][1][1] - c[0][1][0], f.y),
                           lerp2(c[1][0][1] - c[1][0][0], c[1][1][1] - c[1][1][0], f.y), f.z);
    const float gy = lerp2(lerp2(c[0][1][0] - c[0][0][0], c[0][1][1] - c[0][0][1], f.x),
                           lerp2(c[1][1][0] - c[1][0][0], c[1][1][1] - c[1][0][1], f.x), f.z);
    const float gz = lerp2(lerp2(c[1][0][0] - c[0][0][0], c[1][0][1] - c[0][0][1], f.x),
                           lerp2(c[1][1][0] - c[0][1][0], c[1][1][1] - c[0][1][1], f.x), f.y);
    const V3 grad = mk3(gx, gy, gz);
    const float len = norm3(grad);
    if (len <= 0.0F) return false;
    out = div3(grad, len);
    return true;
  }
};

}  // namespace kf

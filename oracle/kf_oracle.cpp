// KinectFusion CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// A C++20 restatement of the reference's single-threaded CPU path
// (maximsolodukhin/KinectFusion, "BasicHostPipeline"), used as the checker for
// the HIP library.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it; the product never links or calls it.
//
// Every function follows the reference host code operation by operation
// (same float expression order, IEEE division and sqrt, no contraction: build
// with -ffp-contract=off), citing the reference file:line it restates.
//
// Deliberate departures, each also made by the HIP library:
//  * ICP normal equations are accumulated in FP64 (the reference sums in
//    FP32); the 6x6 system is solved by an FP64 Cholesky factorisation and the
//    pose composed in FP64 (reference: float Cholesky in
//    src/kinectfusion/cuda/icp_correspondence.cu:169-209 for its device loop).
//  * The stability veto uses a cyclic Jacobi eigen-solver in FP64 instead of
//    Eigen::SelfAdjointEigenSolver<float> (icp_optimizer.cpp:319-347).
//  * Poses are FP64 4x4 matrices; the float transforms handed to the
//    per-pixel code are the FP64 values rounded to float, and the rigid
//    inverse is formed in FP64 (reference: Eigen Matrix4f::inverse()).
//
// Parity of the Eigen-bit-level parts (float eigenvalues, LDLT solution bits,
// AngleAxisf, Matrix4f::inverse) is unpinned by construction: those are
// replaced by the FP64 forms above.

#include <algorithm>
#include <array>
#include <bit>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <unordered_map>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Core math: include/kinectfusion/vector.hpp (host branches only).
// ---------------------------------------------------------------------------
struct V3 {
  float x{}, y{}, z{};
};

inline V3 mk(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return mk(-a.x, -a.y, -a.z); }
inline V3 operator*(V3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(float s, V3 a) { return a * s; }  // vector.hpp:125-128
inline V3 operator/(V3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
// vector.hpp:143-146: ((x*x) + (y*y)) + (z*z)
inline float dot(V3 a, V3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
inline V3 cross(V3 a, V3 b) {  // vector.hpp:148-153
  return mk((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z),
            (a.x * b.y) - (a.y * b.x));
}
inline float norm(V3 v) { return std::sqrt(dot(v, v)); }  // vector.hpp:160-162
// Host branch of vector.hpp:164-174: divide by the length.
inline V3 normalized(V3 v) {
  const float length = norm(v);
  return length > 0.0F ? v / length : V3{};
}
inline bool all_finite(V3 v) {
  return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z);
}
inline V3 nan3() {
  const float n = std::numeric_limits<float>::quiet_NaN();
  return mk(n, n, n);
}
// cuda_compat.hpp:212-235
template <typename T>
inline T cmin(T a, T b) { return b < a ? b : a; }
template <typename T>
inline T cmax(T a, T b) { return a < b ? b : a; }
template <typename T>
inline T cclamp(T v, T lo, T hi) {
  if (v < lo) return lo;
  if (hi < v) return hi;
  return v;
}

// Row-major 3x3 rotation plus translation (vector.hpp:189-229).
struct Rigid {
  V3 rx, ry, rz, t;
};
inline V3 rot(const Rigid& m, V3 p) { return mk(dot(m.rx, p), dot(m.ry, p), dot(m.rz, p)); }
inline V3 apply(const Rigid& m, V3 p) { return rot(m, p) + m.t; }  // vector.hpp:226-229

// A 3x4 row-major float matrix as passed across the ABI.
inline Rigid rigid_from(const float* m) {
  return Rigid{mk(m[0], m[1], m[2]), mk(m[4], m[5], m[6]), mk(m[8], m[9], m[10]),
               mk(m[3], m[7], m[11])};
}

struct Intr {
  float fx, fy, cx, cy;
};

// rgbd.hpp:40-44
inline void project(const Intr& k, V3 p, float& u, float& v) {
  u = (k.fx * p.x / p.z) + k.cx;
  v = (k.fy * p.y / p.z) + k.cy;
}
// rgbd.hpp:53-57
inline V3 back_project(const Intr& k, float px, float py, float depth) {
  return mk((px - k.cx) * depth / k.fx, (py - k.cy) * depth / k.fy, depth);
}

// lround(v) landing in [0, limit): the value checks of
// tsdf_integration.hpp:156-168 and icp_correspondence.hpp:261-272.
inline bool round_index(float v, long limit, long& out) {
  if (!std::isfinite(v)) return false;
  if (v > 1.0e9F || v < -1.0e9F) return false;
  const long r = std::lround(v);
  if (r < 0 || r >= limit) return false;
  out = r;
  return true;
}

// rgbd.hpp:122-133
inline bool depth_in_range(std::uint16_t raw, float scale, float min_d, float max_d,
                           float& meters) {
  if (raw == 0) return false;
  const float m = static_cast<float>(raw) / scale;
  if (m < min_d || m > max_d) return false;
  meters = m;
  return true;
}

// ---------------------------------------------------------------------------
// Voxel stores: include/kinectfusion/volume.hpp:28-184.
// ---------------------------------------------------------------------------
enum VoxelKind : int { kFloatVoxel = 0, kQuantizedVoxel = 1, kBf16Voxel = 2 };

constexpr float kQScale = 32767.0F;              // volume.hpp:29
constexpr float kQInvScale = 1.0F / kQScale;     // volume.hpp:32
constexpr float kWScale = 256.0F;                // volume.hpp:86
constexpr float kWInvScale = 1.0F / kWScale;     // volume.hpp:87
constexpr float kWMax = 65535.0F / kWScale;      // volume.hpp:88

inline std::int16_t q_encode(float t) {  // volume.hpp:37-39
  return static_cast<std::int16_t>(cclamp(t, -1.0F, 1.0F) * kQScale);
}
inline float q_decode(std::int16_t r) { return static_cast<float>(r) * kQInvScale; }
inline std::uint16_t w_encode(float w) {  // volume.hpp:93-95
  return static_cast<std::uint16_t>(cclamp(w, 0.0F, kWMax) * kWScale);
}
inline float w_decode(std::uint16_t r) { return static_cast<float>(r) * kWInvScale; }
inline std::uint16_t bf_encode(float t) {  // volume.hpp:60-62,74-78
  const std::uint32_t bits = std::bit_cast<std::uint32_t>(cclamp(t, -1.0F, 1.0F));
  return static_cast<std::uint16_t>((bits + 0x7FFFU + ((bits >> 16U) & 1U)) >> 16U);
}
inline float bf_decode(std::uint16_t r) {  // volume.hpp:64-68
  return std::bit_cast<float>(static_cast<std::uint32_t>(r) << 16U);
}

struct FloatVoxel {
  float tsdf, weight;
};
struct PackedVoxel {
  std::uint16_t dist, weight;  // int16 or bf16 distance bits, u16 weight
};

// Float view of any stored voxel plus the observed/finite predicates of
// volume.hpp:110-134.
struct Loaded {
  float tsdf;
  float weight;
  bool observed;
  bool finite;
};

struct Volume {
  std::int64_t nx, ny, nz;
  float voxel_size;
  V3 origin;
  float truncation;
  int kind;
  void* voxels;
  float* colors;  // 4 floats per voxel (r, g, b, weight) or null
  // Sparse store (BlockRep, block_rep.hpp:183-307): the pool slot of every
  // 8^3 block (-1 unallocated, grid.hpp:44-59) and voxels/colors are the
  // pool, 512 voxels per slot in intra-block order (grid.hpp:61-69); null
  // for the dense store.
  const std::int32_t* slots = nullptr;

  std::size_t flat(std::int64_t x, std::int64_t y, std::int64_t z) const {  // vector.hpp:62-68
    return static_cast<std::size_t>(((z * ny) + y) * nx + x);
  }
  // Storage index of voxel (x, y, z), or -1 when its block is unallocated
  // (SparseVolumeView::pool_at returns null, block_rep.hpp:87-100).
  std::int64_t index(std::int64_t x, std::int64_t y, std::int64_t z) const {
    if (slots == nullptr) return static_cast<std::int64_t>(flat(x, y, z));
    const std::int64_t bx = (nx + 7) / 8, by = (ny + 7) / 8;
    const std::int32_t slot = slots[((z / 8) * by + (y / 8)) * bx + (x / 8)];
    if (slot < 0) return -1;
    return static_cast<std::int64_t>(slot) * 512 + (((z % 8) * 8 + (y % 8)) * 8 + (x % 8));
  }
  V3 cell_center(std::int64_t x, std::int64_t y, std::int64_t z) const {  // volume.hpp:276-283
    const V3 start = mk(static_cast<float>(x), static_cast<float>(y), static_cast<float>(z));
    return origin + ((start + mk(0.5F, 0.5F, 0.5F)) * voxel_size);
  }
  bool contains(V3 p) const {  // vector.hpp:42-47
    return p.x >= 0.0F && p.y >= 0.0F && p.z >= 0.0F && p.x < static_cast<float>(nx) &&
           p.y < static_cast<float>(ny) && p.z < static_cast<float>(nz);
  }
  bool contains_i(long long x, long long y, long long z) const {  // vector.hpp:51-57
    return x >= 0 && y >= 0 && z >= 0 && x < nx && y < ny && z < nz;
  }
  Loaded load(std::size_t i) const {
    if (kind == kFloatVoxel) {
      const FloatVoxel v = static_cast<const FloatVoxel*>(voxels)[i];
      return Loaded{v.tsdf, v.weight, v.weight > 0.0F, std::isfinite(v.tsdf)};
    }
    const PackedVoxel v = static_cast<const PackedVoxel*>(voxels)[i];
    const float t = kind == kQuantizedVoxel ? q_decode(static_cast<std::int16_t>(v.dist))
                                            : bf_decode(v.dist);
    return Loaded{t, w_decode(v.weight), v.weight != 0, true};
  }
  // BasicVoxel::fused (volume.hpp:137-146)
  void fuse(std::size_t i, float observed, float w, float max_weight) {
    const Loaded cur = load(i);
    const float avg = ((cur.tsdf * cur.weight) + (observed * w)) / (cur.weight + w);
    const float tw = cmin(cur.weight + w, max_weight);
    if (kind == kFloatVoxel) {
      static_cast<FloatVoxel*>(voxels)[i] = FloatVoxel{avg, tw};
    } else {
      const std::uint16_t d = kind == kQuantizedVoxel ? static_cast<std::uint16_t>(q_encode(avg))
                                                      : bf_encode(avg);
      static_cast<PackedVoxel*>(voxels)[i] = PackedVoxel{d, w_encode(tw)};
    }
  }
};

// ---------------------------------------------------------------------------
// Trilinear sampler: trilinear.hpp:85-269, volume_sampler.hpp:36-255.
// ---------------------------------------------------------------------------
struct Sampler {
  const Volume* vol;
  float inv_voxel;  // volume_sampler.hpp:127-128

  explicit Sampler(const Volume* v) : vol(v), inv_voxel(1.0F / v->voxel_size) {}

  V3 grid_coords(V3 p) const { return (p - vol->origin) * inv_voxel; }  // :247-252

  struct GS {
    int bx, by, bz;
    V3 frac;
  };
  static GS grid_sample(V3 g) {  // trilinear.hpp:85-91
    const V3 gg = g - mk(0.5F, 0.5F, 0.5F);
    const V3 fl = mk(std::floor(gg.x), std::floor(gg.y), std::floor(gg.z));
    return GS{static_cast<int>(fl.x), static_cast<int>(fl.y), static_cast<int>(fl.z), gg - fl};
  }
  static float axis_w(float f, int o) { return o == 0 ? 1.0F - f : f; }  // :130-133
  static float corner_w(V3 f, int ox, int oy, int oz) {                    // :136-140
    return axis_w(f.x, ox) * axis_w(f.y, oy) * axis_w(f.z, oz);
  }
  // Index of stencil corner, or -1 when clipped (volume_sampler.hpp:49-56;
  // both the interior and clipped paths address the same voxel).
  long long corner(const GS& s, int ox, int oy, int oz) const {
    const long long x = static_cast<long long>(s.bx) + ox;
    const long long y = static_cast<long long>(s.by) + oy;
    const long long z = static_cast<long long>(s.bz) + oz;
    if (!vol->contains_i(x, y, z)) return -1;
    return static_cast<long long>(vol->index(x, y, z));
  }

  // TrilinearSampler::tsdf + blend_tsdf (volume_sampler.hpp:141-158,
  // trilinear.hpp:144-181).
  bool tsdf(V3 p, bool require_all, float& out) const {
    const V3 g = grid_coords(p);
    if (!vol->contains(g)) return false;
    const GS s = grid_sample(g);
    float acc = 0.0F, wsum = 0.0F;
    int valid = 0;
    for (int oz = 0; oz <= 1; ++oz)
      for (int oy = 0; oy <= 1; ++oy)
        for (int ox = 0; ox <= 1; ++ox) {
          const long long i = corner(s, ox, oy, oz);
          if (i < 0) continue;
          const Loaded c = vol->load(static_cast<std::size_t>(i));
          if (!c.observed || !c.finite) continue;
          const float w = corner_w(s.frac, ox, oy, oz);
          acc += w * c.tsdf;
          wsum += w;
          ++valid;
        }
    if (require_all && valid < 8) return false;
    if (wsum <= 1.0e-6F) return false;
    const float d = acc / wsum;
    if (!std::isfinite(d)) return false;
    out = d;
    return true;
  }

  // blend_color (trilinear.hpp:184-210)
  bool color(V3 p, V3& out) const {
    if (vol->colors == nullptr) return false;
    const V3 g = grid_coords(p);
    if (!vol->contains(g)) return false;
    const GS s = grid_sample(g);
    V3 acc{};
    float wsum = 0.0F;
    for (int oz = 0; oz <= 1; ++oz)
      for (int oy = 0; oy <= 1; ++oy)
        for (int ox = 0; ox <= 1; ++ox) {
          const long long i = corner(s, ox, oy, oz);
          if (i < 0) continue;
          const float* c = vol->colors + 4 * static_cast<std::size_t>(i);
          if (c[3] <= 0.0F) continue;
          const float w = corner_w(s.frac, ox, oy, oz);
          acc = acc + (w * mk(c[0], c[1], c[2]));
          wsum += w;
        }
    if (wsum <= 1.0e-6F) return false;
    out = acc / wsum;
    return true;
  }

  // central_normal (trilinear.hpp:246-269)
  bool normal(V3 p, bool require_all, V3& out) const {
    float g[3];
    for (int axis = 0; axis < 3; ++axis) {
      const V3 off = mk(axis == 0 ? 1.0F : 0.0F, axis == 1 ? 1.0F : 0.0F,
                        axis == 2 ? 1.0F : 0.0F) * vol->voxel_size;
      float plus = 0.0F, minus = 0.0F;
      if (!tsdf(p + off, require_all, plus)) return false;
      if (!tsdf(p - off, require_all, minus)) return false;
      g[axis] = plus - minus;
    }
    const V3 dir = mk(g[0], g[1], g[2]);
    const float len = norm(dir);
    if (len <= 0.0F) return false;
    out = dir / len;
    return true;
  }

  // cell_normal + cell_gradient (volume_sampler.hpp:189-239,
  // trilinear.hpp:214-243)
  bool cell_normal(V3 p, V3& out) const {
    const V3 g = grid_coords(p);
    if (!vol->contains(g)) return false;
    const GS s = grid_sample(g);
    float c[2][2][2];
    bool complete = true;
    for (int oz = 0; oz <= 1 && complete; ++oz)
      for (int oy = 0; oy <= 1 && complete; ++oy)
        for (int ox = 0; ox <= 1; ++ox) {
          const long long i = corner(s, ox, oy, oz);
          if (i < 0) { complete = false; break; }
          const Loaded v = vol->load(static_cast<std::size_t>(i));
          if (!v.observed || !v.finite) { complete = false; break; }
          c[oz][oy][ox] = v.tsdf;
        }
    if (!complete) return normal(p, false, out);
    const V3 f = s.frac;
    auto lerp2 = [](float lo, float hi, float m) { return lo + (m * (hi - lo)); };
    const float gx = lerp2(lerp2(c[0][0][1] - c[0][0][0], c[0][1][1] - c[0][1][0], f.y),
                           lerp2(c[1][0][1] - c[1][0][0], c[1][1][1] - c[1][1][0], f.y), f.z);
    const float gy = lerp2(lerp2(c[0][1][0] - c[0][0][0], c[0][1][1] - c[0][0][1], f.x),
                           lerp2(c[1][1][0] - c[1][0][0], c[1][1][1] - c[1][0][1], f.x), f.z);
    const float gz = lerp2(lerp2(c[1][0][0] - c[0][0][0], c[1][0][1] - c[0][0][1], f.x),
                           lerp2(c[1][1][0] - c[0][1][0], c[1][1][1] - c[0][1][1], f.x), f.y);
    const V3 grad = mk(gx, gy, gz);
    const float len = norm(grad);
    if (len <= 0.0F) return false;
    out = grad / len;
    return true;
  }
};

// rgbd.hpp:84-93
inline std::uint32_t pixel_from_color(V3 c) {
  auto to_byte = [](float v) { return static_cast<std::uint32_t>(cclamp(v, 0.0F, 255.0F)); };
  return to_byte(c.x) | (to_byte(c.y) << 8U) | (to_byte(c.z) << 16U) | (0xFFU << 24U);
}
// rgbd.hpp:97-102
inline V3 color_from_pixel(std::uint32_t p) {
  return mk(static_cast<float>(p & 0xFFU), static_cast<float>((p >> 8U) & 0xFFU),
            static_cast<float>((p >> 16U) & 0xFFU));
}

// ---------------------------------------------------------------------------
// FP64 pose helpers (replace Eigen float inverse / AngleAxis; see header).
// ---------------------------------------------------------------------------
struct Pose64 {
  double m[16];  // row-major 4x4, last row 0 0 0 1
};

inline void rigid_inverse64(const double* a, double* out) {
  // out = [R^T, -R^T t]
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) out[r * 4 + c] = a[c * 4 + r];
  for (int r = 0; r < 3; ++r) {
    out[r * 4 + 3] = -(out[r * 4 + 0] * a[3] + out[r * 4 + 1] * a[7] + out[r * 4 + 2] * a[11]);
  }
  out[12] = 0.0;
  out[13] = 0.0;
  out[14] = 0.0;
  out[15] = 1.0;
}

inline Rigid rigid_from64(const double* m) {
  auto f = [&](int i) { return static_cast<float>(m[i]); };
  return Rigid{mk(f(0), f(1), f(2)), mk(f(4), f(5), f(6)), mk(f(8), f(9), f(10)),
               mk(f(3), f(7), f(11))};
}

}  // namespace

// ===========================================================================
// C ABI for the parity tests (ctypes).
// ===========================================================================
extern "C" {

struct kfo_geometry {
  std::int64_t nx, ny, nz;
  float voxel_size;
  float ox, oy, oz;
  float truncation;
};

struct kfo_intrinsics {
  float fx, fy, cx, cy;
};

struct kfo_integration_options {
  float depth_scale;
  float observation_weight;
  float max_weight;
  float min_depth;
  float max_depth;
  int projective_distance;
  int distance_scaled_truncation;
  float truncation_distance_scale;
};

static Volume make_volume(const kfo_geometry* g, int kind, const void* voxels, const float* colors,
                          const std::int32_t* slots = nullptr) {
  return Volume{g->nx, g->ny, g->nz, g->voxel_size, mk(g->ox, g->oy, g->oz), g->truncation,
                kind, const_cast<void*>(voxels), const_cast<float*>(colors), slots};
}

// Fresh voxels: tsdf = kUnobservedTsdf, weight 0 (volume.hpp:23,107-108);
// colour voxels zero (volume.hpp:169-171).  n voxels (a dense grid, or a
// sparse pool's slots x 512).
void kfo_voxels_reset(int kind, void* voxels, float* colors, std::int64_t count) {
  const std::size_t n = static_cast<std::size_t>(count);
  if (kind == kFloatVoxel) {
    auto* v = static_cast<FloatVoxel*>(voxels);
    for (std::size_t i = 0; i < n; ++i) v[i] = FloatVoxel{1.0F, 0.0F};
  } else {
    const std::uint16_t d = kind == kQuantizedVoxel ? static_cast<std::uint16_t>(q_encode(1.0F))
                                                    : bf_encode(1.0F);
    auto* v = static_cast<PackedVoxel*>(voxels);
    for (std::size_t i = 0; i < n; ++i) v[i] = PackedVoxel{d, 0};
  }
  if (colors != nullptr) std::memset(colors, 0, n * 4 * sizeof(float));
}

void kfo_volume_reset(const kfo_geometry* g, int kind, void* voxels, float* colors) {
  kfo_voxels_reset(kind, voxels, colors, g->nx * g->ny * g->nz);
}

// IntegrationContext::observe/fuse (tsdf_integration.hpp:183-252) with the
// Classic/AngleWeighted rules (:287-342) for one voxel.  Returns 0 when the
// voxel is not fused, 1 when fused, 2 when fused with colour.
static int integrate_voxel(Volume& vol, std::int64_t x, std::int64_t y, std::int64_t z, int rule,
                           const Rigid& w2c, const Intr& intr, const std::uint16_t* depth, int width,
                           int height, const std::uint32_t* color, const float* normals,
                           const kfo_integration_options* opt) {
  // observe (tsdf_integration.hpp:184-224)
  const V3 cp = apply(w2c, vol.cell_center(x, y, z));
  if (cp.z <= 0.0F) return 0;  // :150-152
  float u, v;
  project(intr, cp, u, v);
  long col, row;
  if (!round_index(u, width, col) || !round_index(v, height, row)) return 0;
  float d;
  if (!depth_in_range(depth[row * width + col], opt->depth_scale, opt->min_depth, opt->max_depth, d)) return 0;
  const V3 ray = back_project(intr, static_cast<float>(col), static_cast<float>(row), 1.0F);
  const float lambda = norm(ray);
  if (!all_finite(ray) || lambda == 0.0F) return 0;
  const float trunc = opt->distance_scaled_truncation ? vol.truncation + (opt->truncation_distance_scale * d)
                                                      : vol.truncation;
  const float sd = opt->projective_distance ? d - (norm(cp) / lambda) : d - cp.z;
  if (sd < -trunc) return 0;
  const float tsdf = cclamp(sd / trunc, -1.0F, 1.0F);
  const bool integrate_color = sd <= trunc * 0.5F;
  // weights (:287-342)
  float w = opt->observation_weight;
  if (rule == 1 && normals != nullptr) {
    const float* n = normals + 3 * (static_cast<std::size_t>(row) * width + col);
    const V3 ns = mk(n[0], n[1], n[2]);
    if (all_finite(ns)) {
      const V3 view = -normalized(ray);
      const float cos_theta = cmax(0.0F, dot(normalized(ns), view));
      w *= cos_theta / d;
    }
  }
  // fuse (:230-252)
  if (w <= 0.0F) return 0;
  const std::int64_t at = vol.index(x, y, z);
  if (at < 0) return 0;  // sparse: only allocated blocks are integrated
  const std::size_t i = static_cast<std::size_t>(at);
  vol.fuse(i, tsdf, w, opt->max_weight);
  if (vol.colors != nullptr && color != nullptr && integrate_color) {
    const V3 c = color_from_pixel(color[static_cast<std::size_t>(row) * width + col]);
    float* cv = vol.colors + 4 * i;
    const V3 old = mk(cv[0], cv[1], cv[2]);
    const float cw = cv[3];
    const V3 avg = ((old * cw) + (c * w)) / (cw + w);  // vector.hpp:233-239
    cv[0] = avg.x;
    cv[1] = avg.y;
    cv[2] = avg.z;
    cv[3] = cmin(cw + w, opt->max_weight);
    return 2;
  }
  return 1;
}

// Dense full-sweep integration of one frame: IntegrationSweep<kHost>::run
// (tsdf_integration.hpp:356-370).  rule: 0 classic, 1 angle-weighted.
// world_to_camera: 3x4 row-major float.  counts[0] = voxels fused with w > 0,
// counts[1] = colour voxels fused.
int kfo_integrate(const kfo_geometry* g, int kind, int rule, void* voxels, float* colors,
                  const std::uint16_t* depth, int width, int height, const std::uint32_t* color,
                  const float* normals, const kfo_intrinsics* k, const float* world_to_camera,
                  const kfo_integration_options* opt, std::int64_t* counts) {
  Volume vol = make_volume(g, kind, voxels, colors);
  const Rigid w2c = rigid_from(world_to_camera);
  const Intr intr{k->fx, k->fy, k->cx, k->cy};
  std::int64_t fused = 0, color_fused = 0;
  for (std::int64_t z = 0; z < vol.nz; ++z)
    for (std::int64_t y = 0; y < vol.ny; ++y)
      for (std::int64_t x = 0; x < vol.nx; ++x) {
        const int r = integrate_voxel(vol, x, y, z, rule, w2c, intr, depth, width, height, color, normals, opt);
        fused += r > 0 ? 1 : 0;
        color_fused += r > 1 ? 1 : 0;
      }
  if (counts != nullptr) {
    counts[0] = fused;
    counts[1] = color_fused;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Sparse (block) volume: BlockRep<kHost> (block_rep.hpp:183-307) driven by
// TruncationBandWalk (tsdf_integration.hpp:378-437).  The volume is the
// reference's slot grid (one int32 per 8^3 block, -1 unallocated) plus a
// pool of 512-voxel slots in allocation order; every reader addresses a
// voxel through Volume::index, and a voxel of an unallocated block is a
// missing corner (trilinear.hpp:150-160), as in the reference.
// ---------------------------------------------------------------------------
namespace {
constexpr std::int64_t kBlockEdge = 8;  // grid.hpp:15-17
struct BlockDims {
  std::int64_t bx, by, bz;
  std::int64_t count() const { return bx * by * bz; }
  std::int64_t of_voxel(std::int64_t x, std::int64_t y, std::int64_t z) const {  // grid.hpp:54-59
    return ((z / kBlockEdge) * by + (y / kBlockEdge)) * bx + (x / kBlockEdge);
  }
};
BlockDims block_dims(const Volume& v) {  // grid.hpp:27-34
  return BlockDims{(v.nx + kBlockEdge - 1) / kBlockEdge, (v.ny + kBlockEdge - 1) / kBlockEdge,
                   (v.nz + kBlockEdge - 1) / kBlockEdge};
}
// RigidTransform inverse: {R^T, -R^T t} (vector.hpp:214-224).
Rigid rigid_inverse(const Rigid& m) {
  Rigid r{};
  r.rx = mk(m.rx.x, m.ry.x, m.rz.x);
  r.ry = mk(m.rx.y, m.ry.y, m.rz.y);
  r.rz = mk(m.rx.z, m.ry.z, m.rz.z);
  r.t = -rot(r, m.t);
  return r;
}
}  // namespace

// TruncationBandWalk::visit_pixel (tsdf_integration.hpp:384-428) over every
// pixel in row-major order: the visited block indices in visit order (each
// pixel skips a block equal to its previous visit).  Returns the number of
// visits; writes at most max_visits of them.
std::int64_t kfo_band_walk(const kfo_geometry* g, const std::uint16_t* depth, int width, int height,
                           const kfo_intrinsics* k, const float* world_to_camera,
                           const kfo_integration_options* opt, float step, std::int64_t* visits,
                           std::int64_t max_visits) {
  Volume vol = make_volume(g, kFloatVoxel, nullptr, nullptr);
  const BlockDims blocks = block_dims(vol);
  const Rigid c2w = rigid_inverse(rigid_from(world_to_camera));
  const Intr intr{k->fx, k->fy, k->cx, k->cy};
  const float inv_voxel = 1.0F / vol.voxel_size;
  std::int64_t n = 0;
  for (int row = 0; row < height; ++row)
    for (int col = 0; col < width; ++col) {
      float measured;
      if (!depth_in_range(depth[static_cast<std::size_t>(row) * width + col], opt->depth_scale, opt->min_depth,
                          opt->max_depth, measured))
        continue;
      const float band = opt->distance_scaled_truncation
                             ? vol.truncation + (opt->truncation_distance_scale * measured)
                             : vol.truncation;
      const int samples = static_cast<int>((2.0F * band) / step) + 1;
      std::int64_t previous = -1;
      for (int s = 0; s <= samples; ++s) {
        const float d = (measured - band) + (static_cast<float>(s) * step);
        const V3 world = apply(c2w, back_project(intr, static_cast<float>(col), static_cast<float>(row), d));
        const V3 grid = (world - vol.origin) * inv_voxel;
        if (!vol.contains(grid)) continue;
        const std::int64_t b = blocks.of_voxel(static_cast<std::int64_t>(grid.x), static_cast<std::int64_t>(grid.y),
                                               static_cast<std::int64_t>(grid.z));
        if (b == previous) continue;
        previous = b;
        if (n < max_visits) visits[n] = b;
        ++n;
      }
    }
  return n;
}

// BlockRep<kHost>::integrate (block_rep.hpp:244-307): band walk at half a
// block, allocation in walk order (slots[block] = pool slot or -1;
// state[0] = blocks allocated so far including overflowing requests,
// state[1] = capacity), the frame's work list once per block, then the
// listed blocks' voxels updated as in the dense sweep.  voxels/colors are
// the pool (capacity x 512 voxels).  counts as in kfo_integrate plus
// counts[2] = blocks in the work list.
int kfo_integrate_sparse(const kfo_geometry* g, int kind, int rule, void* voxels, float* colors,
                         std::int32_t* slots, std::int64_t* state, const std::uint16_t* depth, int width,
                         int height, const std::uint32_t* color, const float* normals, const kfo_intrinsics* k,
                         const float* world_to_camera, const kfo_integration_options* opt, std::int64_t* counts) {
  Volume vol = make_volume(g, kind, voxels, colors, slots);
  const BlockDims blocks = block_dims(vol);
  const float step = 0.5F * static_cast<float>(kBlockEdge) * vol.voxel_size;  // :380-383
  std::vector<std::int64_t> visits(static_cast<std::size_t>(width) * height * 8);
  std::int64_t n = kfo_band_walk(g, depth, width, height, k, world_to_camera, opt, step, visits.data(),
                                 static_cast<std::int64_t>(visits.size()));
  if (n > static_cast<std::int64_t>(visits.size())) {
    visits.resize(static_cast<std::size_t>(n));
    n = kfo_band_walk(g, depth, width, height, k, world_to_camera, opt, step, visits.data(), n);
  }
  std::vector<std::uint8_t> marked(static_cast<std::size_t>(blocks.count()), 0);
  std::vector<std::int64_t> work;
  for (std::int64_t i = 0; i < n; ++i) {  // allocate (block_rep.hpp:284-301)
    const std::int64_t b = visits[static_cast<std::size_t>(i)];
    if (slots[b] >= 0) {
      if (marked[b] == 0) {
        marked[b] = 1;
        work.push_back(b);
      }
      continue;
    }
    if (state[0] >= state[1]) {
      ++state[0];  // records the overflow
      continue;
    }
    slots[b] = static_cast<std::int32_t>(state[0]++);
    marked[b] = 1;
    work.push_back(b);
  }
  const Rigid w2c = rigid_from(world_to_camera);
  const Intr intr{k->fx, k->fy, k->cx, k->cy};
  std::int64_t fused = 0, color_fused = 0;
  for (const std::int64_t b : work) {  // BlockVoxels (grid.hpp:230-274)
    const std::int64_t bx = b % blocks.bx, by = (b / blocks.bx) % blocks.by, bz = b / (blocks.bx * blocks.by);
    for (std::int64_t z = bz * kBlockEdge; z < cmin(vol.nz, (bz + 1) * kBlockEdge); ++z)
      for (std::int64_t y = by * kBlockEdge; y < cmin(vol.ny, (by + 1) * kBlockEdge); ++y)
        for (std::int64_t x = bx * kBlockEdge; x < cmin(vol.nx, (bx + 1) * kBlockEdge); ++x) {
          const int r = integrate_voxel(vol, x, y, z, rule, w2c, intr, depth, width, height, color, normals, opt);
          fused += r > 0 ? 1 : 0;
          color_fused += r > 1 ? 1 : 0;
        }
  }
  if (counts != nullptr) {
    counts[0] = fused;
    counts[1] = color_fused;
    counts[2] = static_cast<std::int64_t>(work.size());
  }
  return 0;
}

// BandIntegrationSweep<kHost>::run (tsdf_integration.hpp:444-497), the
// dense kBand mode: the band walk at a one-voxel step (:462-463) marks a
// block bitmap, the set bits are listed in increasing block order, and the
// listed blocks' voxels are updated as in the full sweep; every other voxel
// keeps its value.  counts as kfo_integrate plus counts[2] = blocks listed.
int kfo_integrate_band(const kfo_geometry* g, int kind, int rule, void* voxels, float* colors,
                       const std::uint16_t* depth, int width, int height, const std::uint32_t* color,
                       const float* normals, const kfo_intrinsics* k, const float* world_to_camera,
                       const kfo_integration_options* opt, std::int64_t* counts) {
  Volume vol = make_volume(g, kind, voxels, colors);
  const BlockDims blocks = block_dims(vol);
  std::vector<std::int64_t> visits(static_cast<std::size_t>(width) * height * 16);
  std::int64_t n = kfo_band_walk(g, depth, width, height, k, world_to_camera, opt, vol.voxel_size, visits.data(),
                                 static_cast<std::int64_t>(visits.size()));
  if (n > static_cast<std::int64_t>(visits.size())) {
    visits.resize(static_cast<std::size_t>(n));
    n = kfo_band_walk(g, depth, width, height, k, world_to_camera, opt, vol.voxel_size, visits.data(), n);
  }
  std::vector<std::uint8_t> bitmap(static_cast<std::size_t>(blocks.count()), 0);
  for (std::int64_t i = 0; i < n; ++i) bitmap[static_cast<std::size_t>(visits[static_cast<std::size_t>(i)])] = 1;
  const Rigid w2c = rigid_from(world_to_camera);
  const Intr intr{k->fx, k->fy, k->cx, k->cy};
  std::int64_t fused = 0, color_fused = 0, listed = 0;
  for (std::int64_t b = 0; b < blocks.count(); ++b) {
    if (bitmap[static_cast<std::size_t>(b)] == 0) continue;
    ++listed;
    const std::int64_t bx = b % blocks.bx, by = (b / blocks.bx) % blocks.by, bz = b / (blocks.bx * blocks.by);
    for (std::int64_t z = bz * kBlockEdge; z < cmin(vol.nz, (bz + 1) * kBlockEdge); ++z)
      for (std::int64_t y = by * kBlockEdge; y < cmin(vol.ny, (by + 1) * kBlockEdge); ++y)
        for (std::int64_t x = bx * kBlockEdge; x < cmin(vol.nx, (bx + 1) * kBlockEdge); ++x) {
          const int r = integrate_voxel(vol, x, y, z, rule, w2c, intr, depth, width, height, color, normals, opt);
          fused += r > 0 ? 1 : 0;
          color_fused += r > 1 ? 1 : 0;
        }
  }
  if (counts != nullptr) {
    counts[0] = fused;
    counts[1] = color_fused;
    counts[2] = listed;
  }
  return 0;
}

// Observed-voxel count (volume.hpp:419-432).
std::int64_t kfo_observed_count(const kfo_geometry* g, int kind, const void* voxels) {
  Volume vol = make_volume(g, kind, voxels, nullptr);
  const std::size_t n = static_cast<std::size_t>(g->nx * g->ny * g->nz);
  std::int64_t c = 0;
  for (std::size_t i = 0; i < n; ++i)
    if (vol.load(i).weight > 0.0F) ++c;
  return c;
}

// ---------------------------------------------------------------------------
// Depth processing: depth_processing.hpp:54-284, depth_processing.cpp:97-163.
// ---------------------------------------------------------------------------
void kfo_bilateral(const std::uint16_t* in, int w, int h, float depth_scale, float min_d,
                   float max_d, int radius, float sigma_s, float sigma_r, std::uint16_t* out) {
  const float ss = -0.5F / (sigma_s * sigma_s);  // :118-120
  const float sr = -0.5F / (sigma_r * sigma_r);
  for (int row = 0; row < h; ++row)
    for (int col = 0; col < w; ++col) {
      float cm;
      out[row * w + col] = 0;
      if (!depth_in_range(in[row * w + col], depth_scale, min_d, max_d, cm)) continue;
      const int xb = cmax(col - radius, 0), yb = cmax(row - radius, 0);
      const int xe = cmin(col + radius, w - 1), ye = cmin(row + radius, h - 1);
      float wsum_v = 0.0F, wsum = 0.0F;
      for (int y = yb; y <= ye; ++y)
        for (int x = xb; x <= xe; ++x) {
          const std::uint16_t raw = in[y * w + x];
          float sm;
          if (!depth_in_range(raw, depth_scale, min_d, max_d, sm)) continue;
          const float pd2 = static_cast<float>(((x - col) * (x - col)) + ((y - row) * (y - row)));
          const float dd = sm - cm;
          const float wt = std::exp((pd2 * ss) + (dd * dd * sr));
          wsum_v += wt * static_cast<float>(raw);
          wsum += wt;
        }
      if (wsum <= 0.0F) continue;
      out[row * w + col] = static_cast<std::uint16_t>(std::lround(wsum_v / wsum));
    }
}

void kfo_downsample(const std::uint16_t* in, int w, int h, float depth_scale, float max_jump,
                    std::uint16_t* out) {
  const int ow = w / 2, oh = h / 2;
  for (int row = 0; row < oh; ++row)
    for (int col = 0; col < ow; ++col) {  // depth_processing.hpp:152-187
      std::uint32_t sum = 0, count = 0;
      std::uint16_t mn = 0xFFFF, mx = 0;
      for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
          const std::uint16_t s = in[(row * 2 + dy) * w + (col * 2 + dx)];
          if (s == 0) continue;
          sum += s;
          ++count;
          mn = cmin(mn, s);
          mx = cmax(mx, s);
        }
      std::uint16_t r = 0;
      if (count != 0) {
        const float hi = static_cast<float>(mx) / depth_scale;
        const float lo = static_cast<float>(mn) / depth_scale;
        if (!(hi - lo > max_jump)) r = static_cast<std::uint16_t>((sum + (count / 2U)) / count);
      }
      out[row * ow + col] = r;
    }
}

// VertexProjection with the identity camera pose (depth_processing.hpp:215-226,
// depth_processing.cpp:120-128).
void kfo_vertices(const std::uint16_t* depth, int w, int h, const kfo_intrinsics* k,
                  float depth_scale, float min_d, float max_d, float* out) {
  const Intr intr{k->fx, k->fy, k->cx, k->cy};
  const Rigid id{mk(1, 0, 0), mk(0, 1, 0), mk(0, 0, 1), mk(0, 0, 0)};
  for (int row = 0; row < h; ++row)
    for (int col = 0; col < w; ++col) {
      float d;
      V3 v = nan3();
      if (depth_in_range(depth[row * w + col], depth_scale, min_d, max_d, d))
        v = apply(id, back_project(intr, static_cast<float>(col), static_cast<float>(row), d));
      float* o = out + 3 * (static_cast<std::size_t>(row) * w + col);
      o[0] = v.x;
      o[1] = v.y;
      o[2] = v.z;
    }
}

void kfo_normals(const float* vertices, int w, int h, float max_jump, float* out) {
  auto at = [&](int x, int y) {
    const float* p = vertices + 3 * (static_cast<std::size_t>(y) * w + x);
    return mk(p[0], p[1], p[2]);
  };
  for (int row = 0; row < h; ++row)
    for (int col = 0; col < w; ++col) {  // depth_processing.hpp:251-279
      V3 n = nan3();
      if (!(col == 0 || row == 0 || col + 1 >= w || row + 1 >= h)) {
        const V3 l = at(col - 1, row), r = at(col + 1, row);
        const V3 t = at(col, row - 1), b = at(col, row + 1);
        if (all_finite(at(col, row) + l + r + t + b) && !(std::abs(r.z - l.z) > max_jump) &&
            !(std::abs(b.z - t.z) > max_jump)) {
          const V3 c = cross(b - t, r - l);
          const float len = norm(c);
          if (len > 0.0F) n = c / len;
        }
      }
      float* o = out + 3 * (static_cast<std::size_t>(row) * w + col);
      o[0] = n.x;
      o[1] = n.y;
      o[2] = n.z;
    }
}

// ---------------------------------------------------------------------------
// Raycast: Raycaster::render + SurfaceRaycast (raycasting.hpp:116-300).
// cam_to_world: 3x4 row-major float.  Returns the number of tsdf samples
// taken (the march work) in *samples when non-null.
// ---------------------------------------------------------------------------
int kfo_raycast(const kfo_geometry* g, int kind, const void* voxels, const float* colors,
                const std::int32_t* slots, const float* cam_to_world, const kfo_intrinsics* k, int width,
                int height, float min_depth, float max_depth, float step_scale, int require_all,
                int cell_normals, const float* previous_points, float* points, float* normals_out,
                std::uint32_t* colors_out, std::int64_t* samples) {
  Volume vol = make_volume(g, kind, voxels, colors, slots);
  const Sampler smp(&vol);
  const Rigid pose = rigid_from(cam_to_world);
  const Intr intr{k->fx, k->fy, k->cx, k->cy};
  std::int64_t taken = 0;
  const float base_step = vol.voxel_size * step_scale;  // raycasting.hpp:169
  for (int row = 0; row < height; ++row)
    for (int col = 0; col < width; ++col) {
      const std::size_t pi = static_cast<std::size_t>(row) * width + col;
      // seed_from_previous (raycasting.hpp:122-131): start two truncation
      // bands in front of this pixel's previous surface point
      float seed = min_depth;
      if (previous_points != nullptr) {
        const V3 prev = mk(previous_points[3 * pi], previous_points[3 * pi + 1], previous_points[3 * pi + 2]);
        if (all_finite(prev) && dot(prev, prev) > 0.0F) {
          const float previous_depth = norm(prev - pose.t);
          seed = cmax(min_depth, previous_depth - (2.0F * vol.truncation));
        }
      }
      V3 P = nan3(), N = nan3();
      std::uint32_t C = 0;
      const V3 translated = back_project(intr, static_cast<float>(col), static_cast<float>(row), 1.0F);
      const V3 direction = rot(pose, translated);
      const float dn = norm(direction);
      if (dn > 0.0F) {
        const V3 dir = direction / dn;
        // find_zero_crossing (raycasting.hpp:162-211)
        bool have_prev = false, hit = false;
        float prev_t = 0.0F;
        V3 prev_p{}, surface{};
        float len = seed;
        while (len <= max_depth) {
          const V3 point = pose.t + (len * dir);
          float t;
          ++taken;
          if (!smp.tsdf(point, require_all != 0, t)) {
            have_prev = false;
            len += base_step;
            continue;
          }
          if (have_prev && prev_t > 0.0F && t <= 0.0F) {
            const float gap = prev_t - t;
            const float alpha = gap != 0.0F ? prev_t / gap : 0.0F;
            surface = prev_p + (alpha * (point - prev_p));
            hit = true;
            break;
          }
          if (have_prev && prev_t < 0.0F && t > 0.0F) break;
          have_prev = true;
          prev_t = t;
          prev_p = point;
          const float pre = t * vol.truncation * step_scale;
          const float step = cmax(base_step, pre);
          len += t > 0.0F ? step : base_step;
        }
        if (hit) {
          // write_surface_sample (raycasting.hpp:215-245)
          V3 n;
          const bool ok = cell_normals ? smp.cell_normal(surface, n)
                                       : smp.normal(surface, require_all != 0, n);
          if (ok) {
            P = surface;
            N = n;
            if (colors != nullptr) {
              V3 c;
              if (smp.color(surface, c)) C = pixel_from_color(c);
            } else {
              const V3 view = normalized(pose.t - surface);
              const float inten = 255.0F * cmax(0.0F, dot(n, view));
              C = pixel_from_color(mk(inten, inten, inten));
            }
          }
        }
      }
      points[3 * pi] = P.x;
      points[3 * pi + 1] = P.y;
      points[3 * pi + 2] = P.z;
      normals_out[3 * pi] = N.x;
      normals_out[3 * pi + 1] = N.y;
      normals_out[3 * pi + 2] = N.z;
      colors_out[pi] = C;
    }
  if (samples != nullptr) *samples = taken;
  return 0;
}

// ---------------------------------------------------------------------------
// ICP: CorrespondenceSearch::match (icp_correspondence.hpp:244-294),
// IcpCorrespondence::from_match (:44-52), IcpNormalEquations::accumulate
// (:111-129), find_correspondences (icp_optimizer.cpp:304-317).
// Sums layout: [0..20] upper-triangle J^T J row-major, [21..26] J^T r,
// [27] distance sum, [28] residual sum of squares.
// sums64: FP64 accumulation of the float per-correspondence terms;
// sums32: the reference's sequential FP32 accumulation (may be null).
// ---------------------------------------------------------------------------
std::int64_t kfo_icp_sums(const float* live_v, const float* live_n, int lw, int lh,
                          const float* model_v, const float* model_n, int mw, int mh,
                          const kfo_intrinsics* k, const float* camera, const float* model,
                          float max_dist, float min_dot, double* sums64, float* sums32) {
  const Rigid cam = rigid_from(camera);
  const Rigid mod = rigid_from(model);
  const Intr intr{k->fx, k->fy, k->cx, k->cy};
  double s64[29] = {};
  float s32[29] = {};
  std::int64_t count = 0;
  auto at = [](const float* m, int w, int x, int y) {
    const float* p = m + 3 * (static_cast<std::size_t>(y) * w + x);
    return mk(p[0], p[1], p[2]);
  };
  for (int y = 0; y < lh; ++y)
    for (int x = 0; x < lw; ++x) {
      const V3 lv = at(live_v, lw, x, y), ln = at(live_n, lw, x, y);
      if (!all_finite(lv) || !all_finite(ln)) continue;
      const V3 src = apply(cam, lv);
      const V3 sim = apply(mod, src);
      if (!all_finite(sim) || sim.z <= 0.0F) continue;
      float u, v;
      project(intr, sim, u, v);
      long col, row;
      if (!round_index(u, mw, col) || !round_index(v, mh, row)) continue;
      const V3 mv = at(model_v, mw, static_cast<int>(col), static_cast<int>(row));
      const V3 mn = at(model_n, mw, static_cast<int>(col), static_cast<int>(row));
      if (!all_finite(mv) || !all_finite(mn)) continue;
      const V3 sn = normalized(rot(cam, ln));
      const float dist = norm(src - mv);
      if (dist > max_dist) continue;
      if (dot(sn, mn) < min_dot) continue;
      const V3 r3 = cross(src, mn);
      const float J[6] = {r3.x, r3.y, r3.z, mn.x, mn.y, mn.z};
      const float res = dot(mn, mv - src);
      int e = 0;
      for (int r = 0; r < 6; ++r)
        for (int c = r; c < 6; ++c, ++e) {
          const float p = J[r] * J[c];
          s32[e] += p;
          s64[e] += static_cast<double>(p);
        }
      for (int r = 0; r < 6; ++r) {
        const float p = J[r] * res;
        s32[21 + r] += p;
        s64[21 + r] += static_cast<double>(p);
      }
      s32[27] += dist;
      s64[27] += static_cast<double>(dist);
      const float rr = res * res;
      s32[28] += rr;
      s64[28] += static_cast<double>(rr);
      ++count;
    }
  if (sums64 != nullptr) std::memcpy(sums64, s64, sizeof(s64));
  if (sums32 != nullptr) std::memcpy(sums32, s32, sizeof(s32));
  return count;
}

// Cyclic Jacobi eigenvalues of a symmetric 6x6 FP64 matrix (row-major).
// Replaces Eigen::SelfAdjointEigenSolver<float> (icp_optimizer.cpp:326-333).
void kfo_sym_eig6(const double* a_in, double* evals) {
  double a[6][6];
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) a[r][c] = a_in[r * 6 + c];
  // Sweeps stop once the off-diagonal mass is below (1e-16)^2 of the
  // diagonal's: the eigenvalues are then exact to ~1e-16 of the norm.
  for (int sweep = 0; sweep < 32; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int d = 0; d < 6; ++d) diag += a[d][d] * a[d][d];
    for (int p = 0; p < 6; ++p)
      for (int q = p + 1; q < 6; ++q) off += a[p][q] * a[p][q];
    if (off <= 1.0e-32 * diag) break;
    for (int p = 0; p < 6; ++p)
      for (int q = p + 1; q < 6; ++q) {
        if (a[p][q] == 0.0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0);
        const double s = t * c;
        for (int k2 = 0; k2 < 6; ++k2) {
          const double akp = a[k2][p], akq = a[k2][q];
          a[k2][p] = c * akp - s * akq;
          a[k2][q] = s * akp + c * akq;
        }
        for (int k2 = 0; k2 < 6; ++k2) {
          const double apk = a[p][k2], aqk = a[q][k2];
          a[p][k2] = c * apk - s * aqk;
          a[q][k2] = s * apk + c * aqk;
        }
      }
  }
  for (int i = 0; i < 6; ++i) evals[i] = a[i][i];
}

struct kfo_icp_options {
  std::int64_t min_correspondences;       // icp_optimizer.hpp:14
  float max_point_distance;               // icp_correspondence.hpp:25
  float min_normal_dot;                   // :27
  float min_update_translation;           // icp_optimizer.hpp:15
  float min_update_rotation;              // :16
  float max_update_translation;           // :17
  float max_update_rotation;              // :18
  float min_system_eigenvalue;            // :19
  float max_condition_number;             // :20
  int damping_mode;                       // 0 none, 1 identity, 2 diagonal
  float damping_lambda;
  int adaptive_damping;
  float lambda_increase, lambda_decrease, min_lambda, max_lambda;  // :84-98
};

struct kfo_icp_result {
  int status;  // 0 max-iterations, 1 converged, 2 too few, 3 solve failed, 4 update too large, 5 unconstrained
  int trials;
  int rejected;
  std::int64_t correspondences;
  double mean_distance;
  double mean_squared_residual;
  double min_eigenvalue;
  double condition_number;
  double update_translation;
  double update_rotation;
  double lambda;
};

// FP64 LDL^T with symmetric diagonal pivoting (largest remaining |d| first),
// the factorisation Eigen::LDLT uses in the reference host solve
// (icp_optimizer.cpp:391): a pivot with |d| <= DBL_MIN contributes a zero
// solution component, so a damped semidefinite system still solves; the
// solve fails only on a non-finite solution.
static bool solve6(const double* eq, int damping_mode, double scale, double* x) {
  double a[6][6];
  int e = 0;
  for (int r = 0; r < 6; ++r)
    for (int c = r; c < 6; ++c, ++e) {
      a[r][c] = eq[e];
      a[c][r] = eq[e];
    }
  for (int r = 0; r < 6; ++r) {
    if (damping_mode == 1) a[r][r] += scale;
    else if (damping_mode == 2) a[r][r] += scale * a[r][r];
  }
  int perm[6] = {0, 1, 2, 3, 4, 5};
  double d[6];
  for (int k = 0; k < 6; ++k) {
    int piv = k;
    for (int i = k + 1; i < 6; ++i)
      if (std::fabs(a[i][i]) > std::fabs(a[piv][piv])) piv = i;
    if (piv != k) {
      for (int c = 0; c < 6; ++c) { const double t = a[k][c]; a[k][c] = a[piv][c]; a[piv][c] = t; }
      for (int r = 0; r < 6; ++r) { const double t = a[r][k]; a[r][k] = a[r][piv]; a[r][piv] = t; }
      const int tp = perm[k]; perm[k] = perm[piv]; perm[piv] = tp;
    }
    d[k] = a[k][k];
    const bool zero = std::fabs(d[k]) <= 2.2250738585072014e-308;
    for (int i = k + 1; i < 6; ++i) a[i][k] = zero ? 0.0 : a[i][k] / d[k];
    for (int i = k + 1; i < 6; ++i)
      for (int j = k + 1; j <= i; ++j) {
        a[i][j] -= a[i][k] * d[k] * a[j][k];
        a[j][i] = a[i][j];
      }
  }
  double y[6];
  for (int i = 0; i < 6; ++i) {
    double s = eq[21 + perm[i]];
    for (int q = 0; q < i; ++q) s -= a[i][q] * y[q];
    y[i] = s;
  }
  for (int i = 0; i < 6; ++i) y[i] = std::fabs(d[i]) <= 2.2250738585072014e-308 ? 0.0 : y[i] / d[i];
  double z[6];
  for (int i = 5; i >= 0; --i) {
    double s = y[i];
    for (int q = i + 1; q < 6; ++q) s -= a[q][i] * z[q];
    z[i] = s;
  }
  for (int i = 0; i < 6; ++i) x[perm[i]] = z[i];
  for (int i = 0; i < 6; ++i)
    if (!std::isfinite(x[i])) return false;
  return true;
}

// pose <- exp([w]) pose with t' = R_inc t + t_inc (FP64 Rodrigues; reference:
// icp_optimizer.cpp:349-365 and the composition at :286 /
// cuda/icp_correspondence.cu:298-336).
void kfo_compose_increment(const double* x, double* pose) {
  const double wx = x[0], wy = x[1], wz = x[2];
  const double angle = std::sqrt(wx * wx + wy * wy + wz * wz);
  double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  if (angle > 1.0e-12) {
    const double ux = wx / angle, uy = wy / angle, uz = wz / angle;
    const double s = std::sin(angle), c = std::cos(angle), oc = 1.0 - c;
    R[0] = c + ux * ux * oc;
    R[1] = ux * uy * oc - uz * s;
    R[2] = ux * uz * oc + uy * s;
    R[3] = uy * ux * oc + uz * s;
    R[4] = c + uy * uy * oc;
    R[5] = uy * uz * oc - ux * s;
    R[6] = uz * ux * oc - uy * s;
    R[7] = uz * uy * oc + ux * s;
    R[8] = c + uz * uz * oc;
  }
  double out[16];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c)
      out[r * 4 + c] = R[r * 3 + 0] * pose[0 * 4 + c] + R[r * 3 + 1] * pose[1 * 4 + c] +
                       R[r * 3 + 2] * pose[2 * 4 + c];
    out[r * 4 + 3] = R[r * 3 + 0] * pose[3] + R[r * 3 + 1] * pose[7] + R[r * 3 + 2] * pose[11] +
                     x[3 + r];
  }
  out[12] = 0;
  out[13] = 0;
  out[14] = 0;
  out[15] = 1;
  std::memcpy(pose, out, sizeof(out));
}

// check_system_stability (icp_optimizer.cpp:319-347) on FP64 sums.
static bool stable_system(const double* eq, std::int64_t count, const kfo_icp_options* o,
                          double* min_eig = nullptr, double* cond = nullptr) {
  if (count < o->min_correspondences) return false;
  double m[36];
  int e = 0;
  for (int rr = 0; rr < 6; ++rr)
    for (int c = rr; c < 6; ++c, ++e) {
      m[rr * 6 + c] = eq[e];
      m[c * 6 + rr] = eq[e];
    }
  double ev[6];
  kfo_sym_eig6(m, ev);
  double mn = ev[0], mx = ev[0];
  for (double v : ev) {
    if (!std::isfinite(v)) return false;
    mn = std::min(mn, v);
    mx = std::max(mx, v);
  }
  if (min_eig != nullptr) *min_eig = mn;
  if (!(mn > 0.0 && mx > 0.0)) return false;
  if (cond != nullptr) *cond = mx / mn;
  return mn >= o->min_system_eigenvalue && mx / mn <= o->max_condition_number;
}

// Whole-level Gauss-Newton / LM loop with latched status: the semantics of
// the reference's device loop (cuda/icp_correspondence.cu:216-343 plus the
// host epilogue icp_optimizer.cpp:82-169): every trial reduces, solves and
// composes; the stability spectrum is checked once on the final system.
// pose_inout: initial camera_to_world (FP64 4x4 row-major), replaced by the
// result on success.  model_pose: the pose the model maps were rendered at.
int kfo_icp_track_level(const float* live_v, const float* live_n, int lw, int lh,
                        const float* model_v, const float* model_n, int mw, int mh,
                        const kfo_intrinsics* k, const double* model_pose, double* pose_inout,
                        int iterations, const kfo_icp_options* o, kfo_icp_result* res) {
  double model_inv[16];
  rigid_inverse64(model_pose, model_inv);
  float model_f[12], cam_f[12];
  for (int i = 0; i < 12; ++i) model_f[i] = static_cast<float>(model_inv[i]);
  double pose[16], best_pose[16];
  std::memcpy(pose, pose_inout, sizeof(pose));
  std::memcpy(best_pose, pose, sizeof(pose));
  double eq[29] = {}, best_eq[29] = {};
  std::int64_t eq_count = 0, best_count = 0;
  double best_cost = 0.0, lambda = o->damping_lambda;
  const bool damping = o->damping_mode != 0;
  const bool adaptive = damping && o->adaptive_damping != 0;
  kfo_icp_result r{};
  int status = 0;
  for (int it = 0; it < iterations && status == 0; ++it) {
    for (int i = 0; i < 12; ++i) cam_f[i] = static_cast<float>(pose[i]);
    double sums[29];
    const std::int64_t count = kfo_icp_sums(live_v, live_n, lw, lh, model_v, model_n, mw, mh, k,
                                            cam_f, model_f, o->max_point_distance,
                                            o->min_normal_dot, sums, nullptr);
    if (count < o->min_correspondences) {
      std::memcpy(eq, sums, sizeof(eq));
      eq_count = count;
      status = 2;
      break;
    }
    if (r.trials == 0) lambda = o->damping_lambda;
    bool accepted = true;
    if (adaptive) {
      const double cost = sums[28] / static_cast<double>(count);
      accepted = r.trials == 0 || cost <= best_cost;
      if (accepted) {
        best_cost = cost;
        std::memcpy(best_eq, sums, sizeof(best_eq));
        best_count = count;
        std::memcpy(best_pose, pose, sizeof(pose));
        lambda = std::max(lambda * o->lambda_decrease, static_cast<double>(o->min_lambda));
      } else {
        ++r.rejected;
        std::memcpy(pose, best_pose, sizeof(pose));
        lambda = std::min(lambda * o->lambda_increase, static_cast<double>(o->max_lambda));
      }
      std::memcpy(eq, best_eq, sizeof(eq));
      eq_count = best_count;
    } else {
      std::memcpy(eq, sums, sizeof(eq));
      eq_count = count;
    }
    ++r.trials;
    const int attempts = damping ? 8 : 1;
    double x[6] = {}, scale = lambda, used = lambda, ut = 0.0, ang = 0.0;
    bool solved = false;
    for (int a = 0; a < attempts; ++a) {
      double xs[6];
      if (solve6(eq, o->damping_mode, scale, xs)) {
        solved = true;
        used = scale;
        std::memcpy(x, xs, sizeof(x));
        ang = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        ut = std::sqrt(x[3] * x[3] + x[4] * x[4] + x[5] * x[5]);
        if (!damping || (ut <= o->max_update_translation && ang <= o->max_update_rotation)) break;
      }
      scale = std::min(scale * o->lambda_increase, static_cast<double>(o->max_lambda));
    }
    // A failing trial reports kUnconstrainedSystem when the undamped system
    // is unstable: the reference's host loop vetoes such systems before the
    // solve (icp_optimizer.cpp:263-270).
    if (!solved) {
      status = (!damping && !stable_system(eq, eq_count, o)) ? 5 : 3;
      break;
    }
    lambda = used;
    r.update_translation = ut;
    r.update_rotation = ang;
    if (ut > o->max_update_translation || ang > o->max_update_rotation) {
      status = (!damping && !stable_system(eq, eq_count, o)) ? 5 : 4;
      break;
    }
    kfo_compose_increment(x, pose);
    if (accepted && ut < o->min_update_translation && ang < o->min_update_rotation) status = 1;
  }
  r.lambda = lambda;
  r.correspondences = eq_count;
  if (eq_count > 0) {
    r.mean_distance = eq[27] / static_cast<double>(eq_count);
    r.mean_squared_residual = eq[28] / static_cast<double>(eq_count);
  }
  if (status == 0 || status == 1) {
    // stability once per level on the final system
    const bool stable = stable_system(eq, eq_count, o, &r.min_eigenvalue, &r.condition_number);
    if (!stable && !damping) status = 5;
  }
  r.status = status;
  if (status == 0 || status == 1) std::memcpy(pose_inout, pose, sizeof(pose));
  if (res != nullptr) *res = r;
  return status;
}

// ---------------------------------------------------------------------------
// Marching cubes: MarchingCubes::extract + MeshBuilder
// (marching_cubes.cpp:315-507), dense traversal.
// ---------------------------------------------------------------------------
#include "mc_cases.inc"

struct kfo_mesh {
  std::vector<float> positions, normals, colors;  // 3 per vertex
  std::vector<std::uint32_t> triangles;           // 3 per triangle
};

namespace {
constexpr int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                               {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
constexpr int kEdge[12][3] = {{0, 1, 0}, {1, 2, 1}, {3, 2, 0}, {0, 3, 1}, {4, 5, 0}, {5, 6, 1},
                              {7, 6, 0}, {4, 7, 1}, {0, 4, 2}, {1, 5, 2}, {2, 6, 2}, {3, 7, 2}};
int edge_of(char c) { return c <= '9' ? c - '0' : c - 'a' + 10; }
}  // namespace

// MarchingCubes::extract (marching_cubes.cpp:478-507).  slots == nullptr:
// dense traversal, cells x-fastest; otherwise the sparse traversal: blocks
// with slots[block] >= 0 in flat order, each block's cells x-fastest within
// the cell grid (BlockVoxels over (N - 1)^3).
void* kfo_mesh_extract_traversal(const kfo_geometry* g, int kind, const void* voxels, const float* colors,
                                 float min_weight, const std::int32_t* slots) {
  auto* mesh = new kfo_mesh();
  Volume vol = make_volume(g, kind, voxels, colors, slots);
  if (vol.nx < 2 || vol.ny < 2 || vol.nz < 2) return mesh;
  const Sampler smp(&vol);
  std::unordered_map<std::size_t, std::uint32_t> welded;
  std::vector<V3> pos, nrm, col;
  auto add_cell = [&](std::int64_t cx, std::int64_t cy, std::int64_t cz) {
        float val[8];
        unsigned cube = 0;
        for (int c = 0; c < 8; ++c) {  // add_cell (:367-387)
          const std::int64_t at = vol.index(cx + kCorner[c][0], cy + kCorner[c][1], cz + kCorner[c][2]);
          if (at < 0) return;  // a corner in an unallocated block is missing
          const Loaded v = vol.load(static_cast<std::size_t>(at));
          if (!v.observed || !v.finite || v.weight < min_weight) return;
          val[c] = v.tsdf;
          if (val[c] < 0.0F) cube |= 1U << c;
        }
        if (cube == 0 || cube == 255) return;
        const char* tri = kTriCases[cube];
        auto vertex_on_edge = [&](int edge) -> std::uint32_t {  // :412-448
          const int ne = kEdge[edge][0], fe = kEdge[edge][1];
          const std::int64_t nx = cx + kCorner[ne][0], ny = cy + kCorner[ne][1], nz = cz + kCorner[ne][2];
          const std::size_t key = vol.flat(nx, ny, nz) * 3 + static_cast<std::size_t>(kEdge[edge][2]);
          const auto [it, inserted] = welded.try_emplace(key, static_cast<std::uint32_t>(pos.size()));
          if (!inserted) return it->second;
          const float nv = val[ne], fv = val[fe];
          const V3 np = vol.cell_center(nx, ny, nz);
          const V3 fp = vol.cell_center(cx + kCorner[fe][0], cy + kCorner[fe][1], cz + kCorner[fe][2]);
          const float frac = -nv / (fv - nv);
          const V3 p = np + (frac * (fp - np));
          pos.push_back(p);
          V3 n;
          nrm.push_back(smp.normal(p, false, n) ? n : V3{});
          V3 c;
          col.push_back(smp.color(p, c) ? c : mk(255.0F, 255.0F, 255.0F));
          return it->second;
        };
        for (std::size_t e = 0; tri[e] != '\0'; e += 3) {  // reversed emit (:393-404)
          const std::uint32_t a = vertex_on_edge(edge_of(tri[e]));
          const std::uint32_t b = vertex_on_edge(edge_of(tri[e + 2]));
          const std::uint32_t c = vertex_on_edge(edge_of(tri[e + 1]));
          mesh->triangles.insert(mesh->triangles.end(), {a, b, c});
        }
  };
  if (slots == nullptr) {
    for (std::int64_t cz = 0; cz < vol.nz - 1; ++cz)
      for (std::int64_t cy = 0; cy < vol.ny - 1; ++cy)
        for (std::int64_t cx = 0; cx < vol.nx - 1; ++cx) add_cell(cx, cy, cz);
  } else {
    const BlockDims blocks = block_dims(vol);
    for (std::int64_t b = 0; b < blocks.count(); ++b) {
      if (slots[b] < 0) continue;
      const std::int64_t bx = b % blocks.bx, by = (b / blocks.bx) % blocks.by, bz = b / (blocks.bx * blocks.by);
      for (std::int64_t cz = bz * kBlockEdge; cz < cmin(vol.nz - 1, (bz + 1) * kBlockEdge); ++cz)
        for (std::int64_t cy = by * kBlockEdge; cy < cmin(vol.ny - 1, (by + 1) * kBlockEdge); ++cy)
          for (std::int64_t cx = bx * kBlockEdge; cx < cmin(vol.nx - 1, (bx + 1) * kBlockEdge); ++cx)
            add_cell(cx, cy, cz);
    }
  }
  // fill_missing_normals (:452-467)
  std::unordered_map<std::uint32_t, V3> acc;
  const std::size_t ntri = mesh->triangles.size() / 3;
  for (std::size_t t = 0; t < ntri; ++t) {
    const std::uint32_t* tr = &mesh->triangles[3 * t];
    const V3 face = cross(pos[tr[1]] - pos[tr[0]], pos[tr[2]] - pos[tr[0]]);
    for (int j = 0; j < 3; ++j) {
      const V3 n = nrm[tr[j]];
      if (n.x == 0.0F && n.y == 0.0F && n.z == 0.0F) {
        V3& s = acc[tr[j]];
        s = s + face;
      }
    }
  }
  for (auto& [i, s] : acc) nrm[i] = normalized(s);
  for (std::size_t i = 0; i < pos.size(); ++i) {
    mesh->positions.insert(mesh->positions.end(), {pos[i].x, pos[i].y, pos[i].z});
    mesh->normals.insert(mesh->normals.end(), {nrm[i].x, nrm[i].y, nrm[i].z});
    mesh->colors.insert(mesh->colors.end(), {col[i].x, col[i].y, col[i].z});
  }
  return mesh;
}

void* kfo_mesh_extract(const kfo_geometry* g, int kind, const void* voxels, const float* colors,
                       float min_weight) {
  return kfo_mesh_extract_traversal(g, kind, voxels, colors, min_weight, nullptr);
}

void kfo_mesh_sizes(const void* m, std::int64_t* nverts, std::int64_t* ntris) {
  const auto* mesh = static_cast<const kfo_mesh*>(m);
  *nverts = static_cast<std::int64_t>(mesh->positions.size() / 3);
  *ntris = static_cast<std::int64_t>(mesh->triangles.size() / 3);
}

void kfo_mesh_copy(const void* m, float* positions, float* normals, float* colors,
                   std::uint32_t* triangles) {
  const auto* mesh = static_cast<const kfo_mesh*>(m);
  std::memcpy(positions, mesh->positions.data(), mesh->positions.size() * sizeof(float));
  std::memcpy(normals, mesh->normals.data(), mesh->normals.size() * sizeof(float));
  std::memcpy(colors, mesh->colors.data(), mesh->colors.size() * sizeof(float));
  std::memcpy(triangles, mesh->triangles.data(), mesh->triangles.size() * sizeof(std::uint32_t));
}

void kfo_mesh_free(void* m) { delete static_cast<kfo_mesh*>(m); }

// Codec round trips for the golden tests.
std::int16_t kfo_q_encode(float t) { return q_encode(t); }
float kfo_q_decode(std::int16_t r) { return q_decode(r); }
std::uint16_t kfo_bf_encode(float t) { return bf_encode(t); }
float kfo_bf_decode(std::uint16_t r) { return bf_decode(r); }
std::uint16_t kfo_w_encode(float w) { return w_encode(w); }
float kfo_w_decode(std::uint16_t r) { return w_decode(r); }

}  // extern "C"

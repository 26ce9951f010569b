// Internal declarations of libkf_hip: the context, device buffers and the
// float math shared by the kernels.
//
// Bit-exact parity with the reference CPU path rests on three rules kept by
// every kernel here:
//  * the same float operation order as the reference host code (dot is
//    (x*x + y*y) + z*z, include/kinectfusion/vector.hpp:143-146; cell centres
//    are origin + (i + 0.5) * voxel, volume.hpp:276-283; projection is
//    fx * x / z + cx, rgbd.hpp:40-44);
//  * IEEE division and square root and no FMA contraction (the build uses
//    -ffp-contract=off and hip's correctly rounded fp32 divide/sqrt);
//  * the host form of normalized(): divide by the length (vector.hpp:164-174).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/kf_hip.h"

namespace kf {

constexpr int kMaxLevels = 4;
// integration brick: 16 x 4 x 8 voxels, one wave (a z-column of 8 per lane);
// it is two 8^3 blocks along x, one y-half of each
constexpr int kBrickX = 16;
constexpr int kBrickY = 4;
constexpr int kBrickZ = 8;
// A listed brick: its coordinates packed as bx | by << 8 | bz << 18
// (resolution <= 4096 per axis: bx < 256, by < 1024, bz < 512).
__host__ __device__ inline uint32_t pack_brick(int bx, int by, int bz) {
  return static_cast<uint32_t>(bx) | (static_cast<uint32_t>(by) << 8U) | (static_cast<uint32_t>(bz) << 18U);
}
constexpr int kTile = 16;    // depth-tile edge for the integration cull

// ---------------------------------------------------------------------------
// Float vector math (host + device).
// ---------------------------------------------------------------------------
struct V3 {
  float x, y, z;
};
__host__ __device__ inline V3 mk3(float x, float y, float z) { return V3{x, y, z}; }
__host__ __device__ inline V3 add3(V3 a, V3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ inline V3 sub3(V3 a, V3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ inline V3 scale3(V3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__host__ __device__ inline V3 div3(V3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
__host__ __device__ inline float dot3(V3 a, V3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
__host__ __device__ inline V3 cross3(V3 a, V3 b) {
  return mk3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
__host__ __device__ inline float norm3(V3 v) { return sqrtf(dot3(v, v)); }
__host__ __device__ inline V3 normalize3(V3 v) {
  const float len = norm3(v);
  return len > 0.0F ? div3(v, len) : mk3(0.0F, 0.0F, 0.0F);
}
__host__ __device__ inline bool finite3(V3 v) {
  return isfinite(v.x) && isfinite(v.y) && isfinite(v.z);
}
__host__ __device__ inline float fmin_ref(float a, float b) { return b < a ? b : a; }
__host__ __device__ inline float fmax_ref(float a, float b) { return a < b ? b : a; }
__host__ __device__ inline float fclamp_ref(float v, float lo, float hi) {
  if (v < lo) return lo;
  if (hi < v) return hi;
  return v;
}
__host__ __device__ inline V3 ld3(const float4& v) { return mk3(v.x, v.y, v.z); }
__host__ __device__ inline float4 st3(V3 v) { return make_float4(v.x, v.y, v.z, 0.0F); }
__device__ inline V3 nan3() {
  const float n = __builtin_nanf("");
  return mk3(n, n, n);
}

// Rigid transform [R | t], rows of R.
struct Rigid {
  V3 r0, r1, r2, t;
};
__host__ __device__ inline V3 rot(const Rigid& m, V3 p) {
  return mk3(dot3(m.r0, p), dot3(m.r1, p), dot3(m.r2, p));
}
__host__ __device__ inline V3 xform(const Rigid& m, V3 p) { return add3(rot(m, p), m.t); }
__host__ __device__ inline Rigid rigid_from12(const float* m) {
  return Rigid{mk3(m[0], m[1], m[2]), mk3(m[4], m[5], m[6]), mk3(m[8], m[9], m[10]),
               mk3(m[3], m[7], m[11])};
}
__host__ __device__ inline Rigid rigid_from_pose64(const double* p) {
  auto f = [&](int i) { return static_cast<float>(p[i]); };
  return Rigid{mk3(f(0), f(1), f(2)), mk3(f(4), f(5), f(6)), mk3(f(8), f(9), f(10)),
               mk3(f(3), f(7), f(11))};
}
// out = [R^T, -R^T t] in FP64.
__host__ __device__ inline void rigid_inverse64(const double* a, double* out) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) out[r * 4 + c] = a[c * 4 + r];
  for (int r = 0; r < 3; ++r)
    out[r * 4 + 3] = -(out[r * 4 + 0] * a[3] + out[r * 4 + 1] * a[7] + out[r * 4 + 2] * a[11]);
  out[12] = 0.0;
  out[13] = 0.0;
  out[14] = 0.0;
  out[15] = 1.0;
}

struct Intr {
  float fx, fy, cx, cy;
};
__host__ __device__ inline void project(const Intr& k, V3 p, float& u, float& v) {
  u = (k.fx * p.x / p.z) + k.cx;
  v = (k.fy * p.y / p.z) + k.cy;
}
__host__ __device__ inline V3 back_project(const Intr& k, float px, float py, float d) {
  return mk3((px - k.cx) * d / k.fx, (py - k.cy) * d / k.fy, d);
}
// lround(v) accepted only when it lands in [0, limit).  roundf rounds half
// away from zero like std::lround; NaN fails the comparison.
__host__ __device__ inline bool round_index(float v, int limit, int& out) {
  const float r = roundf(v);
  if (!(r >= 0.0F && r < static_cast<float>(limit))) return false;
  out = static_cast<int>(r);
  return true;
}
__host__ __device__ inline bool depth_in_range(uint16_t raw, float scale, float lo, float hi,
                                               float& m) {
  if (raw == 0) return false;
  m = static_cast<float>(raw) / scale;
  return !(m < lo || m > hi);
}

// ---------------------------------------------------------------------------
// Voxel stores (include/kinectfusion/volume.hpp:28-153).
// ---------------------------------------------------------------------------
constexpr float kQScale = 32767.0F;
constexpr float kQInvScale = 1.0F / 32767.0F;
constexpr float kWScale = 256.0F;
constexpr float kWInvScale = 1.0F / 256.0F;
constexpr float kWMax = 65535.0F / 256.0F;

__host__ __device__ inline uint32_t q_encode(float t) {
  return static_cast<uint16_t>(static_cast<int16_t>(fclamp_ref(t, -1.0F, 1.0F) * kQScale));
}
__host__ __device__ inline float q_decode(uint32_t raw16) {
  return static_cast<float>(static_cast<int16_t>(static_cast<uint16_t>(raw16))) * kQInvScale;
}
__host__ __device__ inline uint32_t bf_encode(float t) {
  const uint32_t b = __builtin_bit_cast(uint32_t, fclamp_ref(t, -1.0F, 1.0F));
  return ((b + 0x7FFFU + ((b >> 16U) & 1U)) >> 16U) & 0xFFFFU;
}
__host__ __device__ inline float bf_decode(uint32_t raw16) {
  return __builtin_bit_cast(float, (raw16 & 0xFFFFU) << 16U);
}
__host__ __device__ inline uint32_t w_encode(float w) {
  return static_cast<uint16_t>(fclamp_ref(w, 0.0F, kWMax) * kWScale);
}
__host__ __device__ inline float w_decode(uint32_t raw16) {
  return static_cast<float>(raw16 & 0xFFFFU) * kWInvScale;
}

// Float view of a stored voxel: tsdf, weight, observed, finite.
template <int Kind>
struct VoxelIO;

template <>
struct VoxelIO<KF_VOXEL_FLOAT> {
  using Raw = float2;
  __device__ static inline void unpack(Raw v, float& t, float& w, bool& obs, bool& fin) {
    t = v.x;
    w = v.y;
    obs = v.y > 0.0F;
    fin = isfinite(v.x);
  }
  __device__ static inline Raw pack(float t, float w) { return make_float2(t, w); }
  __host__ __device__ static inline Raw fresh() { return make_float2(1.0F, 0.0F); }
};

template <>
struct VoxelIO<KF_VOXEL_QUANTIZED> {
  using Raw = uint32_t;  // low 16 bits int16 tsdf, high 16 bits uint16 weight
  __device__ static inline void unpack(Raw v, float& t, float& w, bool& obs, bool& fin) {
    t = q_decode(v & 0xFFFFU);
    w = w_decode(v >> 16U);
    obs = (v >> 16U) != 0U;
    fin = true;
  }
  __device__ static inline Raw pack(float t, float w) { return q_encode(t) | (w_encode(w) << 16U); }
  __host__ __device__ static inline Raw fresh() { return q_encode(1.0F); }
};

template <>
struct VoxelIO<KF_VOXEL_BF16> {
  using Raw = uint32_t;  // low 16 bits bf16 tsdf, high 16 bits uint16 weight
  __device__ static inline void unpack(Raw v, float& t, float& w, bool& obs, bool& fin) {
    t = bf_decode(v & 0xFFFFU);
    w = w_decode(v >> 16U);
    obs = (v >> 16U) != 0U;
    fin = true;
  }
  __device__ static inline Raw pack(float t, float w) { return bf_encode(t) | (w_encode(w) << 16U); }
  __host__ __device__ static inline Raw fresh() { return bf_encode(1.0F); }
};

inline size_t voxel_bytes(int kind) { return kind == KF_VOXEL_FLOAT ? 8 : 4; }

// Dense volume geometry passed to kernels by value.
struct Grid {
  int nx, ny, nz;
  float voxel;
  float ox, oy, oz;
  float trunc;
  __host__ __device__ inline size_t flat(int x, int y, int z) const {
    return (static_cast<size_t>(z) * ny + y) * nx + x;
  }
  __host__ __device__ inline V3 center(int x, int y, int z) const {
    return add3(mk3(ox, oy, oz),
                scale3(add3(mk3(static_cast<float>(x), static_cast<float>(y), static_cast<float>(z)),
                            mk3(0.5F, 0.5F, 0.5F)),
                       voxel));
  }
};

// ---------------------------------------------------------------------------
// Hashed (sparse) volume: 8^3-voxel blocks in a pool of slots, found through
// an open-addressed hash table of block keys with linear probing.  A block
// key is the reference's flat x-major block index (grid.hpp:54-59); voxels
// of a block sit at (z % 8 * 8 + y % 8) * 8 + x % 8 (grid.hpp:61-69).
// ---------------------------------------------------------------------------
constexpr unsigned long long kHashEmpty = ~0ULL;
constexpr uint32_t kNoSlot = 0xFFFFFFFFU;       // key present, no pool slot (pool exhausted)
constexpr uint32_t kPendingSlot = 0xFFFFFFFEU;  // key inserted this frame, slot not yet assigned
constexpr int kBlockVoxels = 512;

__host__ __device__ inline uint32_t intra_of(int x, int y, int z) {
  return static_cast<uint32_t>((((z & 7) * 8) + (y & 7)) * 8 + (x & 7));
}

struct HashView {
  unsigned long long* table;  // (key << 32) | slot, kHashEmpty when free
  uint32_t mask;              // buckets - 1 (buckets a power of two)
  int bx, by, bz;             // blocks per axis
  __host__ __device__ inline uint32_t bucket(uint32_t key) const {
    return static_cast<uint32_t>((static_cast<unsigned long long>(key) * 0x9E3779B97F4A7C15ULL) >> 32) & mask;
  }
  __host__ __device__ inline uint32_t key_of(int x, int y, int z) const {
    return static_cast<uint32_t>(((z >> 3) * by + (y >> 3)) * bx + (x >> 3));
  }
  // Pool slot of a block, or kNoSlot when the block holds no voxels.
  __device__ inline uint32_t find(uint32_t key) const {
    uint32_t h = bucket(key);
    for (uint32_t probe = 0; probe <= mask; ++probe) {
      const unsigned long long e = table[h];
      if (e == kHashEmpty) return kNoSlot;
      if (static_cast<uint32_t>(e >> 32) == key) {
        const uint32_t slot = static_cast<uint32_t>(e);
        return slot == kPendingSlot ? kNoSlot : slot;
      }
      h = (h + 1) & mask;
    }
    return kNoSlot;
  }
};

// ---------------------------------------------------------------------------
// ICP device state, one per context (device memory).
// ---------------------------------------------------------------------------
struct IcpState {
  double pose[16];        // running camera_to_world
  double model_pose[16];  // render pose of the current level's model maps
  float model_inv[12];    // model_pose^-1 (3 x 4, row-major) in float, set with model_pose
  double best_pose[16];
  double eq[29];
  double best_eq[29];
  double best_cost;
  double lambda;
  double update_t, update_r;
  double min_eig, cond;
  long long eq_count, best_count;
  int status;     // per level; see kf_icp_result
  int trials, rejected;
  int frame_failed;
  int failed_level;
  // one-launch levels whose grid-wide wait timed out, since kf_create (never
  // cleared; the host reports each new one once)
  int sync_timeouts;
  float sums32[30];  // last compensated FP32 reduction (29 sums + count)
  double level_pose0[16];  // pose at the start of the current level
  kf_icp_result results[kMaxLevels];
  double prev_pose[16];   // camera_to_world of the previous two frames
  double prev2_pose[16];
  long long frames_done;
  // learned warm start: the constant-velocity seed kept for the fallback,
  // the network's proposal, and how the proposal fared
  double seed_pose[16];
  double proposal[16];
  double proposal_err_t, proposal_err_r, learned_rms;
  int used_learned;
  int pad2;
  // grid-wide protocol of the one-launch level: [0] ticket (zero between
  // trials, and reset when a level begins, after a timed-out launch left it
  // short), [1] epoch (only ever advanced)
  unsigned int level_sync[2];
};

// Prepare a tracking level (one thread): the level's result reads "not run",
// the trial state is cleared and the model pose is the running pose.  In a
// failed frame nothing runs; a frame whose learned run was kept (marked
// failed so that the fallback pass exits at once) keeps its results.
// Prepare a level (one thread): result slot, status and counters cleared,
// model pose = running pose.  The grid-wide ticket is left alone.
__device__ inline void icp_begin_level_state(IcpState* st, int level) {
  if (st->frame_failed != 0) {
    if (st->used_learned == 0) {
      st->results[level] = kf_icp_result{};
      st->results[level].status = -1;
    }
    return;
  }
  st->results[level] = kf_icp_result{};
  st->results[level].status = -1;
  st->status = 0;
  st->trials = 0;
  st->rejected = 0;
  st->best_cost = 0.0;
  st->lambda = 0.0;
  st->eq_count = 0;
  st->best_count = 0;
  st->update_t = 0.0;
  st->update_r = 0.0;
  for (int i = 0; i < 16; ++i) {
    st->model_pose[i] = st->pose[i];
    st->level_pose0[i] = st->pose[i];
    st->best_pose[i] = st->pose[i];
  }
  double inv[16];
  rigid_inverse64(st->model_pose, inv);
  for (int i = 0; i < 12; ++i) st->model_inv[i] = static_cast<float>(inv[i]);
}

// The same, also resetting the grid-wide ticket (a timed-out one-launch
// level may have left it short); runs before any workgroup of the level's
// one-launch form can take a ticket.
__device__ inline void icp_begin_level(IcpState* st, int level) {
  st->level_sync[0] = 0U;
  icp_begin_level_state(st, level);
}

// The learned warm start's check after its run (one thread): keep the run
// when no level failed, the finest level's point-to-plane RMS is at most
// max_rms and it converged or its last increment is within (max_t, max_r),
// marking the frame failed so that the fallback pass exits at once
// (the frame's end clears the mark); otherwise restore the constant-velocity
// seed for the fallback pass.
__device__ inline void learned_check(IcpState* st, int finest, float max_rms, float max_t, float max_r) {
  const kf_icp_result& r = st->results[finest];
  const double rms = r.mean_squared_residual >= 0.0 ? sqrt(r.mean_squared_residual) : 1e300;
  st->learned_rms = rms;
  // a run that ended at its iteration budget still moving more than the
  // increment thresholds needed more iterations than its schedule gives
  const bool settled = r.status == 1 || ((max_t <= 0.0F || r.update_translation <= static_cast<double>(max_t)) &&
                                         (max_r <= 0.0F || r.update_rotation <= static_cast<double>(max_r)));
  if (st->frame_failed == 0 && (r.status == 0 || r.status == 1) && settled && rms <= static_cast<double>(max_rms)) {
    st->used_learned = 1;
    st->frame_failed = 1;
  } else {
    st->used_learned = 0;
    st->frame_failed = 0;
    st->failed_level = -1;
    for (int i = 0; i < 16; ++i) st->pose[i] = st->seed_pose[i];
  }
}

// The learned warm start's end of a frame (one thread): a kept run's frame
// mark is cleared, and the proposal's distance from the final pose recorded.
__device__ inline void learned_end(IcpState* st) {
  if (st->used_learned != 0) {
    st->frame_failed = 0;
    st->failed_level = -1;
  }
  // the proposal's distance from the final pose: E = proposal^-1 pose
  double inv[16];
  rigid_inverse64(st->proposal, inv);
  double E[16];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      double s = 0.0;
      for (int q = 0; q < 4; ++q) s += inv[r * 4 + q] * st->pose[q * 4 + c];
      E[r * 4 + c] = s;
    }
  st->proposal_err_t = sqrt(E[3] * E[3] + E[7] * E[7] + E[11] * E[11]);
  const double ct = fmin(1.0, fmax(-1.0, 0.5 * (E[0] + E[5] + E[10] - 1.0)));
  st->proposal_err_r = acos(ct);
}

struct DeviceIcpParams {
  long long min_corr;
  float max_dist, min_dot;
  float min_t, min_r, max_t, max_r;
  float min_eig, max_cond;
  int damping_mode;
  float lambda0;
  int adaptive;
  float inc, dec, lmin, lmax;
  // the learned warm start's check: largest last increment of a kept run
  // (0: no increment test)
  float learned_max_t = 0.0F, learned_max_r = 0.0F;
};

struct LevelBuffers {
  int w = 0, h = 0;
  Intr k{};
  uint16_t* depth = nullptr;  // level depth (level 0: filtered)
  float4* vert = nullptr;
  float4* norm = nullptr;
  float4* model_pts = nullptr;  // raycast output for this level
  float4* model_nrm = nullptr;
  uint32_t* model_col = nullptr;
};

}  // namespace kf

struct kf_context {
  kf_config cfg{};
  hipStream_t stream = nullptr;
  int device = 0;
  std::string error;
  kf::Grid grid{};
  size_t n_voxels = 0;
  void* voxels = nullptr;
  float4* colors = nullptr;
  int levels_built = 0;
  kf::LevelBuffers lv[kf::kMaxLevels];
  uint16_t* raw_depth = nullptr;  // staging copy of the frame's raw depth (graph input)
  uint32_t* raw_color = nullptr;  // staging copy of the frame's colour (graph input)
  // captured per-frame work (tracking frames); rebuilt when its key changes
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  kf_icp_options graph_opt{};
  bool graph_color = false;
  // integration scratch
  float4* pix = nullptr;          // per pixel: depth m, lambda, weight, valid
  float* tile_max = nullptr;      // per 16x16 tile max usable depth
  uint32_t* bricks = nullptr;     // active brick list
  // dense band integration: blocks the band marked (cleared as they are
  // listed) and their list
  uint32_t* band_bits = nullptr;
  uint32_t* band_list = nullptr;
  // one bit per 8^3 voxel block: any voxel observed / block and its 26
  // neighbours all unobserved (the raycaster skips those)
  uint32_t* obs_bits = nullptr;
  uint32_t* skip_bits = nullptr;
  // free space: 4 bits per block (its y-half x z-half parts), set while every
  // voxel of the part is observed and reads tsdf exactly 1 (set when a pass
  // fuses the whole part to 1, kept while unfused voxels keep their values,
  // cleared when a fused voxel ends elsewhere); a block's free bit = all 4
  // parts set for it and its 26 neighbours (the raycaster steps through those
  // blocks without loads)
  uint32_t* part_bits = nullptr;
  uint32_t* free_bits = nullptr;
  // both classes of block b in one word for the raycaster: bit 2 (b % 16) of
  // word b / 16 is its skip bit, the bit above it its free bit
  uint32_t* class_bits = nullptr;
  // hashed storage: one bit per cell of 4^3 blocks, set when every block of
  // the cell is a skip block (the raycaster steps over such a cell at once)
  uint32_t* coarse_bits = nullptr;
  // hashed storage: per non-skip block, bit (dz + 1) * 9 + (dy + 1) * 3 +
  // (dx + 1) set when the block at that offset holds an observed voxel
  uint32_t* nbr_obs = nullptr;
  size_t part_words = 0;
  size_t block_words = 0;
  uint32_t* counters = nullptr;   // [0] brick count, [1..2] fused counts (u64 pair as 2x u32), spare
  unsigned long long* fused = nullptr;
  // icp scratch
  float* partials = nullptr;      // per-block partial sums (32 floats each)
  float* frame_w2c = nullptr;     // the frame's world_to_camera (12 floats) + skip flag, from the pixel pass
  // bound of the one-launch ICP level's grid-wide wait, in ticks of the
  // 100 MHz wall clock; and the sync_timeouts count already reported
  long long sync_timeout_ticks = 2000000;  // 20 ms
  int seen_timeouts = 0;
  kf::IcpState* icp = nullptr;    // device
  kf::IcpState* icp_host = nullptr;  // pinned mirror for results
  double* trajectory = nullptr;   // device, 16 doubles per frame
  int64_t traj_capacity = 0;
  int64_t frames = 0;
  // timing
  hipEvent_t ev[15] = {};  // 0 start, 1 pyramid, 2.. per level (raycast, icp), 11 end,
                          // 12/13 around the voxel-update kernel, 14 after the proposal
  kf_stage_times last_times{};
  bool have_times = false;
  // hashed volume (KF_STORAGE_HASHED): voxels / colors above are the pool
  unsigned long long* hash_table = nullptr;
  uint32_t* hash_stamp = nullptr;     // frame id of the last visit, per bucket
  uint32_t* block_of_slot = nullptr;  // block key of every pool slot
  uint32_t* work = nullptr;           // pool slots to integrate this frame
  // [0] slots requested (allocated + overflowing), [1] work count,
  // [2] table entries, [3] frame id, [4] overflowing requests
  uint32_t* sparse = nullptr;
  size_t buckets = 0;
  size_t capacity = 0;
  // learned warm start: attached network, previous coarsest-level depth,
  // the proposal's twist
  kf_posenet* posenet = nullptr;
  uint16_t* prev_coarse = nullptr;
  float* twist = nullptr;
  bool timed_learned = false;
};

namespace kf {

// Error plumbing: every ABI call runs inside guard(); HIP failures throw a
// local exception type that the guard turns into a status and a message.
struct Failure {
  kf_status status;
  std::string message;
};

[[noreturn]] void fail(kf_status status, const std::string& message);
void hip_check(hipError_t err, const char* what);
#define KF_HIP(expr) ::kf::hip_check((expr), #expr)

// Launch helpers implemented per stage.
// Staging copies of a frame's raw depth and colour (colour optional).
struct StageCopy {
  uint16_t* depth = nullptr;
  const uint32_t* color_in = nullptr;
  uint32_t* color = nullptr;
};
bool bilateral_active(const kf_config& c);
void launch_bilateral(kf_context* ctx, const uint16_t* raw, StageCopy stage = {});
void launch_level_maps(kf_context* ctx, int level);
// All three levels' maps after the bilateral pass in one launch (3 levels, a
// camera size divisible by 4); false when it does not apply.
bool launch_pyramid3(kf_context* ctx);
// end_learned: the pixel pass also ends the learned warm start's frame.
void launch_integrate(kf_context* ctx, const kf_depth_frame& frame, bool count, bool from_state,
                      bool end_frame = false, bool accumulate = false, bool end_learned = false);
// begin_level >= 0: the launch also prepares that tracking level
// (icp_begin_level, by one thread) -- one launch fewer per level.
void launch_raycast(kf_context* ctx, int level, const double* pose_dev_or_null, const Rigid* pose_host,
                    bool skip_if_failed, int begin_level = -1);
// single_launch: one kernel for the whole level (its workgroups run the
// trials in step, with device-scope fences and a grid-wide wait between
// trials) -- for passes that rarely run and whose launches are mostly early
// exits (the learned warm start's fallback); otherwise two kernels per
// trial, the last trial's second kernel also finishing the level.
void launch_icp_level(kf_context* ctx, const float4* lv_v, const float4* lv_n, int lw, int lh,
                      const float4* md_v, const float4* md_n, int mw, int mh, Intr k,
                      int iterations, const DeviceIcpParams& p, int level_index, bool single_launch = false,
                      float check_max_rms = 0.0F);
void launch_icp_reduce_once(kf_context* ctx, const float4* lv_v, const float4* lv_n, int lw, int lh,
                            const float4* md_v, const float4* md_n, int mw, int mh, Intr k,
                            const float* camera12, const float* model12, float max_dist,
                            float min_dot);
// kf_icp_reduce without a context: the partial rows in the library's static
// per-device array; sums (29 FP64) and count (device) written by the device.
void launch_icp_reduce_static(hipStream_t stream, const float4* lv_v, const float4* lv_n, int lw, int lh,
                              const float4* md_v, const float4* md_n, int mw, int mh, Intr k,
                              const float* camera12, const float* model12, float max_dist, float min_dot,
                              double* sums, int64_t* count);
void launch_icp_begin_level(kf_context* ctx, int level_index);
// The learned warm start's fallback pass (every level: raycast, then the
// trials) as one launch that returns at once when the learned run was kept;
// false when its grid cannot be resident at once on this device (the caller
// then launches each level on its own).
void launch_fill_voxels(kf_context* ctx);
// skip and free bits from the observed and part bits; end_frame: the launch
// also records the frame's pose (trajectory slot, previous poses, counter)
void launch_skip_map(kf_context* ctx, bool end_frame = false);
void launch_rebuild_observed(kf_context* ctx);  // observed bits from the voxels

bool hashed(const kf_context* ctx);
HashView hash_view(const kf_context* ctx);
void launch_fill_pool(kf_context* ctx);  // the table empty, the counters zero
// the dense expansion of a hashed volume into device buffers of the grid's size
void launch_expand_hashed(kf_context* ctx, void* voxels, float4* colors);

// Block grid of the occupancy bitmaps: 8^3 voxels per block, x fastest.
struct BlockMap {
  int bx, by, bz;  // blocks per axis
  __host__ __device__ inline int flat(int x, int y, int z) const { return (z * by + y) * bx + x; }
  __device__ inline bool bit(const uint32_t* bits, int f) const { return (bits[f >> 5] >> (f & 31)) & 1U; }
};
inline BlockMap block_map(const Grid& g) {
  return BlockMap{(g.nx + 7) / 8, (g.ny + 7) / 8, (g.nz + 7) / 8};
}
// cells of 4^3 blocks (32^3 voxels)
inline BlockMap coarse_map(const Grid& g) {
  const BlockMap b = block_map(g);
  return BlockMap{(b.bx + 3) / 4, (b.by + 3) / 4, (b.bz + 3) / 4};
}
DeviceIcpParams to_device_params(const kf_icp_options& o);

// Pose network (kf_posenet.hip): proposal for one depth pair on stream s.
void posenet_propose(kf_posenet* net, hipStream_t s, const uint16_t* prev, const uint16_t* cur, float* twist);
bool posenet_input_matches(int width, int height);
// Before a context uses the network: its depth scale must be the context's,
// and no call on the network's stream may still run.
void posenet_prepare_attach(kf_posenet* net, float depth_scale);

}  // namespace kf

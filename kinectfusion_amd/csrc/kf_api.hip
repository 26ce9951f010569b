// C ABI of libkf_hip (include/kf_hip.h): context lifecycle, buffers, and the
// per-frame orchestration that keeps the running pose on the device.
//
// Reference behaviour mirrored here:
//  * Pipeline::create / validation (src/kinectfusion/pipeline.cpp:129-147,
//    volume.hpp:499-508, depth_processing.cpp:74-91) -- but with no host
//    fallback: a missing gfx950 device is an error;
//  * Reconstruction::initialize / process_frame / track_pose
//    (src/app/reconstruction.cpp:143-250): the first frame integrates at the
//    identity, later frames track coarse-to-fine and integrate only when
//    every level succeeded; a failing level leaves its initial pose.
#include <algorithm>
#include <cstring>
#include <vector>
#include <new>
#include <string>

#include "kf_internal.hpp"

namespace kf {

[[noreturn]] void fail(kf_status status, const std::string& message) {
  throw Failure{status, message};
}

void hip_check(hipError_t err, const char* what) {
  if (err != hipSuccess) {
    const kf_status s = err == hipErrorOutOfMemory ? KF_ERR_OUT_OF_MEMORY : KF_ERR_DEVICE;
    fail(s, std::string(what) + ": " + hipGetErrorString(err));
  }
}

namespace {

thread_local std::string g_create_error;

template <typename F>
kf_status guard(kf_context* ctx, F&& body) {
  try {
    body();
    return KF_OK;
  } catch (const Failure& f) {
    if (ctx != nullptr) ctx->error = f.message;
    else g_create_error = f.message;
    return f.status;
  } catch (const std::bad_alloc&) {
    if (ctx != nullptr) ctx->error = "host allocation failed";
    return KF_ERR_OUT_OF_MEMORY;
  } catch (...) {
    if (ctx != nullptr) ctx->error = "unexpected exception";
    return KF_ERR_DEVICE;
  }
}

void require(bool ok, const char* msg) {
  if (!ok) fail(KF_ERR_INVALID_ARGUMENT, msg);
}

template <typename T>
T* dalloc(size_t count) {
  void* p = nullptr;
  KF_HIP(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
  return static_cast<T*>(p);
}

void dfree(void* p) {
  if (p != nullptr) (void)hipFree(p);
}

// A tracked frame's start (one thread): the frame marks are cleared and the
// pose seeded from the previous pose, or by constant velocity.
__device__ void frame_begin(IcpState* st, int init_mode) {
  st->frame_failed = 0;
  st->failed_level = -1;
  st->used_learned = 0;
  if (init_mode == KF_INIT_CONSTANT_VELOCITY && st->frames_done >= 2) {
    // T_k = T_{k-1} (T_{k-2}^-1 T_{k-1}) in FP64
    double inv2[16], rel[16], out[16];
    rigid_inverse64(st->prev2_pose, inv2);
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) {
        double s = 0.0;
        for (int q = 0; q < 4; ++q) s += inv2[r * 4 + q] * st->prev_pose[q * 4 + c];
        rel[r * 4 + c] = s;
      }
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) {
        double s = 0.0;
        for (int q = 0; q < 4; ++q) s += st->prev_pose[r * 4 + q] * rel[q * 4 + c];
        out[r * 4 + c] = s;
      }
    for (int i = 0; i < 16; ++i) st->pose[i] = out[i];
  } else {
    for (int i = 0; i < 16; ++i) st->pose[i] = st->prev_pose[i];
  }
}

__global__ void frame_begin_kernel(IcpState* st, int init_mode) {
  if (threadIdx.x == 0) frame_begin(st, init_mode);
}


// The staging copies of a frame's inputs for the captured graph, both in one
// launch (four pixels per thread).
__global__ void __launch_bounds__(256) stage_inputs_kernel(const uint16_t* __restrict__ depth,
                                                            uint16_t* __restrict__ depth_out,
                                                            const uint32_t* __restrict__ color,
                                                            uint32_t* __restrict__ color_out, size_t npx) {
  const size_t i0 = (static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x) * 4;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const size_t i = i0 + k;
    if (i >= npx) break;
    depth_out[i] = depth[i];
    if (color != nullptr) color_out[i] = color[i];
  }
}

__global__ void set_identity_kernel(IcpState* st) {
  if (threadIdx.x != 0) return;
  for (int i = 0; i < 16; ++i) {
    const double v = (i % 5 == 0) ? 1.0 : 0.0;
    st->pose[i] = v;
    st->prev_pose[i] = v;
    st->prev2_pose[i] = v;
  }
  st->frame_failed = 0;
  st->failed_level = -1;
  st->frames_done = 0;
  st->level_sync[0] = 0U;
  for (int l = 0; l < kMaxLevels; ++l) {
    st->results[l] = kf_icp_result{};
    st->results[l].status = -1;
  }
}

// FP64 rotation of the twist omega (Rodrigues), row-major 3x3.
__device__ void rodrigues64(const double w[3], double R[9]) {
  const double th = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  double a, b;
  if (th < 1e-12) {
    a = 1.0;
    b = 0.5;
  } else {
    a = sin(th) / th;
    b = (1.0 - cos(th)) / (th * th);
  }
  const double K[9] = {0.0, -w[2], w[1], w[2], 0.0, -w[0], -w[1], w[0], 0.0};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double k2 = 0.0;
      for (int q = 0; q < 3; ++q) k2 += K[r * 3 + q] * K[q * 3 + c];
      R[r * 3 + c] = (r == c ? 1.0 : 0.0) + a * K[r * 3 + c] + b * k2;
    }
}

// The learned seed, at the frame's start: the constant-velocity seed is kept
// for the fallback, then pose = prev_pose [R(omega) | t].
__global__ void learned_seed_kernel(IcpState* st, const float* twist) {
  if (threadIdx.x != 0) return;
  // the grid-wide ticket of the fallback's one-launch levels starts the
  // frame at zero (a timed-out launch may have left it short)
  st->level_sync[0] = 0U;
  frame_begin(st, KF_INIT_CONSTANT_VELOCITY);
  const double w[3] = {twist[0], twist[1], twist[2]};
  double R[9];
  rodrigues64(w, R);
  double D[16] = {R[0], R[1], R[2], twist[3], R[3], R[4], R[5], twist[4], R[6], R[7], R[8], twist[5],
                  0.0,  0.0,  0.0,  1.0};
  for (int i = 0; i < 16; ++i) st->seed_pose[i] = st->pose[i];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      double s = 0.0;
      for (int q = 0; q < 4; ++q) s += st->prev_pose[r * 4 + q] * D[q * 4 + c];
      st->proposal[r * 4 + c] = s;
      st->pose[r * 4 + c] = s;
    }
  for (int l = 0; l < kMaxLevels; ++l) {
    st->results[l] = kf_icp_result{};
    st->results[l].status = -1;
  }
  st->used_learned = 0;
  st->learned_rms = -1.0;
}

template <int Kind>
__global__ void count_observed_kernel(const typename VoxelIO<Kind>::Raw* v, size_t n,
                                      unsigned long long* out) {
  unsigned long long c = 0;
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    float t, w;
    bool obs, fin;
    VoxelIO<Kind>::unpack(v[i], t, w, obs, fin);
    if (w > 0.0F) ++c;
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, c);
}

__global__ void zero_u64_kernel(unsigned long long* p) {
  if (threadIdx.x == 0) *p = 0ULL;
}

void validate(const kf_config& c) {
  // VolumeGeometry (volume.hpp:499-508)
  require(c.resolution[0] > 1 && c.resolution[1] > 1 && c.resolution[2] > 1,
          "Volume resolution must be at least 2 along each axis");
  require(c.resolution[0] <= 4096 && c.resolution[1] <= 4096 && c.resolution[2] <= 4096,
          "Volume resolution above 4096 per axis is not supported");
  require(c.voxel_size > 0.0F, "Voxel size must be positive");
  require(c.truncation_distance > 0.0F, "Truncation distance must be positive");
  require(std::isfinite(c.origin[0]) && std::isfinite(c.origin[1]) && std::isfinite(c.origin[2]),
          "Volume origin must be finite");
  require(c.voxel_store >= 0 && c.voxel_store <= 2, "Unknown voxel store");
  require(c.color_store == KF_COLOR_FLOAT || c.color_store == KF_COLOR_NONE, "Unknown color store");
  require(c.storage == KF_STORAGE_DENSE || c.storage == KF_STORAGE_HASHED, "Unknown storage layout");
  if (c.storage == KF_STORAGE_HASHED) {
    // the hash holds every allocated block plus at most one frame's new
    // requests once the pool is full, at a load of at most 7/8
    const int64_t b = c.hash_buckets;
    require(b >= 1024 && b <= (int64_t{1} << 30) && (b & (b - 1)) == 0,
            "Hash buckets must be a power of two in [2^10, 2^30]");
    require(c.block_capacity >= 1 && c.block_capacity <= (int64_t{1} << 26), "Block capacity must be in [1, 2^26]");
    require(2 * c.block_capacity <= b, "Hash buckets must be at least twice the block capacity");
  }
  require(c.tsdf_rule == KF_RULE_CLASSIC || c.tsdf_rule == KF_RULE_ANGLE_WEIGHTED, "Unknown TSDF rule");
  require(c.integration_mode == KF_INTEGRATION_FULL || c.integration_mode == KF_INTEGRATION_BAND,
          "Unknown integration mode");
  // Pipeline::require_valid_storage (src/kinectfusion/pipeline.cpp:108-117)
  require(c.storage != KF_STORAGE_HASHED || c.integration_mode == KF_INTEGRATION_BAND,
          "Sparse storage implies band integration (nothing is allocated far from surfaces); set "
          "integration_mode = KF_INTEGRATION_BAND");
  // TsdfIntegrator::validated (tsdf_integration.cpp) and depth options
  require(c.depth_scale > 0.0F, "Depth scale must be positive");
  require(c.min_depth >= 0.0F && c.max_depth > c.min_depth, "Depth range is invalid");
  require(c.max_weight > 0.0F, "Max weight must be positive");
  require(c.levels >= 1 && c.levels <= kMaxLevels, "Pyramid levels must be in [1, 4]");
  require(c.max_normal_depth_jump >= 0.0F, "Normal depth jump threshold must be non-negative");
  require(c.max_downsample_depth_jump >= 0.0F, "Downsample depth jump threshold must be non-negative");
  require(c.bilateral_radius >= 0 && c.bilateral_radius <= 8, "Bilateral radius must be in [0, 8]");
  require(c.bilateral_spatial_sigma > 0.0F, "Bilateral spatial sigma must be positive");
  require(c.bilateral_depth_sigma > 0.0F, "Bilateral depth sigma must be positive");
  // RaycastOptions (raycasting.cpp:15-29)
  require(c.ray_min_depth >= 0.0F && c.ray_max_depth > c.ray_min_depth, "Raycast depth range is invalid");
  require(c.step_scale > 0.0F, "Raycast step scale must be positive");
  require((c.ray_max_depth - c.ray_min_depth) / (c.voxel_size * c.step_scale) <= 1.0e6F,
          "Raycast march would exceed 1e6 steps per ray");
  require(c.width > 0 && c.height > 0 && c.width <= 8192 && c.height <= 8192, "Camera size is invalid");
  require(c.fx > 0.0F && c.fy > 0.0F, "Focal lengths must be positive");
  require(c.init_mode >= KF_INIT_PREVIOUS && c.init_mode <= KF_INIT_LEARNED, "Unknown init mode");
  {
    // 0 iterations skip a level (and its raycast); at least one level runs
    int running = 0;
    for (int l = 0; l < c.levels; ++l) {
      require(c.icp_iterations[l] >= 0, "ICP iterations must be >= 0");
      running += c.icp_iterations[l] > 0 ? 1 : 0;
    }
    require(running > 0, "ICP must run at least one level");
  }
  if (c.init_mode == KF_INIT_LEARNED) {
    int any = 0;
    for (int l = 0; l < c.levels; ++l) {
      require(c.learned_iterations[l] >= 0, "Learned-seed iterations must be >= 0");
      any += c.learned_iterations[l] > 0 ? 1 : 0;
    }
    require(any > 0, "Learned-seed iterations must run at least one level");
    require(c.learned_max_rms > 0.0F, "Learned-seed RMS threshold must be positive");
    require(c.learned_max_update_translation >= 0.0F && c.learned_max_update_rotation >= 0.0F,
            "Learned-seed increment thresholds must be >= 0");
  }
}

// PipelineConfig::sparse_block_capacity 0 = a quarter of the grid's blocks
// (pipeline.hpp:40); hash_buckets 0 = the smallest power of two that is at
// least twice the capacity (and at least 2^10).
void resolve_hashed_sizes(kf_config& c) {
  if (c.storage != KF_STORAGE_HASHED) return;
  if (c.block_capacity == 0) {
    const int64_t nb = ((c.resolution[0] + 7) / 8) * ((c.resolution[1] + 7) / 8) * ((c.resolution[2] + 7) / 8);
    c.block_capacity = std::max<int64_t>(1, nb / 4);
  }
  if (c.hash_buckets == 0) {
    int64_t b = 1024;
    while (b < 2 * c.block_capacity && b < (int64_t{1} << 30)) b *= 2;
    c.hash_buckets = b;
  }
}

// Raycast maps start zero-filled: a zero point is "no previous surface" for
// RaycastOptions::seed_from_previous (raycasting.hpp:122-131).
void clear_maps(kf_context* ctx) {
  for (int l = 0; l < ctx->cfg.levels; ++l) {
    const LevelBuffers& L = ctx->lv[l];
    const size_t n = static_cast<size_t>(L.w) * L.h;
    KF_HIP(hipMemsetAsync(L.model_pts, 0, n * sizeof(float4), ctx->stream));
    KF_HIP(hipMemsetAsync(L.model_nrm, 0, n * sizeof(float4), ctx->stream));
    KF_HIP(hipMemsetAsync(L.model_col, 0, n * sizeof(uint32_t), ctx->stream));
  }
}

void upload_pose(kf_context* ctx, const double* pose) {
  KF_HIP(hipMemcpyAsync(ctx->icp->pose, pose, 16 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
}

void fetch_state(kf_context* ctx) {
  KF_HIP(hipMemcpyAsync(ctx->icp_host, ctx->icp, sizeof(IcpState), hipMemcpyDeviceToHost, ctx->stream));
  KF_HIP(hipStreamSynchronize(ctx->stream));
}

// A one-launch ICP level whose grid-wide wait timed out has marked its frame
// failed (not integrated, not reported tracked) and counted the event; the
// first synchronising call after it reports it, once, as a device error.
// Needs icp_host->sync_timeouts fresh from the device.
void report_sync_timeouts(kf_context* ctx) {
  const int n = ctx->icp_host->sync_timeouts;
  if (n == ctx->seen_timeouts) return;
  ctx->seen_timeouts = n;
  fail(KF_ERR_DEVICE,
       "ICP: a one-launch tracking level timed out waiting for its trials (is another context tracking on "
       "this device at the same time?); the frame was not integrated");
}

void fetch_state_checked(kf_context* ctx) {
  fetch_state(ctx);
  report_sync_timeouts(ctx);
}

// filtered: level 0's depth already holds the filtered frame
void build_pyramid(kf_context* ctx, const uint16_t* depth, bool filtered = false) {
  if (!filtered) launch_bilateral(ctx, depth);
  if (!launch_pyramid3(ctx))
    for (int l = 0; l < ctx->cfg.levels; ++l) launch_level_maps(ctx, l);
  ctx->levels_built = ctx->cfg.levels;
}

kf_depth_frame frame_from_pyramid(kf_context* ctx, const uint16_t* depth, const uint32_t* color) {
  kf_depth_frame f{};
  f.depth = depth;
  f.color = color;
  f.normals = reinterpret_cast<const float*>(ctx->lv[0].norm);
  f.width = ctx->cfg.width;
  f.height = ctx->cfg.height;
  f.fx = ctx->cfg.fx;
  f.fy = ctx->cfg.fy;
  f.cx = ctx->cfg.cx;
  f.cy = ctx->cfg.cy;
  const float id[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  std::memcpy(f.world_to_camera, id, sizeof(id));
  return f;
}

void record(kf_context* ctx, int slot) {
  if (ctx->cfg.stage_timing) KF_HIP(hipEventRecord(ctx->ev[slot], ctx->stream));
}

// Coarse-to-fine tracking from the device pose (track_pose).
// Coarse-to-fine tracking with per-level iteration counts (0 skips a level
// and its raycast).
// check_max_rms > 0: the finest level run also makes the learned warm
// start's check.
void track_levels(kf_context* ctx, const DeviceIcpParams& p, bool timing, const int32_t* iterations,
                  bool single_launch = false, float check_max_rms = 0.0F) {
  const kf_config& c = ctx->cfg;
  int slot = 2;
  int finest = 0;
  while (finest + 1 < c.levels && iterations[finest] == 0) ++finest;
  for (int l = c.levels - 1; l >= 0; --l) {
    LevelBuffers& L = ctx->lv[l];
    if (iterations[l] > 0) launch_raycast(ctx, l, nullptr, nullptr, true, l);  // + begin the level
    if (timing) record(ctx, slot++);
    if (iterations[l] > 0) {
      launch_icp_level(ctx, L.vert, L.norm, L.w, L.h, L.model_pts, L.model_nrm, L.w, L.h, L.k, iterations[l], p,
                       l, single_launch, l == finest ? check_max_rms : 0.0F);
    }
    if (timing) record(ctx, slot++);
  }
}

bool learned_active(const kf_context* ctx) {
  return ctx->cfg.init_mode == KF_INIT_LEARNED && ctx->posenet != nullptr;
}

void drop_graph(kf_context* ctx) {
  if (ctx->graph_exec != nullptr) (void)hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->graph != nullptr) (void)hipGraphDestroy(ctx->graph);
  ctx->graph_exec = nullptr;
  ctx->graph = nullptr;
}

// Reconstruction::initialize / process_frame on the device, enqueued on the
// context stream (eagerly, or into a stream capture): pyramid, the pose seed,
// coarse-to-fine tracking (not on the first frame), integration, and the
// trajectory record.  No call here waits on the device.
void enqueue_frame(kf_context* ctx, const uint16_t* depth, const uint32_t* color, const kf_icp_options& opt,
                   bool first, bool filtered = false) {
  const bool timing = ctx->cfg.stage_timing != 0;
  record(ctx, 0);
  build_pyramid(ctx, depth, filtered);
  record(ctx, 1);
  const kf_depth_frame frame = frame_from_pyramid(ctx, depth, color);
  if (first) {
    // initialize(): integrate the first frame at the identity
    hipLaunchKernelGGL(frame_begin_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->icp,
                       static_cast<int>(KF_INIT_PREVIOUS));
    for (int s = 2; s < 2 + 2 * ctx->cfg.levels && timing; ++s) record(ctx, s);
  } else if (learned_active(ctx)) {
    // learned warm start: propose, track from the proposal, fall back to the
    // constant-velocity seed with the full schedule when it does not hold
    const kf_config& c = ctx->cfg;
    posenet_propose(ctx->posenet, ctx->stream, ctx->prev_coarse, ctx->lv[c.levels - 1].depth, ctx->twist);
    hipLaunchKernelGGL(learned_seed_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->icp, ctx->twist);
    record(ctx, 14);
    DeviceIcpParams p = to_device_params(opt);
    p.learned_max_t = c.learned_max_update_translation;
    p.learned_max_r = c.learned_max_update_rotation;
    track_levels(ctx, p, timing, c.learned_iterations, false, c.learned_max_rms);  // + the check
    // the fallback runs only when the check failed (every launch of a kept
    // frame returns at once): per level, its raycast, then one launch that
    // runs the level's trials in step
    track_levels(ctx, p, false, c.icp_iterations, true);
    record(ctx, 2 * c.levels + 1);  // the fallback run counts as tracking time
  } else {
    hipLaunchKernelGGL(frame_begin_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->icp,
                       ctx->cfg.init_mode == KF_INIT_CONSTANT_VELOCITY ? KF_INIT_CONSTANT_VELOCITY
                                                                     : KF_INIT_PREVIOUS);
    track_levels(ctx, to_device_params(opt), timing, ctx->cfg.icp_iterations);
  }
  ctx->timed_learned = timing && !first && learned_active(ctx);
  // + the learned run's end (in the pixel pass) and the frame's end (in the
  // skip-map pass, which also keeps this frame's coarse depth for the next
  // proposal)
  launch_integrate(ctx, frame, ctx->cfg.count_fused != 0, true, true, true, !first && learned_active(ctx));
  KF_HIP(hipGetLastError());
}

void copy_results(const kf_context* ctx, kf_icp_result* results) {
  for (int l = 0; l < ctx->cfg.levels; ++l) results[l] = ctx->icp_host->results[l];
}

}  // namespace
}  // namespace kf

using namespace kf;

extern "C" {

void kf_config_defaults(kf_config* c) {
  std::memset(c, 0, sizeof(*c));
  // src/app/app_options.hpp:22-28 and the library option defaults
  c->resolution[0] = c->resolution[1] = c->resolution[2] = 512;
  c->voxel_size = 0.01F;
  c->origin[0] = c->origin[1] = -0.5F * 512 * 0.01F;
  c->origin[2] = -2.56F;
  c->truncation_distance = 0.05F;
  c->voxel_store = KF_VOXEL_FLOAT;
  c->color_store = KF_COLOR_FLOAT;
  c->storage = KF_STORAGE_DENSE;
  c->tsdf_rule = KF_RULE_ANGLE_WEIGHTED;
  c->depth_scale = 5000.0F;
  c->observation_weight = 1.0F;
  c->max_weight = 196.0F;
  c->min_depth = 0.4F;
  c->max_depth = 8.0F;
  c->projective_distance = 1;
  c->distance_scaled_truncation = 0;
  c->truncation_distance_scale = 0.01F;
  c->ray_min_depth = 0.4F;
  c->ray_max_depth = 8.0F;
  c->step_scale = 1.0F;
  c->require_all_corners = 0;
  c->cell_normals = 0;
  c->levels = 3;
  c->max_normal_depth_jump = 0.1F;
  c->max_downsample_depth_jump = 0.1F;
  c->bilateral_filter = 1;
  c->bilateral_radius = 2;
  c->bilateral_spatial_sigma = 2.0F;
  c->bilateral_depth_sigma = 0.08F;
  c->width = 640;
  c->height = 480;
  c->fx = 525.0F;
  c->fy = 525.0F;
  c->cx = 319.5F;
  c->cy = 239.5F;
  c->hash_buckets = 1 << 20;
  c->block_capacity = 500000;
  c->init_mode = KF_INIT_PREVIOUS;
  c->icp_iterations[0] = 10;
  c->icp_iterations[1] = 5;
  c->icp_iterations[2] = 4;
  c->icp_iterations[3] = 4;
  c->stage_timing = 0;
  c->count_fused = 0;
  c->learned_iterations[0] = 4;
  c->learned_iterations[1] = 0;
  c->learned_iterations[2] = 0;
  c->learned_iterations[3] = 0;
  c->learned_max_rms = 0.01F;
  c->learned_max_update_translation = 1.0e-3F;
  c->learned_max_update_rotation = 1.0e-3F;
}

void kf_icp_options_defaults(kf_icp_options* o) {
  std::memset(o, 0, sizeof(*o));
  o->min_correspondences = 64;       // icp_optimizer.hpp:14-20
  o->max_point_distance = 0.05F;     // icp_correspondence.hpp:25-34
  o->min_normal_dot = 0.9659258F;
  o->min_update_translation = 1.0e-5F;
  o->min_update_rotation = 1.0e-5F;
  o->max_update_translation = 0.15F;
  o->max_update_rotation = 0.35F;
  o->min_system_eigenvalue = 1.0e-6F;
  o->max_condition_number = 1.0e6F;
  o->damping_mode = KF_DAMPING_NONE;
  o->damping_lambda = 1.0e-3F;
  o->adaptive_damping = 0;
  o->lambda_increase = 10.0F;
  o->lambda_decrease = 0.1F;
  o->min_lambda = 1.0e-9F;
  o->max_lambda = 1.0e6F;
}

kf_status kf_create(const kf_config* cfg, kf_context** out) {
  if (out == nullptr || cfg == nullptr) {
    g_create_error = "kf_create: null argument";
    return KF_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  kf_context* ctx = nullptr;
  const kf_status s = guard(nullptr, [&] {
    kf_config resolved = *cfg;
    resolve_hashed_sizes(resolved);
    validate(resolved);
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
      fail(KF_ERR_UNSUPPORTED, "no HIP device is available; this library has no host fallback");
    int dev = 0;
    KF_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop{};
    KF_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      fail(KF_ERR_UNSUPPORTED, std::string("device is ") + prop.gcnArchName + "; built for gfx950 only");
    ctx = new kf_context();
    ctx->cfg = resolved;
    ctx->device = dev;
    const kf_config& c = ctx->cfg;
    KF_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ctx->grid = Grid{static_cast<int>(c.resolution[0]), static_cast<int>(c.resolution[1]),
                     static_cast<int>(c.resolution[2]), c.voxel_size, c.origin[0], c.origin[1],
                     c.origin[2], c.truncation_distance};
    if (c.storage == KF_STORAGE_HASHED) {
      // BlockRep (block_rep.hpp:183-307): a pool of 8^3-voxel blocks and the
      // hash from block keys to pool slots
      ctx->buckets = static_cast<size_t>(c.hash_buckets);
      ctx->capacity = static_cast<size_t>(c.block_capacity);
      ctx->n_voxels = ctx->capacity * kBlockVoxels;
      ctx->hash_table = dalloc<unsigned long long>(ctx->buckets);
      ctx->hash_stamp = dalloc<uint32_t>(ctx->buckets);
      ctx->block_of_slot = dalloc<uint32_t>(ctx->capacity);
      ctx->work = dalloc<uint32_t>(ctx->capacity);
      ctx->sparse = dalloc<uint32_t>(8);
    } else {
      ctx->n_voxels = static_cast<size_t>(c.resolution[0]) * c.resolution[1] * c.resolution[2];
    }
    ctx->voxels = dalloc<uint8_t>(ctx->n_voxels * voxel_bytes(c.voxel_store));
    if (c.color_store == KF_COLOR_FLOAT) ctx->colors = dalloc<float4>(ctx->n_voxels);
    int w = c.width, h = c.height;
    for (int l = 0; l < c.levels; ++l) {
      LevelBuffers& L = ctx->lv[l];
      L.w = w;
      L.h = h;
      const float s2 = 1.0F / static_cast<float>(1U << l);  // CameraIntrinsics::scaled (rgbd.hpp:66-70)
      L.k = Intr{c.fx * s2, c.fy * s2, c.cx * s2, c.cy * s2};
      const size_t n = static_cast<size_t>(w) * h;
      L.depth = dalloc<uint16_t>(n);
      L.vert = dalloc<float4>(n);
      L.norm = dalloc<float4>(n);
      L.model_pts = dalloc<float4>(n);
      L.model_nrm = dalloc<float4>(n);
      L.model_col = dalloc<uint32_t>(n);
      w /= 2;
      h /= 2;
      require(w > 0 || l + 1 == c.levels, "Too many pyramid levels for the camera size");
    }
    const size_t npx = static_cast<size_t>(c.width) * c.height;
    ctx->pix = dalloc<float4>(npx);
    ctx->raw_depth = dalloc<uint16_t>(npx);
    ctx->raw_color = dalloc<uint32_t>(npx);
    {
      // the depth maximum of each 16 px tile
      const int tx = (c.width + kTile - 1) / kTile, ty = (c.height + kTile - 1) / kTile;
      ctx->tile_max = dalloc<float>(static_cast<size_t>(tx) * ty);
    }
    if (c.storage == KF_STORAGE_DENSE) {
      const size_t nb = static_cast<size_t>((c.resolution[0] + kBrickX - 1) / kBrickX) *
                        ((c.resolution[1] + kBrickY - 1) / kBrickY) *
                        ((c.resolution[2] + kBrickZ - 1) / kBrickZ);
      ctx->bricks = dalloc<uint32_t>(nb);
    }
    {
      const BlockMap bm = block_map(ctx->grid);
      ctx->block_words = (static_cast<size_t>(bm.bx) * bm.by * bm.bz + 31) / 32;
      ctx->obs_bits = dalloc<uint32_t>(ctx->block_words);
      ctx->skip_bits = dalloc<uint32_t>(ctx->block_words);
      KF_HIP(hipMemsetAsync(ctx->obs_bits, 0, ctx->block_words * 4, ctx->stream));
      ctx->part_words = (static_cast<size_t>(bm.bx) * bm.by * bm.bz + 7) / 8;
      ctx->part_bits = dalloc<uint32_t>(ctx->part_words);
      ctx->free_bits = dalloc<uint32_t>(ctx->block_words);
      ctx->class_bits = dalloc<uint32_t>(2 * ctx->block_words);
      KF_HIP(hipMemsetAsync(ctx->part_bits, 0, ctx->part_words * 4, ctx->stream));
      if (c.storage == KF_STORAGE_HASHED) {
        const BlockMap cm = coarse_map(ctx->grid);
        ctx->coarse_bits = dalloc<uint32_t>((static_cast<size_t>(cm.bx) * cm.by * cm.bz + 31) / 32);
        const size_t nb = static_cast<size_t>(bm.bx) * bm.by * bm.bz;
        ctx->nbr_obs = dalloc<uint32_t>(nb);
        KF_HIP(hipMemsetAsync(ctx->nbr_obs, 0, nb * 4, ctx->stream));
      }
      if (c.storage == KF_STORAGE_DENSE && c.integration_mode == KF_INTEGRATION_BAND) {
        ctx->band_bits = dalloc<uint32_t>(ctx->block_words);
        KF_HIP(hipMemsetAsync(ctx->band_bits, 0, ctx->block_words * 4, ctx->stream));
        ctx->band_list = dalloc<uint32_t>(static_cast<size_t>(bm.bx) * bm.by * bm.bz);
      }
    }
    ctx->counters = dalloc<uint32_t>(4);
    ctx->fused = dalloc<unsigned long long>(4);
    KF_HIP(hipMemsetAsync(ctx->fused, 0, 4 * sizeof(unsigned long long), ctx->stream));
    ctx->partials = dalloc<float>(512 * 32);
    ctx->frame_w2c = dalloc<float>(16);
    ctx->icp = dalloc<IcpState>(1);
    KF_HIP(hipMemsetAsync(ctx->icp, 0, sizeof(IcpState), ctx->stream));
    KF_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->icp_host), sizeof(IcpState), hipHostMallocDefault));
    if (c.init_mode == KF_INIT_LEARNED) {
      const LevelBuffers& L = ctx->lv[c.levels - 1];
      ctx->prev_coarse = dalloc<uint16_t>(static_cast<size_t>(L.w) * L.h);
      KF_HIP(hipMemsetAsync(ctx->prev_coarse, 0, sizeof(uint16_t) * L.w * L.h, ctx->stream));
      ctx->twist = dalloc<float>(8);
    }
    ctx->traj_capacity = 1 << 15;  // frames before the first regrowth
    ctx->trajectory = dalloc<double>(ctx->traj_capacity * 16);
    for (auto& e : ctx->ev) KF_HIP(hipEventCreate(&e));
    clear_maps(ctx);
    launch_fill_voxels(ctx);
    if (hashed(ctx)) launch_fill_pool(ctx);
    launch_skip_map(ctx);
    hipLaunchKernelGGL(set_identity_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->icp);
    KF_HIP(hipStreamSynchronize(ctx->stream));
  });
  if (s != KF_OK) {
    if (ctx != nullptr) {
      g_create_error = ctx->error;
      kf_destroy(ctx);
    }
    return s;
  }
  *out = ctx;
  return KF_OK;
}

void kf_destroy(kf_context* ctx) {
  if (ctx == nullptr) return;
  if (ctx->stream != nullptr) (void)hipStreamSynchronize(ctx->stream);
  drop_graph(ctx);
  dfree(ctx->raw_depth);
  dfree(ctx->raw_color);
  dfree(ctx->voxels);
  dfree(ctx->colors);
  for (auto& L : ctx->lv) {
    dfree(L.depth);
    dfree(L.vert);
    dfree(L.norm);
    dfree(L.model_pts);
    dfree(L.model_nrm);
    dfree(L.model_col);
  }
  dfree(ctx->pix);
  dfree(ctx->tile_max);
  dfree(ctx->bricks);
  dfree(ctx->band_bits);
  dfree(ctx->band_list);
  dfree(ctx->obs_bits);
  dfree(ctx->skip_bits);
  dfree(ctx->part_bits);
  dfree(ctx->free_bits);
  dfree(ctx->class_bits);
  dfree(ctx->coarse_bits);
  dfree(ctx->nbr_obs);
  dfree(ctx->counters);
  dfree(ctx->fused);
  dfree(ctx->partials);
  dfree(ctx->frame_w2c);
  dfree(ctx->icp);
  dfree(ctx->trajectory);
  dfree(ctx->prev_coarse);
  dfree(ctx->twist);
  dfree(ctx->hash_table);
  dfree(ctx->hash_stamp);
  dfree(ctx->block_of_slot);
  dfree(ctx->work);
  dfree(ctx->sparse);
  if (ctx->icp_host != nullptr) (void)hipHostFree(ctx->icp_host);
  for (auto& e : ctx->ev)
    if (e != nullptr) (void)hipEventDestroy(e);
  if (ctx->stream != nullptr) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* kf_last_error(const kf_context* ctx) {
  return ctx != nullptr ? ctx->error.c_str() : g_create_error.c_str();
}

kf_status kf_synchronize(kf_context* ctx) {
  return guard(ctx, [&] {
    KF_HIP(hipMemcpyAsync(&ctx->icp_host->sync_timeouts, &ctx->icp->sync_timeouts, sizeof(int),
                          hipMemcpyDeviceToHost, ctx->stream));
    KF_HIP(hipStreamSynchronize(ctx->stream));
    report_sync_timeouts(ctx);
  });
}

kf_status kf_set_sync_timeout(kf_context* ctx, int64_t ticks) {
  return guard(ctx, [&] {
    require(ticks >= 0, "kf_set_sync_timeout: ticks must be >= 0");
    KF_HIP(hipStreamSynchronize(ctx->stream));
    ctx->sync_timeout_ticks = ticks;
    drop_graph(ctx);  // the captured frame holds the bound as a kernel argument
  });
}

kf_status kf_reset(kf_context* ctx) {
  return guard(ctx, [&] {
    launch_fill_voxels(ctx);
    if (hashed(ctx)) launch_fill_pool(ctx);
    KF_HIP(hipMemsetAsync(ctx->obs_bits, 0, ctx->block_words * 4, ctx->stream));
    KF_HIP(hipMemsetAsync(ctx->part_bits, 0, ctx->part_words * 4, ctx->stream));
    launch_skip_map(ctx);
    KF_HIP(hipMemsetAsync(ctx->fused, 0, 4 * sizeof(unsigned long long), ctx->stream));
    clear_maps(ctx);
    hipLaunchKernelGGL(set_identity_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->icp);
    KF_HIP(hipGetLastError());
    ctx->frames = 0;
    ctx->levels_built = 0;
  });
}

kf_status kf_pyramid_build(kf_context* ctx, const uint16_t* depth, int32_t* levels) {
  return guard(ctx, [&] {
    require(depth != nullptr, "kf_pyramid_build: null depth");
    build_pyramid(ctx, depth);
    if (levels != nullptr) *levels = ctx->levels_built;
  });
}

kf_status kf_pyramid_level(kf_context* ctx, int32_t level, kf_level_view* out) {
  return guard(ctx, [&] {
    if (level < 0 || level >= ctx->levels_built) fail(KF_ERR_STATE, "kf_pyramid_level: level not built");
    const LevelBuffers& L = ctx->lv[level];
    out->width = L.w;
    out->height = L.h;
    out->fx = L.k.fx;
    out->fy = L.k.fy;
    out->cx = L.k.cx;
    out->cy = L.k.cy;
    out->depth = L.depth;
    out->vertices = reinterpret_cast<const float*>(L.vert);
    out->normals = reinterpret_cast<const float*>(L.norm);
  });
}

kf_status kf_integrate(kf_context* ctx, const kf_depth_frame* frame, int64_t* counts) {
  return guard(ctx, [&] {
    require(frame != nullptr && frame->depth != nullptr, "kf_integrate: null frame or depth");
    require(frame->width > 0 && frame->height > 0, "kf_integrate: empty frame");
    require(frame->fx > 0.0F && frame->fy > 0.0F, "kf_integrate: invalid intrinsics");
    launch_integrate(ctx, *frame, counts != nullptr, false);
    if (counts != nullptr) {
      unsigned long long host[2];
      KF_HIP(hipMemcpyAsync(host, ctx->fused, sizeof(host), hipMemcpyDeviceToHost, ctx->stream));
      KF_HIP(hipStreamSynchronize(ctx->stream));
      counts[0] = static_cast<int64_t>(host[0]);
      counts[1] = static_cast<int64_t>(host[1]);
    }
  });
}

kf_status kf_raycast(kf_context* ctx, const double* camera_to_world, int32_t level, kf_maps_view* out) {
  return guard(ctx, [&] {
    require(camera_to_world != nullptr, "kf_raycast: null pose");
    require(level >= 0 && level < ctx->cfg.levels, "kf_raycast: level out of range");
    const Rigid pose = rigid_from_pose64(camera_to_world);
    launch_raycast(ctx, level, nullptr, &pose, false);
    const LevelBuffers& L = ctx->lv[level];
    if (out != nullptr) {
      out->width = L.w;
      out->height = L.h;
      out->points = reinterpret_cast<const float*>(L.model_pts);
      out->normals = reinterpret_cast<const float*>(L.model_nrm);
      out->colors = L.model_col;
    }
  });
}

kf_status kf_icp_reduce(kf_context* ctx, const float* lv, const float* ln, int32_t lw, int32_t lh,
                        const float* mv, const float* mn, int32_t mw, int32_t mh, const float* k,
                        const float* camera, const float* model, float max_dist, float min_dot,
                        double* sums, int64_t* count) {
  return guard(ctx, [&] {
    require(lv && ln && mv && mn && k && camera && model && sums && count, "kf_icp_reduce: null argument");
    require(lw > 0 && lh > 0 && mw > 0 && mh > 0, "kf_icp_reduce: empty maps");
    launch_icp_reduce_once(ctx, reinterpret_cast<const float4*>(lv), reinterpret_cast<const float4*>(ln),
                           lw, lh, reinterpret_cast<const float4*>(mv),
                           reinterpret_cast<const float4*>(mn), mw, mh, Intr{k[0], k[1], k[2], k[3]},
                           camera, model, max_dist, min_dot);
    fetch_state_checked(ctx);
    for (int i = 0; i < 29; ++i) sums[i] = static_cast<double>(ctx->icp_host->sums32[i]);
    *count = static_cast<int64_t>(__builtin_bit_cast(uint32_t, ctx->icp_host->sums32[29]));
  });
}

kf_status kf_icp_reduce_static(const float* lv, const float* ln, int32_t lw, int32_t lh, const float* mv,
                               const float* mn, int32_t mw, int32_t mh, const float* k, const float* camera,
                               const float* model, float max_dist, float min_dot, double* sums, int64_t* count,
                               void* stream) {
  return guard(nullptr, [&] {
    require(lv && ln && mv && mn && k && camera && model && sums && count, "kf_icp_reduce_static: null argument");
    require(lw > 0 && lh > 0 && mw > 0 && mh > 0, "kf_icp_reduce_static: empty maps");
    require(static_cast<int64_t>(lw) * lh <= KF_ICP_STATIC_MAX_PIXELS,
            "kf_icp_reduce_static: the live maps exceed KF_ICP_STATIC_MAX_PIXELS pixels");
    require(static_cast<int64_t>(mw) * mh <= KF_ICP_STATIC_MAX_PIXELS,
            "kf_icp_reduce_static: the model maps exceed KF_ICP_STATIC_MAX_PIXELS pixels");
    launch_icp_reduce_static(static_cast<hipStream_t>(stream), reinterpret_cast<const float4*>(lv),
                             reinterpret_cast<const float4*>(ln), lw, lh, reinterpret_cast<const float4*>(mv),
                             reinterpret_cast<const float4*>(mn), mw, mh, Intr{k[0], k[1], k[2], k[3]}, camera,
                             model, max_dist, min_dot, sums, count);
  });
}

kf_status kf_icp_track_maps(kf_context* ctx, const float* lv, const float* ln, int32_t lw, int32_t lh,
                            const float* mv, const float* mn, int32_t mw, int32_t mh, const float* k,
                            const double* model_pose, double* pose_inout, int32_t iterations,
                            const kf_icp_options* opt, kf_icp_result* result) {
  return guard(ctx, [&] {
    require(lv && ln && mv && mn && k && model_pose && pose_inout && opt, "kf_icp_track_maps: null argument");
    require(lw > 0 && lh > 0 && mw > 0 && mh > 0, "kf_icp_track_maps: empty maps");
    require(iterations >= 1, "kf_icp_track_maps: iterations must be >= 1");
    const DeviceIcpParams p = to_device_params(*opt);
    // running pose = model pose for the level setup, then the initial pose
    KF_HIP(hipMemsetAsync(&ctx->icp->frame_failed, 0, sizeof(int), ctx->stream));
    upload_pose(ctx, model_pose);
    launch_icp_begin_level(ctx, 0);
    upload_pose(ctx, pose_inout);
    KF_HIP(hipMemcpyAsync(ctx->icp->level_pose0, pose_inout, 16 * sizeof(double), hipMemcpyHostToDevice,
                          ctx->stream));
    KF_HIP(hipMemcpyAsync(ctx->icp->best_pose, pose_inout, 16 * sizeof(double), hipMemcpyHostToDevice,
                          ctx->stream));
    launch_icp_level(ctx, reinterpret_cast<const float4*>(lv), reinterpret_cast<const float4*>(ln), lw, lh,
                     reinterpret_cast<const float4*>(mv), reinterpret_cast<const float4*>(mn), mw, mh,
                     Intr{k[0], k[1], k[2], k[3]}, iterations, p, 0);
    fetch_state(ctx);
    const kf_icp_result& r = ctx->icp_host->results[0];
    if (result != nullptr) *result = r;
    if (r.status == 0 || r.status == 1) std::memcpy(pose_inout, ctx->icp_host->pose, 16 * sizeof(double));
    // leave the context's tracking state untouched by this standalone call
    KF_HIP(hipMemcpyAsync(ctx->icp->pose, ctx->icp->prev_pose, 16 * sizeof(double),
                          hipMemcpyDeviceToDevice, ctx->stream));
    ctx->icp_host->frame_failed = 0;
    KF_HIP(hipMemsetAsync(&ctx->icp->frame_failed, 0, sizeof(int), ctx->stream));
    report_sync_timeouts(ctx);
  });
}

kf_status kf_track(kf_context* ctx, const kf_icp_options* opt, double* pose_inout, kf_icp_result* results) {
  return guard(ctx, [&] {
    require(opt != nullptr && pose_inout != nullptr, "kf_track: null argument");
    if (ctx->levels_built == 0) fail(KF_ERR_STATE, "kf_track: build the pyramid first");
    const DeviceIcpParams p = to_device_params(*opt);
    upload_pose(ctx, pose_inout);
    KF_HIP(hipMemsetAsync(&ctx->icp->frame_failed, 0, sizeof(int), ctx->stream));
    track_levels(ctx, p, false, ctx->cfg.icp_iterations);
    fetch_state(ctx);
    std::memcpy(pose_inout, ctx->icp_host->pose, 16 * sizeof(double));
    if (results != nullptr) copy_results(ctx, results);
    KF_HIP(hipMemcpyAsync(ctx->icp->pose, ctx->icp->prev_pose, 16 * sizeof(double),
                          hipMemcpyDeviceToDevice, ctx->stream));
    KF_HIP(hipMemsetAsync(&ctx->icp->frame_failed, 0, sizeof(int), ctx->stream));
    report_sync_timeouts(ctx);
  });
}

kf_status kf_process_frame(kf_context* ctx, const uint16_t* depth, const uint32_t* color,
                           const kf_icp_options* opt, kf_frame_result* result) {
  return guard(ctx, [&] {
    require(depth != nullptr && opt != nullptr, "kf_process_frame: null argument");
    if (ctx->frames >= ctx->traj_capacity) {
      double* bigger = dalloc<double>(ctx->traj_capacity * 2 * 16);
      KF_HIP(hipMemcpyAsync(bigger, ctx->trajectory, ctx->traj_capacity * 16 * sizeof(double),
                            hipMemcpyDeviceToDevice, ctx->stream));
      KF_HIP(hipStreamSynchronize(ctx->stream));
      dfree(ctx->trajectory);
      ctx->trajectory = bigger;
      ctx->traj_capacity *= 2;
      drop_graph(ctx);  // the captured frame writes the old trajectory buffer
    }
    const bool timing = ctx->cfg.stage_timing != 0;
    const bool eager = timing || ctx->cfg.count_fused != 0 || ctx->frames == 0;
    if (eager) {
      enqueue_frame(ctx, depth, color, *opt, ctx->frames == 0);
    } else {
      // tracking frames replay one captured graph over staging copies of the
      // inputs; the bilateral filter, launched before it, makes the copies
      // in its own pass
      const bool has_color = color != nullptr;
      const bool filtered = bilateral_active(ctx->cfg);
      if (filtered) {
        launch_bilateral(ctx, depth, StageCopy{ctx->raw_depth, color, ctx->raw_color});
      } else {
        const size_t npx = static_cast<size_t>(ctx->cfg.width) * ctx->cfg.height;
        hipLaunchKernelGGL(stage_inputs_kernel, dim3(static_cast<unsigned>((npx + 1023) / 1024)), dim3(256), 0,
                           ctx->stream, depth, ctx->raw_depth, color, ctx->raw_color, npx);
        KF_HIP(hipGetLastError());
      }
      if (ctx->graph_exec == nullptr || has_color != ctx->graph_color ||
          std::memcmp(&ctx->graph_opt, opt, sizeof(kf_icp_options)) != 0) {
        drop_graph(ctx);
        KF_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        try {
          enqueue_frame(ctx, ctx->raw_depth, has_color ? ctx->raw_color : nullptr, *opt, false, filtered);
        } catch (...) {
          hipGraph_t broken = nullptr;
          (void)hipStreamEndCapture(ctx->stream, &broken);
          if (broken != nullptr) (void)hipGraphDestroy(broken);
          throw;
        }
        KF_HIP(hipStreamEndCapture(ctx->stream, &ctx->graph));
        KF_HIP(hipGraphInstantiate(&ctx->graph_exec, ctx->graph, nullptr, nullptr, 0));
        ctx->graph_opt = *opt;
        ctx->graph_color = has_color;
      }
      KF_HIP(hipGraphLaunch(ctx->graph_exec, ctx->stream));
    }
    if (timing) KF_HIP(hipEventRecord(ctx->ev[11], ctx->stream));
    ctx->frames += 1;
    ctx->have_times = timing;
    if (result != nullptr) {
      fetch_state(ctx);
      std::memset(result, 0, sizeof(*result));
      result->failed_level = ctx->icp_host->frame_failed ? ctx->icp_host->failed_level : -1;
      result->tracked = ctx->icp_host->frame_failed ? 0 : 1;
      std::memcpy(result->pose, ctx->icp_host->pose, sizeof(result->pose));
      if (ctx->frames > 1) copy_results(ctx, result->level);
      if (ctx->frames > 1 && learned_active(ctx)) {
        result->used_learned = ctx->icp_host->used_learned;
        result->proposal_translation_error = ctx->icp_host->proposal_err_t;
        result->proposal_rotation_error = ctx->icp_host->proposal_err_r;
        result->learned_rms = ctx->icp_host->learned_rms;
        std::memcpy(result->proposal_pose, ctx->icp_host->proposal, sizeof(result->proposal_pose));
      }
      report_sync_timeouts(ctx);
    }
  });
}

kf_status kf_trajectory(kf_context* ctx, double* poses, int64_t capacity, int64_t* count) {
  return guard(ctx, [&] {
    const int64_t n = std::min<int64_t>(capacity, ctx->frames);
    if (n > 0)
      KF_HIP(hipMemcpyAsync(poses, ctx->trajectory, n * 16 * sizeof(double), hipMemcpyDeviceToHost,
                            ctx->stream));
    KF_HIP(hipStreamSynchronize(ctx->stream));
    if (count != nullptr) *count = ctx->frames;
  });
}

kf_status kf_last_stage_times(kf_context* ctx, kf_stage_times* out) {
  return guard(ctx, [&] {
    if (!ctx->have_times) fail(KF_ERR_STATE, "stage timing is off or no frame was processed");
    KF_HIP(hipEventSynchronize(ctx->ev[11]));
    auto ms = [&](int a, int b) {
      float t = 0.0F;
      KF_HIP(hipEventElapsedTime(&t, ctx->ev[a], ctx->ev[b]));
      return t;
    };
    kf_stage_times t{};
    t.pyramid_ms = ms(0, 1);
    // events 2.. alternate: after raycast, after icp (per level); the first
    // raycast starts at event 1
    int prev = 1;
    if (ctx->timed_learned) {
      t.proposal_ms = ms(1, 14);
      prev = 14;
    }
    for (int l = 0; l < ctx->cfg.levels; ++l) {
      t.raycast_ms += ms(prev, 2 + 2 * l);
      t.icp_ms += ms(2 + 2 * l, 3 + 2 * l);
      prev = 3 + 2 * l;
    }
    t.integrate_ms = ms(prev, 11);
    t.integrate_kernel_ms = ms(12, 13);
    t.total_ms = ms(0, 11);
    *out = t;
  });
}

kf_status kf_fused_counts(kf_context* ctx, int64_t* counts) {
  return guard(ctx, [&] {
    require(counts != nullptr, "kf_fused_counts: null argument");
    unsigned long long host[4];
    KF_HIP(hipMemcpyAsync(host, ctx->fused, sizeof(host), hipMemcpyDeviceToHost, ctx->stream));
    KF_HIP(hipStreamSynchronize(ctx->stream));
    counts[0] = static_cast<int64_t>(host[2]);
    counts[1] = static_cast<int64_t>(host[3]);
  });
}

kf_status kf_observed_voxel_count(kf_context* ctx, int64_t* count) {
  return guard(ctx, [&] {
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->fused);
    const int grid = 2048;
    switch (ctx->cfg.voxel_store) {
      case KF_VOXEL_QUANTIZED:
        hipLaunchKernelGGL(count_observed_kernel<KF_VOXEL_QUANTIZED>, dim3(grid), dim3(256), 0, ctx->stream,
                           static_cast<const uint32_t*>(ctx->voxels), ctx->n_voxels, ctx->fused);
        break;
      case KF_VOXEL_BF16:
        hipLaunchKernelGGL(count_observed_kernel<KF_VOXEL_BF16>, dim3(grid), dim3(256), 0, ctx->stream,
                           static_cast<const uint32_t*>(ctx->voxels), ctx->n_voxels, ctx->fused);
        break;
      default:
        hipLaunchKernelGGL(count_observed_kernel<KF_VOXEL_FLOAT>, dim3(grid), dim3(256), 0, ctx->stream,
                           static_cast<const float2*>(ctx->voxels), ctx->n_voxels, ctx->fused);
        break;
    }
    KF_HIP(hipGetLastError());
    unsigned long long host = 0;
    KF_HIP(hipMemcpyAsync(&host, ctx->fused, sizeof(host), hipMemcpyDeviceToHost, ctx->stream));
    KF_HIP(hipStreamSynchronize(ctx->stream));
    *count = static_cast<int64_t>(host);
  });
}

kf_status kf_volume_bytes(const kf_context* ctx, size_t* voxel_bytes_out, size_t* color_bytes) {
  if (ctx == nullptr) return KF_ERR_INVALID_ARGUMENT;
  const kf_config& c = ctx->cfg;
  const size_t dense = static_cast<size_t>(c.resolution[0]) * c.resolution[1] * c.resolution[2];
  if (voxel_bytes_out != nullptr) *voxel_bytes_out = dense * voxel_bytes(c.voxel_store);
  if (color_bytes != nullptr) *color_bytes = ctx->colors != nullptr ? dense * sizeof(float4) : 0;
  return KF_OK;
}

kf_status kf_volume_download(kf_context* ctx, void* voxels, float* colors) {
  return guard(ctx, [&] {
    if (hashed(ctx)) {
      // stage the dense expansion on the device, then copy it out
      size_t vb = 0, cb = 0;
      kf_volume_bytes(ctx, &vb, &cb);
      void* dv = dalloc<uint8_t>(vb);
      float4* dc = cb != 0 ? dalloc<float4>(cb / sizeof(float4)) : nullptr;
      try {
        launch_expand_hashed(ctx, dv, dc);
        if (voxels != nullptr) KF_HIP(hipMemcpyAsync(voxels, dv, vb, hipMemcpyDeviceToHost, ctx->stream));
        if (colors != nullptr && dc != nullptr)
          KF_HIP(hipMemcpyAsync(colors, dc, cb, hipMemcpyDeviceToHost, ctx->stream));
        KF_HIP(hipStreamSynchronize(ctx->stream));
      } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        dfree(dv);
        dfree(dc);
        throw;
      }
      dfree(dv);
      dfree(dc);
      return;
    }
    if (voxels != nullptr)
      KF_HIP(hipMemcpyAsync(voxels, ctx->voxels, ctx->n_voxels * voxel_bytes(ctx->cfg.voxel_store),
                            hipMemcpyDeviceToHost, ctx->stream));
    if (colors != nullptr && ctx->colors != nullptr)
      KF_HIP(hipMemcpyAsync(colors, ctx->colors, ctx->n_voxels * sizeof(float4), hipMemcpyDeviceToHost,
                            ctx->stream));
    KF_HIP(hipStreamSynchronize(ctx->stream));
  });
}

kf_status kf_volume_upload(kf_context* ctx, const void* voxels, const float* colors) {
  return guard(ctx, [&] {
    if (hashed(ctx)) fail(KF_ERR_UNSUPPORTED, "kf_volume_upload: a hashed volume is built by integration only");
    if (voxels != nullptr)
      KF_HIP(hipMemcpyAsync(ctx->voxels, voxels, ctx->n_voxels * voxel_bytes(ctx->cfg.voxel_store),
                            hipMemcpyHostToDevice, ctx->stream));
    if (colors != nullptr && ctx->colors != nullptr)
      KF_HIP(hipMemcpyAsync(ctx->colors, colors, ctx->n_voxels * sizeof(float4), hipMemcpyHostToDevice,
                            ctx->stream));
    // uploaded voxels carry no free-space record until they are integrated again
    KF_HIP(hipMemsetAsync(ctx->part_bits, 0, ctx->part_words * 4, ctx->stream));
    launch_rebuild_observed(ctx);
    KF_HIP(hipStreamSynchronize(ctx->stream));
  });
}

kf_status kf_volume_device(kf_context* ctx, void** voxels, float** colors) {
  if (ctx == nullptr) return KF_ERR_INVALID_ARGUMENT;
  if (voxels != nullptr) *voxels = ctx->voxels;
  if (colors != nullptr) *colors = reinterpret_cast<float*>(ctx->colors);
  return KF_OK;
}

}  // extern "C"

extern "C" kf_status kf_hashed_stats(kf_context* ctx, int64_t* stats) {
  return guard(ctx, [&] {
    require(stats != nullptr, "kf_hashed_stats: null argument");
    require(hashed(ctx), "kf_hashed_stats: the volume is dense");
    uint32_t h[8];
    KF_HIP(hipMemcpyAsync(h, ctx->sparse, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    KF_HIP(hipStreamSynchronize(ctx->stream));
    stats[0] = std::min<int64_t>(h[0], static_cast<int64_t>(ctx->capacity));
    stats[1] = h[4];
    stats[2] = h[2];
    stats[3] = h[1];
  });
}

extern "C" kf_status kf_hashed_blocks(kf_context* ctx, uint32_t* keys, int64_t capacity, int64_t* count) {
  return guard(ctx, [&] {
    require(hashed(ctx), "kf_hashed_blocks: the volume is dense");
    int64_t stats[4];
    const kf_status st = kf_hashed_stats(ctx, stats);
    if (st != KF_OK) fail(st, ctx->error);
    const int64_t n = stats[0];
    if (count != nullptr) *count = n;
    if (keys == nullptr || n == 0) return;
    std::vector<uint32_t> host(static_cast<size_t>(n));
    KF_HIP(hipMemcpyAsync(host.data(), ctx->block_of_slot, n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    KF_HIP(hipStreamSynchronize(ctx->stream));
    std::sort(host.begin(), host.end());
    std::memcpy(keys, host.data(), std::min<int64_t>(n, capacity) * sizeof(uint32_t));
  });
}

extern "C" kf_status kf_hashed_download(kf_context* ctx, uint32_t* keys, void* voxels, float* colors,
                                        int64_t capacity, int64_t* count) {
  return guard(ctx, [&] {
    require(hashed(ctx), "kf_hashed_download: the volume is dense");
    require(capacity >= 0, "kf_hashed_download: negative capacity");
    int64_t stats[4];
    const kf_status st = kf_hashed_stats(ctx, stats);
    if (st != KF_OK) fail(st, ctx->error);
    const int64_t n = stats[0];
    if (count != nullptr) *count = n;
    const int64_t m = std::min(n, capacity);
    if (m == 0) return;
    // slots 0..n-1 hold the allocated blocks; stage them, then reorder by key
    std::vector<uint32_t> key_of(static_cast<size_t>(n));
    KF_HIP(hipMemcpyAsync(key_of.data(), ctx->block_of_slot, n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    KF_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<int64_t> order(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return key_of[a] < key_of[b]; });
    const size_t vb = kBlockVoxels * voxel_bytes(ctx->cfg.voxel_store), cb = kBlockVoxels * sizeof(float4);
    auto gather = [&](const void* dev, size_t per_block, void* out) {
      std::vector<uint8_t> stage(static_cast<size_t>(n) * per_block);
      KF_HIP(hipMemcpyAsync(stage.data(), dev, stage.size(), hipMemcpyDeviceToHost, ctx->stream));
      KF_HIP(hipStreamSynchronize(ctx->stream));
      for (int64_t j = 0; j < m; ++j)
        std::memcpy(static_cast<uint8_t*>(out) + j * per_block, stage.data() + order[j] * per_block, per_block);
    };
    if (keys != nullptr)
      for (int64_t j = 0; j < m; ++j) keys[j] = key_of[order[j]];
    if (voxels != nullptr) gather(ctx->voxels, vb, voxels);
    if (colors != nullptr && ctx->colors != nullptr) gather(ctx->colors, cb, colors);
  });
}

extern "C" kf_status kf_block_bitmaps(kf_context* ctx, uint32_t* observed, uint32_t* skip, uint32_t* free_space,
                                      int64_t* words) {
  return guard(ctx, [&] {
    if (words != nullptr) *words = static_cast<int64_t>(ctx->block_words);
    if (observed != nullptr)
      KF_HIP(hipMemcpyAsync(observed, ctx->obs_bits, ctx->block_words * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (skip != nullptr)
      KF_HIP(hipMemcpyAsync(skip, ctx->skip_bits, ctx->block_words * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (free_space != nullptr)
      KF_HIP(hipMemcpyAsync(free_space, ctx->free_bits, ctx->block_words * 4, hipMemcpyDeviceToHost, ctx->stream));
    // the raycaster's class words hold the same skip and free bits
    std::vector<uint32_t> sk(ctx->block_words), fr(ctx->block_words), cls(2 * ctx->block_words);
    KF_HIP(hipMemcpyAsync(sk.data(), ctx->skip_bits, ctx->block_words * 4, hipMemcpyDeviceToHost, ctx->stream));
    KF_HIP(hipMemcpyAsync(fr.data(), ctx->free_bits, ctx->block_words * 4, hipMemcpyDeviceToHost, ctx->stream));
    KF_HIP(hipMemcpyAsync(cls.data(), ctx->class_bits, ctx->block_words * 8, hipMemcpyDeviceToHost, ctx->stream));
    KF_HIP(hipStreamSynchronize(ctx->stream));
    const BlockMap bm = block_map(ctx->grid);
    const size_t nblocks = static_cast<size_t>(bm.bx) * bm.by * bm.bz;
    for (size_t b = 0; b < nblocks; ++b) {
      const uint32_t c = (cls[b >> 4] >> (2 * (b & 15))) & 3U;
      const uint32_t want = ((sk[b >> 5] >> (b & 31)) & 1U) | (((fr[b >> 5] >> (b & 31)) & 1U) << 1);
      if (c != want) fail(KF_ERR_DEVICE, "kf_block_bitmaps: the class words disagree with the skip and free maps");
    }
  });
}

extern "C" kf_status kf_attach_posenet(kf_context* ctx, kf_posenet* net) {
  return guard(ctx, [&] {
    if (net != nullptr) {
      require(ctx->cfg.init_mode == KF_INIT_LEARNED && ctx->prev_coarse != nullptr,
              "kf_attach_posenet: the context was not created with KF_INIT_LEARNED");
      const LevelBuffers& L = ctx->lv[ctx->cfg.levels - 1];
      require(ctx->cfg.levels >= 3 && posenet_input_matches(L.w, L.h),
              "kf_attach_posenet: the coarsest pyramid level must be 160 x 120");
      // the proposal reads the coarse depth in the context's units, and the
      // frame graph reads the weights that training on the network's
      // stream writes: no training step may still be in flight
      posenet_prepare_attach(net, ctx->cfg.depth_scale);
    }
    KF_HIP(hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    ctx->posenet = net;
  });
}

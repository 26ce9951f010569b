/*
 * kf_hip.h -- C ABI of the MI355X (gfx950) KinectFusion library libkf_hip.so.
 *
 * The library is the HIP memory space behind the reference's stage
 * interfaces (maximsolodukhin/KinectFusion):
 *   Pipeline              include/kinectfusion/pipeline.hpp:64-161
 *   PyramidSource         include/kinectfusion/depth_processing.hpp:422-463
 *   ProjectiveIcpTracker  include/kinectfusion/icp_optimizer.hpp:107-130
 *   MarchingCubes         include/kinectfusion/marching_cubes.hpp:16-34
 *
 * Conventions
 *  - Every call returns kf_status; KF_OK is 0.  No C++ exception crosses
 *    the ABI.  kf_last_error(ctx) describes the last failure (ctx may be
 *    NULL for failures of kf_create).
 *  - A context is confined to one host thread and owns one HIP stream; all
 *    device work of the context is ordered on it.  Calls that return host
 *    data synchronise that stream; the others only enqueue work.
 *  - Pointers documented "device" are HIP device pointers (inputs resident
 *    in HBM); "host" pointers are ordinary host memory.
 *  - Poses are camera_to_world, row-major 4x4 FP64.  The 3x4 float
 *    transforms are row-major [R | t].
 *  - Images are row-major (index y * width + x); depth is uint16 raw sensor
 *    units; colour is packed RGBA8 (r in the low byte).
 *  - Vertex/normal/point maps on the device are float4 (x, y, z, 0); invalid
 *    entries hold quiet NaN (include/kinectfusion/vector.hpp:100-103).
 *  - Dense voxels are stored x-fastest (index (z * ny + y) * nx + x,
 *    include/kinectfusion/vector.hpp:62-68): float store = {float tsdf,
 *    float weight}; quantized store = {int16 tsdf, uint16 weight}; bf16
 *    store = {uint16 bf16 tsdf bits, uint16 weight}
 *    (include/kinectfusion/volume.hpp:28-153).  Colour voxels are float4
 *    (r, g, b, weight) (volume.hpp:169-184).
 */
#ifndef KF_HIP_H
#define KF_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum kf_status {
  KF_OK = 0,
  KF_ERR_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference */
  KF_ERR_OUT_OF_MEMORY = 2,
  KF_ERR_DEVICE = 3,           /* std::runtime_error from cuda::check */
  KF_ERR_UNSUPPORTED = 4,      /* std::logic_error: not built / no device */
  KF_ERR_STATE = 5             /* call order violated (e.g. no pyramid yet) */
} kf_status;

/* VoxelStore, ColorStore, StorageLayout (pipeline.hpp:24-26). */
enum { KF_VOXEL_FLOAT = 0, KF_VOXEL_QUANTIZED = 1, KF_VOXEL_BF16 = 2 };
enum { KF_COLOR_FLOAT = 0, KF_COLOR_NONE = 1 };
enum { KF_STORAGE_DENSE = 0, KF_STORAGE_HASHED = 1 };
/* TsdfRuleVariant (tsdf_integration.hpp:349). */
enum { KF_RULE_CLASSIC = 0, KF_RULE_ANGLE_WEIGHTED = 1 };
/* IcpDampingMode (icp_correspondence.hpp:58). */
enum { KF_DAMPING_NONE = 0, KF_DAMPING_IDENTITY = 1, KF_DAMPING_DIAGONAL = 2 };
/* IntegrationMode (tsdf_integration.hpp:29): the full sweep carves free
 * space everywhere; the band integrates only the 8^3 blocks that the
 * truncation band of some pixel crosses (TruncationBandWalk at a one-voxel
 * step, the host BandIntegrationSweep, tsdf_integration.hpp:444-497), and
 * voxels outside the band keep their values.  A hashed volume integrates
 * its band (at half-block steps, block_rep.hpp:287) and requires
 * KF_INTEGRATION_BAND, as the reference's sparse storage does
 * (Pipeline::require_valid_storage, src/kinectfusion/pipeline.cpp:108-117). */
enum { KF_INTEGRATION_FULL = 0, KF_INTEGRATION_BAND = 1 };
/* How kf_process_frame seeds ICP. */
enum { KF_INIT_PREVIOUS = 0, KF_INIT_CONSTANT_VELOCITY = 1, KF_INIT_LEARNED = 2 };

/* Everything needed to compose one pipeline: PipelineConfig
 * (pipeline.hpp:29-46) flattened, with TsdfIntegrationOptions
 * (tsdf_integration.hpp:31-50), RaycastOptions (raycasting.hpp:66-78),
 * DepthProcessingOptions (depth_processing.hpp:31-46) and the level-0
 * camera.  kf_config_defaults fills the reference defaults. */
typedef struct kf_config {
  /* VolumeGeometry (volume.hpp:265-287) */
  int64_t resolution[3];
  float voxel_size;
  float origin[3];
  float truncation_distance;
  int32_t voxel_store;
  int32_t color_store;
  int32_t storage;
  int32_t tsdf_rule;
  /* TsdfIntegrationOptions */
  float depth_scale;
  float observation_weight;
  float max_weight;
  float min_depth;
  float max_depth;
  int32_t projective_distance;
  int32_t distance_scaled_truncation;
  float truncation_distance_scale;
  int32_t integration_mode;  /* KF_INTEGRATION_*; KF_INTEGRATION_BAND with hashed storage */
  /* RaycastOptions */
  float ray_min_depth;
  float ray_max_depth;
  float step_scale;
  int32_t require_all_corners; /* CornerPolicy::kRequireAll */
  int32_t cell_normals;
  /* RaycastOptions::seed_from_previous (raycasting.hpp:77,122-131): a ray
   * starts max(ray_min_depth, |p - camera| - 2 truncation) in front of the
   * point p its pixel found in the previous raycast into the same level
   * slot, when that point is finite and non-zero; maps start zero-filled
   * (kf_create, kf_reset). */
  int32_t seed_from_previous;
  /* DepthProcessingOptions */
  int32_t levels;
  float max_normal_depth_jump;
  float max_downsample_depth_jump;
  int32_t bilateral_filter;
  int32_t bilateral_radius;
  float bilateral_spatial_sigma;
  float bilateral_depth_sigma;
  /* level-0 camera */
  int32_t width;
  int32_t height;
  float fx, fy, cx, cy;
  /* hashed volume (KF_STORAGE_HASHED).  block_capacity 0: a quarter of
   * the grid's 8^3 blocks (PipelineConfig::sparse_block_capacity,
   * pipeline.hpp:40); hash_buckets 0: the smallest power of two that is at
   * least twice the capacity. */
  int64_t hash_buckets;
  int64_t block_capacity;
  /* kf_process_frame */
  int32_t init_mode;
  int32_t icp_iterations[4]; /* per level; reference default 10/5/4; 0 skips a level */
  int32_t stage_timing;      /* record HIP events per stage in kf_process_frame */
  int32_t count_fused;       /* kf_process_frame accumulates fused-voxel counts */
  /* KF_INIT_LEARNED (no reference counterpart): per-level iterations run from
   * the network's proposal (0 skips the level and its raycast), and the
   * thresholds that keep the proposal's result: the finest level's final
   * point-to-plane RMS (m), and its last increment (m, rad) -- a run still
   * moving more than that at its last iteration needed more iterations than
   * its schedule gives (0 disables the increment test).  Otherwise the
   * frame is tracked again from the constant-velocity seed with
   * icp_iterations. */
  int32_t learned_iterations[4];
  float learned_max_rms;
  float learned_max_update_translation;
  float learned_max_update_rotation;
} kf_config;

/* ProjectiveIcpOptions (icp_optimizer.hpp:82-102). */
typedef struct kf_icp_options {
  int64_t min_correspondences;
  float max_point_distance;
  float min_normal_dot;
  float min_update_translation;
  float min_update_rotation;
  float max_update_translation;
  float max_update_rotation;
  float min_system_eigenvalue;
  float max_condition_number;
  int32_t damping_mode;
  float damping_lambda;
  int32_t adaptive_damping;
  float lambda_increase, lambda_decrease, min_lambda, max_lambda;
} kf_icp_options;

/* IcpOutcome diagnostics (icp_optimizer.hpp:43-62).  status: 0 max
 * iterations, 1 converged, 2 too few correspondences, 3 solve failed,
 * 4 update too large, 5 unconstrained system, -1 level skipped. */
typedef struct kf_icp_result {
  int32_t status;
  int32_t trials;
  int32_t rejected;
  int32_t pad;
  int64_t correspondences;
  double mean_distance;
  double mean_squared_residual;
  double min_eigenvalue;
  double condition_number;
  double update_translation;
  double update_rotation;
  double lambda;
} kf_icp_result;

/* One pyramid level as ICP consumes it (PyramidLevel,
 * depth_processing.hpp:415-418): device views valid until the next
 * kf_pyramid_build. */
typedef struct kf_level_view {
  int32_t width, height;
  float fx, fy, cx, cy;
  const uint16_t* depth;  /* filtered (level 0) / downsampled depth */
  const float* vertices;  /* float4 per pixel */
  const float* normals;   /* float4 per pixel */
} kf_level_view;

/* Raycast output (SurfaceMaps, volume.hpp:208-219): device views owned by
 * the context, valid until the next raycast into the same level slot. */
typedef struct kf_maps_view {
  int32_t width, height;
  const float* points;    /* float4 */
  const float* normals;   /* float4 */
  const uint32_t* colors; /* RGBA8 */
} kf_maps_view;

/* A frame to integrate (DepthFrame, tsdf_integration.hpp:68-89). */
typedef struct kf_depth_frame {
  const uint16_t* depth;  /* device, raw depth */
  const uint32_t* color;  /* device or NULL */
  const float* normals;   /* device float4 or NULL (angle rule falls back) */
  int32_t width, height;
  float fx, fy, cx, cy;
  float world_to_camera[12];
} kf_depth_frame;

/* Per-stage device time of the last kf_process_frame, from HIP events. */
typedef struct kf_stage_times {
  float pyramid_ms;
  float raycast_ms;
  float icp_ms;
  float integrate_ms;        /* prep + cull + integrate kernels */
  float integrate_kernel_ms; /* the voxel-update kernel alone */
  float total_ms;
  float proposal_ms;         /* pose network + seed (KF_INIT_LEARNED) */
} kf_stage_times;

/* Result of kf_process_frame (Reconstruction::process_frame,
 * src/app/reconstruction.cpp:184-216). */
typedef struct kf_frame_result {
  int32_t tracked;        /* 1 when every level succeeded and it integrated */
  int32_t failed_level;   /* -1 or the first failing level */
  int32_t used_learned;   /* learned warm start accepted (KF_INIT_LEARNED) */
  int32_t pad;
  double pose[16];
  kf_icp_result level[4];
  /* KF_INIT_LEARNED: the proposal's distance from the final pose and the
   * final RMS of the tracking run from it */
  double proposal_translation_error; /* m */
  double proposal_rotation_error;    /* rad */
  double learned_rms;                /* m */
  /* KF_INIT_LEARNED: the seed tracking started from, T_previous [R(omega) | t]
   * of the network's twist (camera_to_world, row-major FP64) */
  double proposal_pose[16];
} kf_frame_result;

typedef struct kf_context kf_context;
typedef struct kf_mesh kf_mesh;
typedef struct kf_posenet kf_posenet;

/* Pose-proposal network of the learned warm start (no reference
 * counterpart): training and inference options. */
typedef struct kf_posenet_config {
  int32_t batch_capacity; /* largest batch of pairs per call */
  int32_t pad;
  uint64_t seed;          /* weight initialisation */
  float learning_rate;    /* Adam */
  float beta1, beta2, epsilon;
  float depth_scale;      /* raw depth units per metre */
} kf_posenet_config;

/* ---- lifecycle: Pipeline::create (pipeline.hpp:118-121, pipeline.cpp:129-147).
 * No silent host fallback: KF_ERR_UNSUPPORTED when no gfx950 device is
 * usable. */
void kf_config_defaults(kf_config* cfg);
void kf_icp_options_defaults(kf_icp_options* opt);
kf_status kf_create(const kf_config* cfg, kf_context** out);
void kf_destroy(kf_context* ctx);
const char* kf_last_error(const kf_context* ctx);
/* Waits for the context's stream.  Also reports, once, a tracking level of
 * the learned warm start's fallback pass whose grid-wide wait timed out
 * (KF_ERR_DEVICE; that frame was marked failed and not integrated). */
kf_status kf_synchronize(kf_context* ctx);
/* Bound of that grid-wide wait in ticks of the 100 MHz device wall clock
 * (default 2,000,000 = 20 ms, far beyond any trial).  No reference
 * counterpart: a diagnostic that lets a test force the timeout path. */
kf_status kf_set_sync_timeout(kf_context* ctx, int64_t ticks);
/* Fresh volume: tsdf 1, weight 0 (volume.hpp:23,107-108). */
kf_status kf_reset(kf_context* ctx);

/* ---- PyramidSource::build / level (depth_processing.hpp:432-435;
 * depth_processing.cpp:138-163).  depth: device raw depth of the config's
 * width x height.  Returns the level count in *levels. */
kf_status kf_pyramid_build(kf_context* ctx, const uint16_t* depth, int32_t* levels);
kf_status kf_pyramid_level(kf_context* ctx, int32_t level, kf_level_view* out);

/* ---- Pipeline::integrate (pipeline.hpp:78-80; TsdfIntegrator,
 * tsdf_integration.hpp:501-524; dense full sweep, tsdf_integration.hpp:356-370).
 * counts (host, may be NULL): when non-NULL the kernel also counts the
 * voxels fused (counts[0]) and colour voxels fused (counts[1]) and the call
 * synchronises. */
kf_status kf_integrate(kf_context* ctx, const kf_depth_frame* frame, int64_t* counts);

/* ---- Pipeline::raycast (pipeline.hpp:82; SurfaceRaycast,
 * raycasting.hpp:116-245).  Renders into level slot `level` with the
 * config camera scaled by 2^-level. */
kf_status kf_raycast(kf_context* ctx, const double* camera_to_world, int32_t level,
                     kf_maps_view* out);

/* ---- ICP (CorrespondenceSearch::match + IcpNormalEquations,
 * icp_correspondence.hpp:104-294).  One correspondence sweep over the
 * given device maps (float4), reduced on the device with compensated FP32
 * sums.  sums (host): 21 upper-triangle J^T J, 6 J^T r, distance sum,
 * residual sum of squares.  *count (host): accepted correspondences. */
kf_status kf_icp_reduce(kf_context* ctx, const float* live_vertices, const float* live_normals,
                        int32_t live_width, int32_t live_height, const float* model_points,
                        const float* model_normals, int32_t model_width, int32_t model_height,
                        const float* model_intrinsics, const float* camera, const float* model,
                        float max_point_distance, float min_normal_dot, double* sums,
                        int64_t* count);

/* Allocation-free form of kf_icp_reduce.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It takes
 * no context and calls no allocator: the sweep's partial rows (at most 256
 * rows of 32 floats, 32 KiB) live in a static array of the library that each
 * device holds until the process ends.  Live and model maps of at most
 * KF_ICP_STATIC_MAX_PIXELS pixels each; larger maps are rejected
 * (KF_ERR_INVALID_ARGUMENT).  Inputs as kf_icp_reduce (maps on the device,
 * model_intrinsics / camera / model on the host); outputs on the DEVICE:
 * sums (29 doubles, the FP32 compensated sums widened) and count (one
 * int64).  stream: a hipStream_t, or NULL for the default stream; the call
 * only enqueues.  Every call on one device shares that device's rows, so
 * calls that overlap on one device (other streams, other threads) race;
 * each call fully writes the rows it reads, so nothing needs clearing.
 * Failures are described by kf_last_error(NULL). */
#define KF_ICP_STATIC_MAX_PIXELS (1 << 24)
kf_status kf_icp_reduce_static(const float* live_vertices, const float* live_normals, int32_t live_width,
                               int32_t live_height, const float* model_points, const float* model_normals,
                               int32_t model_width, int32_t model_height, const float* model_intrinsics,
                               const float* camera, const float* model, float max_point_distance,
                               float min_normal_dot, double* sums, int64_t* count, void* stream);

/* ProjectiveIcpTracker::estimate_pose in its whole-level device-loop form
 * (icp_optimizer.cpp:82-169, cuda/icp_correspondence.cu:216-343), with the
 * 6x6 Cholesky solve and SE(3) composition in FP64 on the device and one
 * synchronisation at the end.  pose_inout (host): initial camera_to_world,
 * replaced by the estimate on success. */
kf_status kf_icp_track_maps(kf_context* ctx, const float* live_vertices, const float* live_normals,
                            int32_t live_width, int32_t live_height, const float* model_points,
                            const float* model_normals, int32_t model_width, int32_t model_height,
                            const float* model_intrinsics, const double* model_pose,
                            double* pose_inout, int32_t iterations, const kf_icp_options* opt,
                            kf_icp_result* result);

/* Coarse-to-fine tracking of the current pyramid against the volume
 * (Reconstruction::track_pose, src/app/reconstruction.cpp:218-250): per
 * level, raycast the model at the running pose, then the device loop.  One
 * synchronisation.  pose_inout (host) as above; results[levels] (host). */
kf_status kf_track(kf_context* ctx, const kf_icp_options* opt, double* pose_inout,
                   kf_icp_result* results);

/* ---- whole frame (Reconstruction::process_frame, reconstruction.cpp:184-216):
 * pyramid, tracking seeded per cfg.init_mode, integration when tracking
 * succeeded.  Inputs are device pointers; the running pose stays on the
 * device.  result (host, may be NULL): when non-NULL the call synchronises
 * and reports the frame; when NULL nothing waits on the device.
 * The first frame of a context (or after kf_reset) integrates at the
 * identity pose without tracking (reconstruction.cpp:143-182). */
kf_status kf_process_frame(kf_context* ctx, const uint16_t* depth, const uint32_t* color,
                           const kf_icp_options* opt, kf_frame_result* result);
/* Poses of all frames processed so far (host, 16 doubles each). */
kf_status kf_trajectory(kf_context* ctx, double* poses, int64_t capacity, int64_t* count);
kf_status kf_last_stage_times(kf_context* ctx, kf_stage_times* out);
/* Voxels fused (counts[0]) and colour voxels fused (counts[1]) by every
 * kf_process_frame since kf_create / kf_reset when cfg.count_fused is set
 * (the exact N of the integration's algorithmic bytes).  Synchronises. */
kf_status kf_fused_counts(kf_context* ctx, int64_t* counts);

/* ---- volume access: Pipeline::observed_voxel_count / host_view
 * (pipeline.hpp:87-91). */
kf_status kf_observed_voxel_count(kf_context* ctx, int64_t* count);
kf_status kf_volume_bytes(const kf_context* ctx, size_t* voxel_bytes, size_t* color_bytes);
kf_status kf_volume_download(kf_context* ctx, void* voxels, float* colors);
kf_status kf_volume_upload(kf_context* ctx, const void* voxels, const float* colors);
/* Device pointers of the dense volume (float4 colours may be NULL). */
kf_status kf_volume_device(kf_context* ctx, void** voxels, float** colors);

/* ---- hashed volume (cfg.storage == KF_STORAGE_HASHED): BlockRep
 * (block_rep.hpp:183-307) with a hash of cfg.hash_buckets buckets over a
 * pool of cfg.block_capacity 8^3-voxel blocks.  Integration marks the
 * truncation band of every pixel at half-block steps
 * (TruncationBandWalk, tsdf_integration.hpp:378-428), allocates new blocks,
 * and updates only the blocks visited in the frame.  kf_volume_download
 * returns the dense expansion (unallocated voxels fresh, as
 * BlockRep::host_dense_view); kf_volume_upload is not available.
 * stats[0] blocks allocated, [1] requests dropped because the pool or the
 * hash was full (BlockRep::overflowed), [2] hash entries, [3] blocks
 * integrated by the last frame. */
kf_status kf_hashed_stats(kf_context* ctx, int64_t* stats);
/* Keys (flat x-major block indices, grid.hpp:44-59) of the allocated blocks
 * in increasing order; *count receives their number. */
kf_status kf_hashed_blocks(kf_context* ctx, uint32_t* keys, int64_t capacity, int64_t* count);
/* The allocated blocks themselves (SparseVolumeSnapshot,
 * block_rep.hpp:170-182, without the dense expansion): keys as
 * kf_hashed_blocks (host, increasing), then each block's 512 voxels in that
 * order (host, the store's voxel layout, intra-block index
 * (z % 8 * 8 + y % 8) * 8 + x % 8, grid.hpp:61-69) and, when the volume has
 * colour, its 512 float4 colour voxels (host, may be NULL).  At most
 * `capacity` blocks are copied; *count receives the number allocated. */
kf_status kf_hashed_download(kf_context* ctx, uint32_t* keys, void* voxels, float* colors, int64_t capacity,
                             int64_t* count);

/* ---- block maps used to skip samples (no counterpart in the reference; a
 * diagnostic): one bit per 8^3 voxel block, flat x-major block order.
 * observed: some voxel of the block is observed; skip: the block and its 26
 * neighbours are unobserved; free_space: the block and its 26 neighbours
 * hold only observed voxels reading tsdf exactly 1.  *words receives the
 * number of 32-bit words of each map; NULL maps are not copied.  The call
 * also checks the raycaster's per-block class words (skip and free bits
 * side by side) against both maps: KF_ERR_DEVICE if they disagree. */
kf_status kf_block_bitmaps(kf_context* ctx, uint32_t* observed, uint32_t* skip, uint32_t* free_space, int64_t* words);

/* ---- MarchingCubes::extract (marching_cubes.cpp:478-507) on the device:
 * cell classification, stream-compacted triangles and the reference's
 * first-encounter vertex welding order. */
kf_status kf_mesh_extract(kf_context* ctx, float min_weight, kf_mesh** out);
kf_status kf_mesh_size(const kf_mesh* mesh, int64_t* vertices, int64_t* triangles);
kf_status kf_mesh_copy(const kf_mesh* mesh, float* positions, float* normals, float* colors,
                       uint32_t* triangles);
void kf_mesh_free(kf_mesh* mesh);

/* ---- learned warm start.  No counterpart in the reference: the network
 * replaces the tracker's initial pose (Reconstruction::track_pose,
 * src/app/reconstruction.cpp:218-250), and its fallback reads the final
 * IcpNormalEquations::mean_squared_residual (icp_correspondence.hpp:133-139).
 * Input: pairs of 160 x 120 raw depth images (previous, current), the
 * coarsest level of the 640 x 480 pyramid, fed as metres minus 1.5 (0 where
 * invalid); output: the twist (omega rad, t m) with T_current =
 * T_previous [R(omega) | t], the network's six outputs scaled by
 * (0.1, 0.1, 0.1, 0.03, 0.03, 0.03).  Parameters are packed
 * per layer as W[out][in] then b[out], convolution inputs ordered
 * (ky * 3 + kx) * Cin + ci, the flattened 4 x 5 x 128 map in (y, x, c) order.
 * Buffers passed to these calls are device pointers unless noted; calls on
 * one network must not overlap. */
void kf_posenet_config_defaults(kf_posenet_config* cfg);
int64_t kf_posenet_parameter_count(void);
kf_status kf_posenet_create(const kf_posenet_config* cfg, kf_posenet** out);
void kf_posenet_destroy(kf_posenet* net);
const char* kf_posenet_last_error(const kf_posenet* net);
/* FP32 master parameters to / from host memory; setting them clears the Adam
 * moments. */
kf_status kf_posenet_get_parameters(kf_posenet* net, float* params);
kf_status kf_posenet_set_parameters(kf_posenet* net, const float* params);
/* pairs: [batch][2][120][160] u16; twists: [batch][6] floats.  Synchronous. */
kf_status kf_posenet_forward(kf_posenet* net, const uint16_t* pairs, int32_t batch, float* twists);
/* Mean squared error of the normalised output against targets [batch][6]
 * (twists): loss (host) and the gradient of every parameter (host, may be
 * NULL), without updating.  Synchronous. */
kf_status kf_posenet_gradients(kf_posenet* net, const uint16_t* pairs, const float* targets, int32_t batch,
                               float* loss, float* grads);
/* One Adam step on the batch.  loss (host) may be NULL: then nothing waits. */
kf_status kf_posenet_train_step(kf_posenet* net, const uint16_t* pairs, const float* targets, int32_t batch,
                                float* loss);
/* Adam step size of the following kf_posenet_train_step calls (schedules). */
kf_status kf_posenet_set_learning_rate(kf_posenet* net, float learning_rate);
kf_status kf_posenet_synchronize(kf_posenet* net);
/* Use `net` for KF_INIT_LEARNED frames of `ctx` (NULL detaches).  The
 * pipeline must have at least 3 levels with a 160 x 120 coarsest level.
 * The network must outlive its use by the context. */
kf_status kf_attach_posenet(kf_context* ctx, kf_posenet* net);

#ifdef __cplusplus
}
#endif

#endif /* KF_HIP_H */

// Raycaster (reference: SurfaceRaycast::render_pixel / find_zero_crossing /
// write_surface_sample, include/kinectfusion/raycasting.hpp:116-245;
// TrilinearSampler, volume_sampler.hpp:116-255; trilinear helpers,
// trilinear.hpp:85-269).
//
// One thread per pixel, 16 x 16 pixels per workgroup so a wave's rays are a
// compact 16 x 4 footprint and march through neighbouring voxels together.
// Corner blends visit the 2^3 stencil in z -> y -> x order with the
// reference's weight products, so every sample is bit-identical to the CPU
// sampler.
#include "kf_march.hpp"

namespace kf {
namespace {

template <int Kind, typename Store>
__global__ void __launch_bounds__(256) raycast_kernel(Sampler<Kind, Store> vol, RayParams rp, Rigid pose_arg,
                                                       IcpState* state, float4* __restrict__ pts,
                                                       float4* __restrict__ nrm, uint32_t* __restrict__ col,
                                                       const uint32_t* __restrict__ class_bits,
                                                       const uint32_t* __restrict__ nbr_obs, BlockMap bm,
                                                       const uint32_t* __restrict__ coarse_bits, BlockMap cm,
                                                       int begin_level) {
  Rigid pose = pose_arg;
  // the level's preparation touches neither frame_failed nor pose, which
  // every workgroup reads below
  if (begin_level >= 0 && state != nullptr && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0)
    icp_begin_level(state, begin_level);
  if (state != nullptr) {
    if (state->frame_failed) return;
    pose = rigid_from_pose64(state->pose);
  }
  // Workgroups are dealt to the 8 XCDs round-robin by their linear id, and
  // each XCD has its own L2: XCD k takes the k-th contiguous run of tiles
  // (row-major), so the rays that share volume blocks share an L2.  The map
  // is a bijection of [0, tiles) for any tile count.
  const int tiles = static_cast<int>(gridDim.x * gridDim.y);
  const int id = static_cast<int>(blockIdx.y * gridDim.x + blockIdx.x);
  const int q = tiles >> 3, r = tiles & 7, k = id & 7;
  const int tile = k * q + min(k, r) + (id >> 3);
  int x, y;
  tile_pixel(tile, static_cast<int>(gridDim.x), static_cast<int>(threadIdx.x), x, y);
  if (x >= rp.w || y >= rp.h) return;
  raycast_pixel(vol, rp, pose, x, y, pts, nrm, col, class_bits, nbr_obs, bm, coarse_bits, cm);
}

template <int Kind>
void launch_t(kf_context* ctx, const RayParams& rp, const Rigid& pose, IcpState* state,
              LevelBuffers& L, int begin_level) {
  using Raw = typename VoxelIO<Kind>::Raw;
  const dim3 grid((rp.w + 15) / 16, (rp.h + 15) / 16);
  if (hashed(ctx)) {
    Sampler<Kind, HashedStore<Kind>> vol{
        HashedStore<Kind>{hash_view(ctx), static_cast<const Raw*>(ctx->voxels), ctx->colors}, ctx->grid,
        1.0F / ctx->grid.voxel};
    hipLaunchKernelGGL((raycast_kernel<Kind, HashedStore<Kind>>), grid, dim3(256), 0, ctx->stream, vol, rp, pose,
                       state, L.model_pts, L.model_nrm, L.model_col, ctx->class_bits, ctx->nbr_obs, block_map(ctx->grid), ctx->coarse_bits, coarse_map(ctx->grid),
                       begin_level);
    return;
  }
  Sampler<Kind> vol{DenseStore<Kind>{static_cast<const Raw*>(ctx->voxels), ctx->colors}, ctx->grid,
                    1.0F / ctx->grid.voxel};
  hipLaunchKernelGGL((raycast_kernel<Kind, DenseStore<Kind>>), grid, dim3(256), 0, ctx->stream, vol, rp, pose, state,
                     L.model_pts, L.model_nrm, L.model_col, ctx->class_bits, ctx->nbr_obs, block_map(ctx->grid), ctx->coarse_bits, coarse_map(ctx->grid),
                     begin_level);
}

}  // namespace

void launch_raycast(kf_context* ctx, int level, const double* /*unused*/, const Rigid* pose_host,
                    bool from_state, int begin_level) {
  const kf_config& c = ctx->cfg;
  LevelBuffers& L = ctx->lv[level];
  const RayParams rp = ray_params(ctx, level);
  const Rigid pose = pose_host != nullptr ? *pose_host : Rigid{};
  IcpState* state = from_state ? ctx->icp : nullptr;
  switch (c.voxel_store) {
    case KF_VOXEL_QUANTIZED:
      launch_t<KF_VOXEL_QUANTIZED>(ctx, rp, pose, state, L, begin_level);
      break;
    case KF_VOXEL_BF16:
      launch_t<KF_VOXEL_BF16>(ctx, rp, pose, state, L, begin_level);
      break;
    default:
      launch_t<KF_VOXEL_FLOAT>(ctx, rp, pose, state, L, begin_level);
      break;
  }
  KF_HIP(hipGetLastError());
}

}  // namespace kf

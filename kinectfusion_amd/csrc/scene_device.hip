// Device renderer of the seeded synthetic RGB-D source: the same per-pixel
// function as the host renderer (scene_geometry.hpp), one thread per pixel,
// many frames per launch.  Used to generate training pairs for the pose
// network at a rate the host renderer cannot reach.
#include <hip/hip_runtime.h>

#define KF_SCENE_FN __host__ __device__ inline
#include "scene_geometry.hpp"

namespace {

__global__ void __launch_bounds__(256) render_kernel(const kfscene::SceneDesc* sc, const double* poses,
                                                     const std::uint64_t* keys, int noise, std::uint16_t* depth,
                                                     std::uint32_t* rgba, float* exact) {
  const int W = sc->width, H = sc->height;
  const int px = blockIdx.x * blockDim.x + threadIdx.x;
  if (px >= W * H) return;
  const int f = blockIdx.y;
  const int u = px % W, v = px / W;
  std::uint16_t d = 0;
  std::uint32_t c = 0;
  double z = 0.0;
  kfscene::render_pixel(*sc, poses + 16 * f, keys[f], noise, u, v, d, c, z);
  const size_t i = static_cast<size_t>(f) * W * H + px;
  depth[i] = d;
  if (rgba != nullptr) rgba[i] = c;
  if (exact != nullptr) exact[i] = static_cast<float>(z);
}

}  // namespace

extern "C" {

// Renders n frames of scene `s` (a kf_scene_create handle) from host poses
// (n row-major camera_to_world 4x4) with host noise keys, into device
// buffers depth[n][H][W] (and rgba / exact z when non-null).  Synchronous.
// Returns 0 or the HIP error code.
int kf_scene_render_device(const void* s, const double* poses, const std::uint64_t* keys, int n, int noise,
                           std::uint16_t* depth, std::uint32_t* rgba, float* exact) {
  if (s == nullptr || poses == nullptr || keys == nullptr || depth == nullptr || n <= 0 || n > 65535) {
    return static_cast<int>(hipErrorInvalidValue);
  }
  const auto* sc = static_cast<const kfscene::SceneDesc*>(s);
  void* buf = nullptr;
  const size_t pose_bytes = sizeof(double) * 16 * static_cast<size_t>(n);
  const size_t key_bytes = sizeof(std::uint64_t) * static_cast<size_t>(n);
  hipError_t err = hipMalloc(&buf, sizeof(kfscene::SceneDesc) + pose_bytes + key_bytes);
  if (err != hipSuccess) return static_cast<int>(err);
  auto* d_sc = static_cast<kfscene::SceneDesc*>(buf);
  auto* d_poses = reinterpret_cast<double*>(static_cast<char*>(buf) + sizeof(kfscene::SceneDesc));
  auto* d_keys = reinterpret_cast<std::uint64_t*>(reinterpret_cast<char*>(d_poses) + pose_bytes);
  err = hipMemcpy(d_sc, sc, sizeof(kfscene::SceneDesc), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(d_poses, poses, pose_bytes, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(d_keys, keys, key_bytes, hipMemcpyHostToDevice);
  if (err == hipSuccess) {
    const dim3 grid((sc->width * sc->height + 255) / 256, n);
    hipLaunchKernelGGL(render_kernel, grid, dim3(256), 0, nullptr, d_sc, d_poses, d_keys, noise, depth, rgba,
                       exact);
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipDeviceSynchronize();
  const hipError_t free_err = hipFree(buf);
  return static_cast<int>(err != hipSuccess ? err : free_err);
}

}  // extern "C"

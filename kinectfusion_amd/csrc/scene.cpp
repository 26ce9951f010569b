// Seeded synthetic RGB-D source: analytic ray casting of box rooms, spheres
// and boxes along smooth camera trajectories, with the axial depth-noise model
// sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2 m, 1 % dropped pixels, uint16
// millimetre depth and per-surface RGBA8 colour.
//
// The reference reads TUM PNG sequences (src/kinectfusion/virtual_sensor.cpp);
// this source replaces that input with frames whose ground truth is exact.
// Depth is the camera-frame z of the first hit along the pixel-centre ray,
// the convention of CameraIntrinsics::back_project (include/kinectfusion/
// rgbd.hpp:53-57).  World = the camera frame of frame 0.
//
// Randomness is counter-based (SplitMix64 of seed/frame/pixel) with
// Box-Muller normals, so any frame renders identically on any host and in
// any order.  The per-pixel renderer lives in scene_geometry.hpp, shared with
// the device renderer.

#define KF_SCENE_FN inline
#include "scene_geometry.hpp"

#include <algorithm>
#include <cstring>

namespace {

using kfscene::Box;
using kfscene::D3;
using kfscene::SceneDesc;
using kfscene::Sphere;
using kfscene::splitmix64;
using kfscene::unit;

inline D3 operator+(D3 a, D3 b) { return kfscene::add(a, b); }
inline D3 operator-(D3 a, D3 b) { return kfscene::sub(a, b); }
inline D3 operator*(D3 a, double s) { return kfscene::scale(a, s); }
using kfscene::dot;

struct Rng {
  std::uint64_t state;
  std::uint64_t next() { return state = splitmix64(state); }
  double uniform(double lo, double hi) { return lo + (hi - lo) * unit(next()); }
};

std::uint32_t random_rgba(Rng& rng) {
  const auto c = [&]() { return static_cast<std::uint32_t>(rng.uniform(40.0, 250.0)); };
  const std::uint32_t r = c(), g = c(), b = c();
  return r | (g << 8U) | (b << 16U) | (0xFFU << 24U);
}

void euler_to_rot(double rx, double ry, double rz, double* R) {
  // R = Rz(rz) * Ry(ry) * Rx(rx), row-major 3x3
  const double cx = std::cos(rx), sx = std::sin(rx), cy = std::cos(ry), sy = std::sin(ry);
  const double cz = std::cos(rz), sz = std::sin(rz);
  R[0] = cz * cy;
  R[1] = cz * sy * sx - sz * cx;
  R[2] = cz * sy * cx + sz * sx;
  R[3] = sz * cy;
  R[4] = sz * sy * sx + cz * cx;
  R[5] = sz * sy * cx - cz * sx;
  R[6] = -sy;
  R[7] = cy * sx;
  R[8] = cy * cx;
}

}  // namespace

extern "C" {

// kind 0: "room"  -- box room 1.2 m wide (x in [-0.6, 0.6], y in [-0.5, 0.5],
//                    z in [-0.4, 1.8]) with 6 spheres of radius U(0.05, 0.2) m
//                    whose surfaces lie 0.5-1.2 m in front of the first camera.
// kind 1: "hall"  -- 6 m x 3 m x 4 m room (x, y, z) with 20 boxes and 20
//                    spheres; the camera pans (+-18 deg) and walks along x
//                    (+-0.38 m) back and forth, one sweep per 300 frames.
void* kf_scene_create(int kind, std::uint64_t seed, int width, int height, double fx, double fy,
                      double cx, double cy) {
  auto* sc = new SceneDesc();
  std::memset(sc, 0, sizeof(SceneDesc));
  sc->kind = kind;
  sc->seed = seed;
  sc->width = width;
  sc->height = height;
  sc->fx = fx;
  sc->fy = fy;
  sc->cx = cx;
  sc->cy = cy;
  Rng rng{splitmix64(seed * 0x632BE59BD9B4E019ULL + 17U + static_cast<std::uint64_t>(kind))};
  if (kind == 0) {
    sc->room_lo = {-0.6, -0.5, -0.4};
    sc->room_hi = {0.6, 0.5, 1.8};
    for (auto& w : sc->wall_rgba) w = random_rgba(rng);
    for (int i = 0; i < 6; ++i) {
      Sphere s{};
      s.r = rng.uniform(0.05, 0.2);
      // the sphere's surface lies 0.5-1.2 m from the first camera
      const double dist = rng.uniform(0.5, 1.2) + s.r;
      const double u = rng.uniform(0.15, 0.85) * width, v = rng.uniform(0.15, 0.85) * height;
      D3 d{(u - cx) / fx, (v - cy) / fy, 1.0};
      d = d * (1.0 / std::sqrt(dot(d, d)));
      s.c = d * dist;
      // keep the sphere inside the room with a small gap to the walls
      s.c.x = std::clamp(s.c.x, sc->room_lo.x + s.r + 0.02, sc->room_hi.x - s.r - 0.02);
      s.c.y = std::clamp(s.c.y, sc->room_lo.y + s.r + 0.02, sc->room_hi.y - s.r - 0.02);
      s.rgba = random_rgba(rng);
      sc->spheres[sc->n_spheres++] = s;
    }
    // per-frame translation <= sqrt(3) * 0.0055 m, rotation <= 3 * 0.16 deg
    for (int a = 0; a < 3; ++a) {
      sc->t_freq[a] = rng.uniform(0.03, 0.06);
      // the forward swing is kept small so surfaces stay beyond the 0.4 m
      // minimum depth
      sc->t_amp[a] = 0.0055 / sc->t_freq[a] * rng.uniform(0.6, 1.0) * (a == 2 ? 0.3 : 1.0);
      sc->t_phase[a] = rng.uniform(0.0, 6.283185307179586);
      sc->r_freq[a] = rng.uniform(0.03, 0.06);
      sc->r_amp[a] = (0.16 * 3.141592653589793 / 180.0) / sc->r_freq[a] * rng.uniform(0.6, 1.0);
      sc->r_phase[a] = rng.uniform(0.0, 6.283185307179586);
    }
    sc->pan_amp = 0.0;
    sc->walk_amp = 0.0;
    sc->sweep_period = 0.0;
  } else {
    sc->room_lo = {-3.0, -1.5, -1.0};
    sc->room_hi = {3.0, 1.5, 3.0};
    for (auto& w : sc->wall_rgba) w = random_rgba(rng);
    for (int i = 0; i < 40; ++i) {
      // objects stay clear of the camera corridor |x| <= 1.6, |z| <= 0.5
      D3 c{};
      for (int tries = 0; tries < 100; ++tries) {
        c = {rng.uniform(-2.6, 2.6), rng.uniform(-1.2, 1.2), rng.uniform(-0.7, 2.7)};
        if (!(std::fabs(c.z) < 0.9 && std::fabs(c.x) < 2.0)) break;
      }
      if (i % 2 == 0) {
        Box b{};
        const D3 h{rng.uniform(0.1, 0.3), rng.uniform(0.1, 0.3), rng.uniform(0.1, 0.3)};
        b.lo = c - h;
        b.hi = c + h;
        b.rgba = random_rgba(rng);
        sc->boxes[sc->n_boxes++] = b;
      } else {
        Sphere s{};
        s.c = c;
        s.r = rng.uniform(0.1, 0.35);
        s.rgba = random_rgba(rng);
        sc->spheres[sc->n_spheres++] = s;
      }
    }
    // per-frame translation <= (0.38 * 2 pi / 300 + 0.0015) m along x and
    // 0.0015 m along y and z (<= 9.7 mm in all), rotation <= 18 deg * 2 pi /
    // 300 + 3 * 0.05 deg (<= 0.47 deg)
    for (int a = 0; a < 3; ++a) {
      sc->t_freq[a] = rng.uniform(0.02, 0.04);
      sc->t_amp[a] = 0.0015 / sc->t_freq[a];
      sc->t_phase[a] = rng.uniform(0.0, 6.283185307179586);
      sc->r_freq[a] = rng.uniform(0.02, 0.04);
      sc->r_amp[a] = (0.05 * 3.141592653589793 / 180.0) / sc->r_freq[a];
      sc->r_phase[a] = rng.uniform(0.0, 6.283185307179586);
    }
    sc->pan_amp = 18.0 * 3.141592653589793 / 180.0;
    sc->walk_amp = 0.38;
    sc->sweep_period = 300.0;
  }
  return sc;
}

void kf_scene_destroy(void* s) { delete static_cast<SceneDesc*>(s); }

// Ground-truth camera_to_world of frame k (row-major 4x4, FP64); frame 0 is
// the identity.
void kf_scene_pose(const void* s, int frame, double* pose) {
  const SceneDesc& sc = *static_cast<const SceneDesc*>(s);
  const double k = frame;
  double t[3], e[3];
  for (int a = 0; a < 3; ++a) {
    t[a] = sc.t_amp[a] * (std::sin(sc.t_freq[a] * k + sc.t_phase[a]) - std::sin(sc.t_phase[a]));
    e[a] = sc.r_amp[a] * (std::sin(sc.r_freq[a] * k + sc.r_phase[a]) - std::sin(sc.r_phase[a]));
  }
  // hall sweep: pan about y and walk along x, back and forth
  if (sc.sweep_period > 0.0) {
    const double phase = 6.283185307179586 * k / sc.sweep_period;
    e[1] += sc.pan_amp * std::sin(phase);
    t[0] += sc.walk_amp * std::sin(phase);
  }
  double R[9];
  euler_to_rot(e[0], e[1], e[2], R);
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) pose[r * 4 + c] = R[r * 3 + c];
    pose[r * 4 + 3] = t[r];
  }
  pose[12] = 0.0;
  pose[13] = 0.0;
  pose[14] = 0.0;
  pose[15] = 1.0;
}

// Renders the view from camera_to_world `pose`.  depth_mm: uint16 millimetres
// (0 = invalid); rgba: packed RGBA8 (may be null); true_depth: exact z in
// metres before noise (may be null).  noise_key selects the noise stream
// (use the frame index); noise == 0 renders the exact depth, rounded to mm.
void kf_scene_render(const void* s, const double* pose, std::uint64_t noise_key, int noise,
                     std::uint16_t* depth_mm, std::uint32_t* rgba, float* true_depth) {
  const SceneDesc& sc = *static_cast<const SceneDesc*>(s);
  const int W = sc.width, H = sc.height;
#pragma omp parallel for schedule(static)
  for (int v = 0; v < H; ++v) {
    for (int u = 0; u < W; ++u) {
      const std::size_t i = static_cast<std::size_t>(v) * W + u;
      std::uint16_t d = 0;
      std::uint32_t c = 0;
      double z = 0.0;
      kfscene::render_pixel(sc, pose, noise_key, noise, u, v, d, c, z);
      depth_mm[i] = d;
      if (rgba != nullptr) rgba[i] = c;
      if (true_depth != nullptr) true_depth[i] = static_cast<float>(z);
    }
  }
}

// The plain scene description (kfscene::SceneDesc) for the device renderer.
const void* kf_scene_description(const void* s, std::size_t* bytes) {
  if (bytes != nullptr) *bytes = sizeof(SceneDesc);
  return s;
}

}  // extern "C"

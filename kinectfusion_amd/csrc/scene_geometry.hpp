// Scene description and per-pixel rendering of the seeded synthetic RGB-D
// source, shared by the host renderer (scene.cpp, OpenMP) and the device
// renderer (scene_device.hip).  The includer defines KF_SCENE_FN: plain
// `inline` on the host, `__host__ __device__ inline` in HIP code.
//
// Depth is the camera-frame z of the first hit along the pixel-centre ray,
// the convention of CameraIntrinsics::back_project (include/kinectfusion/
// rgbd.hpp:53-57).  Noise: sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2 m,
// Gaussian (Box-Muller), 1 % dropped pixels, millimetre rounding.
// Randomness is counter-based (SplitMix64 of seed, frame key and pixel), so
// any frame renders identically in any order.
#pragma once

#include <cmath>
#include <cstdint>

#ifndef KF_SCENE_FN
#error "define KF_SCENE_FN before including scene_geometry.hpp"
#endif

namespace kfscene {

constexpr int kMaxSpheres = 64;
constexpr int kMaxBoxes = 64;

struct D3 {
  double x, y, z;
};
KF_SCENE_FN D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
KF_SCENE_FN D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
KF_SCENE_FN D3 scale(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
KF_SCENE_FN double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
KF_SCENE_FN double dmax(double a, double b) { return a < b ? b : a; }
KF_SCENE_FN double dmin(double a, double b) { return b < a ? b : a; }

KF_SCENE_FN std::uint64_t splitmix64(std::uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
KF_SCENE_FN double unit(std::uint64_t bits) {  // [0, 1) with 53 bits
  return static_cast<double>(bits >> 11) * (1.0 / 9007199254740992.0);
}

struct Sphere {
  D3 c;
  double r;
  std::uint32_t rgba;
};
struct Box {
  D3 lo, hi;
  std::uint32_t rgba;
};

// Plain data: copied as is to the device.
struct SceneDesc {
  int kind;
  std::uint64_t seed;
  int width, height;
  double fx, fy, cx, cy;
  D3 room_lo, room_hi;
  std::uint32_t wall_rgba[6];
  int n_spheres, n_boxes;
  Sphere spheres[kMaxSpheres];
  Box boxes[kMaxBoxes];
  // trajectory parameters
  double t_amp[3], t_freq[3], t_phase[3];
  double r_amp[3], r_freq[3], r_phase[3];
  double pan_amp;    // hall sweep: pan about the vertical axis, pan_amp sin(2 pi k / sweep_period) rad
  double walk_amp;   // hall sweep: walk along x, walk_amp sin(2 pi k / sweep_period) m
  double sweep_period;  // frames of one back-and-forth sweep (0: no sweep)
};

// Ray/sphere: smallest s > eps with |o + s d - c| = r.
KF_SCENE_FN bool hit_sphere(const Sphere& s, D3 o, D3 d, double& t) {
  const D3 oc = sub(o, s.c);
  const double b = dot(oc, d), c = dot(oc, oc) - s.r * s.r, a = dot(d, d);
  const double disc = b * b - a * c;
  if (disc < 0.0) return false;
  const double q = sqrt(disc);
  double s0 = (-b - q) / a;
  if (s0 <= 1e-9) s0 = (-b + q) / a;
  if (s0 <= 1e-9) return false;
  t = s0;
  return true;
}

// Ray/box (outside surface): slab entry.
KF_SCENE_FN bool hit_box(const Box& bx, D3 o, D3 d, double& t) {
  double tmin = -1e300, tmax = 1e300;
  const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  const double lo[3] = {bx.lo.x, bx.lo.y, bx.lo.z}, hi[3] = {bx.hi.x, bx.hi.y, bx.hi.z};
  for (int a = 0; a < 3; ++a) {
    if (fabs(dd[a]) < 1e-15) {
      if (oo[a] < lo[a] || oo[a] > hi[a]) return false;
      continue;
    }
    double t0 = (lo[a] - oo[a]) / dd[a], t1 = (hi[a] - oo[a]) / dd[a];
    if (t0 > t1) {
      const double sw = t0;
      t0 = t1;
      t1 = sw;
    }
    tmin = dmax(tmin, t0);
    tmax = dmin(tmax, t1);
  }
  if (tmax < tmin || tmin <= 1e-9) return false;
  t = tmin;
  return true;
}

// Ray from inside the room to its walls: exit distance and wall index.
KF_SCENE_FN bool hit_room(const SceneDesc& sc, D3 o, D3 d, double& t, int& wall) {
  const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  const double lo[3] = {sc.room_lo.x, sc.room_lo.y, sc.room_lo.z};
  const double hi[3] = {sc.room_hi.x, sc.room_hi.y, sc.room_hi.z};
  double best = 1e300;
  int bw = -1;
  for (int a = 0; a < 3; ++a) {
    if (dd[a] > 1e-15) {
      const double s = (hi[a] - oo[a]) / dd[a];
      if (s > 1e-9 && s < best) {
        best = s;
        bw = 2 * a + 1;
      }
    } else if (dd[a] < -1e-15) {
      const double s = (lo[a] - oo[a]) / dd[a];
      if (s > 1e-9 && s < best) {
        best = s;
        bw = 2 * a;
      }
    }
  }
  if (bw < 0) return false;
  t = best;
  wall = bw;
  return true;
}

// One pixel (u, v) of the view from camera_to_world `pose` (row-major 4x4).
// Writes the millimetre depth (0 = invalid), the colour (0 where nothing is
// hit) and the exact z before noise.
KF_SCENE_FN void render_pixel(const SceneDesc& sc, const double* pose, std::uint64_t noise_key, int noise, int u,
                              int v, std::uint16_t& depth_mm, std::uint32_t& rgba, double& exact_z) {
  const D3 o{pose[3], pose[7], pose[11]};
  const D3 dc{(u - sc.cx) / sc.fx, (v - sc.cy) / sc.fy, 1.0};
  const D3 d{pose[0] * dc.x + pose[1] * dc.y + pose[2] * dc.z, pose[4] * dc.x + pose[5] * dc.y + pose[6] * dc.z,
             pose[8] * dc.x + pose[9] * dc.y + pose[10] * dc.z};
  double best = 1e300;
  std::uint32_t col = 0;
  int wall = -1;
  double tw;
  if (hit_room(sc, o, d, tw, wall)) {
    best = tw;
    col = sc.wall_rgba[wall];
  }
  for (int i = 0; i < sc.n_spheres; ++i) {
    double t;
    if (hit_sphere(sc.spheres[i], o, d, t) && t < best) {
      best = t;
      col = sc.spheres[i].rgba;
    }
  }
  for (int i = 0; i < sc.n_boxes; ++i) {
    double t;
    if (hit_box(sc.boxes[i], o, d, t) && t < best) {
      best = t;
      col = sc.boxes[i].rgba;
    }
  }
  const std::uint64_t idx = static_cast<std::uint64_t>(v) * static_cast<std::uint64_t>(sc.width) +
                            static_cast<std::uint64_t>(u);
  // camera-frame z of the hit equals the ray parameter (dc.z == 1)
  const double z = best < 1e299 ? best : 0.0;
  exact_z = z;
  double zn = z;
  const std::uint64_t key =
      splitmix64(sc.seed ^ (noise_key * 0xD1B54A32D192ED03ULL) ^ (idx * 0x9E3779B97F4A7C15ULL));
  bool drop = false;
  if (noise != 0 && z > 0.0) {
    const std::uint64_t a = splitmix64(key), b = splitmix64(a), c = splitmix64(b);
    const double u1 = 1.0 - unit(a), u2 = unit(b);  // u1 in (0, 1]
    const double gauss = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    const double sigma = 0.0012 + 0.0019 * (z - 0.4) * (z - 0.4);
    zn = z + sigma * gauss;
    drop = unit(c) < 0.01;
  }
  double mm = floor(zn * 1000.0 + 0.5);
  if (!(mm > 0.0) || z <= 0.0 || drop) mm = 0.0;
  if (mm > 65535.0) mm = 0.0;
  depth_mm = static_cast<std::uint16_t>(mm);
  rgba = z > 0.0 ? col : 0U;
}

}  // namespace kfscene

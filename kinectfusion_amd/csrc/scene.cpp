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
// any order.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct D3 {
  double x, y, z;
};
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

inline std::uint64_t splitmix64(std::uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
inline double unit(std::uint64_t bits) {  // [0, 1) with 53 bits
  return static_cast<double>(bits >> 11) * (1.0 / 9007199254740992.0);
}

struct Rng {
  std::uint64_t state;
  std::uint64_t next() { return state = splitmix64(state); }
  double uniform(double lo, double hi) { return lo + (hi - lo) * unit(next()); }
};

struct Sphere {
  D3 c;
  double r;
  std::uint32_t rgba;
};
struct Box {
  D3 lo, hi;
  std::uint32_t rgba;
};

struct Scene {
  int kind;
  std::uint64_t seed;
  int width, height;
  double fx, fy, cx, cy;
  D3 room_lo, room_hi;
  std::uint32_t wall_rgba[6];
  std::vector<Sphere> spheres;
  std::vector<Box> boxes;
  // trajectory parameters
  double t_amp[3], t_freq[3], t_phase[3];
  double r_amp[3], r_freq[3], r_phase[3];
  double pan_rate;  // radians per frame about the vertical axis (hall sweep)
  double walk_rate; // metres per frame along x (hall sweep)
};

std::uint32_t random_rgba(Rng& rng) {
  const auto c = [&]() { return static_cast<std::uint32_t>(rng.uniform(40.0, 250.0)); };
  const std::uint32_t r = c(), g = c(), b = c();
  return r | (g << 8U) | (b << 16U) | (0xFFU << 24U);
}

// Ray/sphere: smallest s > eps with |o + s d - c| = r.
inline bool hit_sphere(const Sphere& s, D3 o, D3 d, double& t) {
  const D3 oc = o - s.c;
  const double b = dot(oc, d), c = dot(oc, oc) - s.r * s.r, a = dot(d, d);
  const double disc = b * b - a * c;
  if (disc < 0.0) return false;
  const double q = std::sqrt(disc);
  double s0 = (-b - q) / a;
  if (s0 <= 1e-9) s0 = (-b + q) / a;
  if (s0 <= 1e-9) return false;
  t = s0;
  return true;
}
// Ray/box (outside surface): slab entry.
inline bool hit_box(const Box& bx, D3 o, D3 d, double& t) {
  double tmin = -1e300, tmax = 1e300;
  const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  const double lo[3] = {bx.lo.x, bx.lo.y, bx.lo.z}, hi[3] = {bx.hi.x, bx.hi.y, bx.hi.z};
  for (int a = 0; a < 3; ++a) {
    if (std::fabs(dd[a]) < 1e-15) {
      if (oo[a] < lo[a] || oo[a] > hi[a]) return false;
      continue;
    }
    double t0 = (lo[a] - oo[a]) / dd[a], t1 = (hi[a] - oo[a]) / dd[a];
    if (t0 > t1) std::swap(t0, t1);
    tmin = std::max(tmin, t0);
    tmax = std::min(tmax, t1);
  }
  if (tmax < tmin || tmin <= 1e-9) return false;
  t = tmin;
  return true;
}
// Ray from inside the room to its walls: exit distance and wall index.
inline bool hit_room(const Scene& sc, D3 o, D3 d, double& t, int& wall) {
  const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  const double lo[3] = {sc.room_lo.x, sc.room_lo.y, sc.room_lo.z};
  const double hi[3] = {sc.room_hi.x, sc.room_hi.y, sc.room_hi.z};
  double best = 1e300;
  int bw = -1;
  for (int a = 0; a < 3; ++a) {
    if (dd[a] > 1e-15) {
      const double s = (hi[a] - oo[a]) / dd[a];
      if (s > 1e-9 && s < best) { best = s; bw = 2 * a + 1; }
    } else if (dd[a] < -1e-15) {
      const double s = (lo[a] - oo[a]) / dd[a];
      if (s > 1e-9 && s < best) { best = s; bw = 2 * a; }
    }
  }
  if (bw < 0) return false;
  t = best;
  wall = bw;
  return true;
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
//                    spheres; the camera pans and walks along x.
void* kf_scene_create(int kind, std::uint64_t seed, int width, int height, double fx, double fy,
                      double cx, double cy) {
  auto* sc = new Scene();
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
      sc->spheres.push_back(s);
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
    sc->pan_rate = 0.0;
    sc->walk_rate = 0.0;
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
        sc->boxes.push_back(b);
      } else {
        Sphere s{};
        s.c = c;
        s.r = rng.uniform(0.1, 0.35);
        s.rgba = random_rgba(rng);
        sc->spheres.push_back(s);
      }
    }
    for (int a = 0; a < 3; ++a) {
      sc->t_freq[a] = rng.uniform(0.02, 0.04);
      sc->t_amp[a] = 0.002 / sc->t_freq[a];
      sc->t_phase[a] = rng.uniform(0.0, 6.283185307179586);
      sc->r_freq[a] = rng.uniform(0.02, 0.04);
      sc->r_amp[a] = (0.05 * 3.141592653589793 / 180.0) / sc->r_freq[a];
      sc->r_phase[a] = rng.uniform(0.0, 6.283185307179586);
    }
    sc->pan_rate = 0.3 * 3.141592653589793 / 180.0;
    sc->walk_rate = 0.004;
  }
  return sc;
}

void kf_scene_destroy(void* s) { delete static_cast<Scene*>(s); }

// Ground-truth camera_to_world of frame k (row-major 4x4, FP64); frame 0 is
// the identity.
void kf_scene_pose(const void* s, int frame, double* pose) {
  const Scene& sc = *static_cast<const Scene*>(s);
  const double k = frame;
  double t[3], e[3];
  for (int a = 0; a < 3; ++a) {
    t[a] = sc.t_amp[a] * (std::sin(sc.t_freq[a] * k + sc.t_phase[a]) - std::sin(sc.t_phase[a]));
    e[a] = sc.r_amp[a] * (std::sin(sc.r_freq[a] * k + sc.r_phase[a]) - std::sin(sc.r_phase[a]));
  }
  // hall sweep: pan about y, walk along x (bounded back-and-forth)
  e[1] += sc.pan_rate * k;
  t[0] += sc.walk_rate * k;
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
  const Scene& sc = *static_cast<const Scene*>(s);
  const D3 o{pose[3], pose[7], pose[11]};
  const int W = sc.width, H = sc.height;
#pragma omp parallel for schedule(static)
  for (int v = 0; v < H; ++v) {
    for (int u = 0; u < W; ++u) {
      const D3 dc{(u - sc.cx) / sc.fx, (v - sc.cy) / sc.fy, 1.0};
      const D3 d{pose[0] * dc.x + pose[1] * dc.y + pose[2] * dc.z,
                 pose[4] * dc.x + pose[5] * dc.y + pose[6] * dc.z,
                 pose[8] * dc.x + pose[9] * dc.y + pose[10] * dc.z};
      double best = 1e300;
      std::uint32_t col = 0;
      int wall = -1;
      double tw;
      if (hit_room(sc, o, d, tw, wall)) {
        best = tw;
        col = sc.wall_rgba[wall];
      }
      for (const Sphere& sp : sc.spheres) {
        double t;
        if (hit_sphere(sp, o, d, t) && t < best) {
          best = t;
          col = sp.rgba;
        }
      }
      for (const Box& bx : sc.boxes) {
        double t;
        if (hit_box(bx, o, d, t) && t < best) {
          best = t;
          col = bx.rgba;
        }
      }
      const std::size_t i = static_cast<std::size_t>(v) * W + u;
      // camera-frame z of the hit equals the ray parameter (dc.z == 1)
      const double z = best < 1e299 ? best : 0.0;
      if (true_depth != nullptr) true_depth[i] = static_cast<float>(z);
      double zn = z;
      std::uint64_t key = splitmix64(sc.seed ^ (noise_key * 0xD1B54A32D192ED03ULL) ^
                                     (static_cast<std::uint64_t>(i) * 0x9E3779B97F4A7C15ULL));
      bool drop = false;
      if (noise != 0 && z > 0.0) {
        const std::uint64_t a = splitmix64(key), b = splitmix64(a), c = splitmix64(b);
        const double u1 = 1.0 - unit(a), u2 = unit(b);  // u1 in (0, 1]
        const double gauss = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        const double sigma = 0.0012 + 0.0019 * (z - 0.4) * (z - 0.4);
        zn = z + sigma * gauss;
        drop = unit(c) < 0.01;
      }
      double mm = std::floor(zn * 1000.0 + 0.5);
      if (!(mm > 0.0) || z <= 0.0 || drop) mm = 0.0;
      if (mm > 65535.0) mm = 0.0;
      depth_mm[i] = static_cast<std::uint16_t>(mm);
      if (rgba != nullptr) rgba[i] = z > 0.0 ? col : 0U;
    }
  }
}

}  // extern "C"

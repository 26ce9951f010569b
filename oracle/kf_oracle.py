"""Python face of the KinectFusion CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads oracle/libkf_oracle.so (built from oracle/kf_oracle.cpp by
oracle/Makefile) and exposes numpy-friendly wrappers.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may use this
module; the product (kinectfusion_amd) never imports it.

Each wrapper restates one reference CPU function; the C++ source cites the
reference file:line it follows.  `ate_umeyama` restates the trajectory
evaluator (src/app/trajectory_eval.cpp:77-128) in numpy.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB: C.CDLL | None = None

FLOAT, QUANTIZED, BF16 = 0, 1, 2
CLASSIC, ANGLE = 0, 1


class Geometry(C.Structure):
    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64), ("voxel_size", C.c_float),
                ("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float), ("truncation", C.c_float)]


class Intrinsics(C.Structure):
    _fields_ = [("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float)]


class IntegrationOptions(C.Structure):
    _fields_ = [("depth_scale", C.c_float), ("observation_weight", C.c_float), ("max_weight", C.c_float),
                ("min_depth", C.c_float), ("max_depth", C.c_float), ("projective_distance", C.c_int),
                ("distance_scaled_truncation", C.c_int), ("truncation_distance_scale", C.c_float)]


class IcpOptions(C.Structure):
    _fields_ = [("min_correspondences", C.c_int64), ("max_point_distance", C.c_float),
                ("min_normal_dot", C.c_float), ("min_update_translation", C.c_float),
                ("min_update_rotation", C.c_float), ("max_update_translation", C.c_float),
                ("max_update_rotation", C.c_float), ("min_system_eigenvalue", C.c_float),
                ("max_condition_number", C.c_float), ("damping_mode", C.c_int), ("damping_lambda", C.c_float),
                ("adaptive_damping", C.c_int), ("lambda_increase", C.c_float), ("lambda_decrease", C.c_float),
                ("min_lambda", C.c_float), ("max_lambda", C.c_float)]


class IcpResult(C.Structure):
    _fields_ = [("status", C.c_int), ("trials", C.c_int), ("rejected", C.c_int),
                ("correspondences", C.c_int64), ("mean_distance", C.c_double),
                ("mean_squared_residual", C.c_double), ("min_eigenvalue", C.c_double),
                ("condition_number", C.c_double), ("update_translation", C.c_double),
                ("update_rotation", C.c_double), ("lambda_", C.c_double)]


def icp_options(**kw) -> IcpOptions:
    """ProjectiveIcpOptions defaults (icp_optimizer.hpp:14-20, icp_correspondence.hpp:25-34)."""
    o = IcpOptions(64, 0.05, 0.9659258, 1e-5, 1e-5, 0.15, 0.35, 1e-6, 1e6, 0, 1e-3, 0, 10.0, 0.1, 1e-9, 1e6)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def integration_options(depth_scale=5000.0, **kw) -> IntegrationOptions:
    """TsdfIntegrationOptions defaults (tsdf_integration.hpp:31-41)."""
    o = IntegrationOptions(depth_scale, 1.0, 196.0, 0.4, 8.0, 1, 0, 0.01)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def build() -> Path:
    """Compile libkf_oracle.so with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return HERE / "libkf_oracle.so"


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        path = HERE / "libkf_oracle.so"
        if not path.exists():
            build()
        _LIB = C.CDLL(str(path))
        _declare(_LIB)
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _declare(L: C.CDLL) -> None:
    vp, i32, i64, f32 = C.c_void_p, C.c_int, C.c_int64, C.c_float
    L.kfo_volume_reset.argtypes = [C.POINTER(Geometry), i32, vp, vp]
    L.kfo_voxels_reset.argtypes = [i32, vp, vp, i64]
    L.kfo_integrate_band.argtypes = [C.POINTER(Geometry), i32, i32, vp, vp, vp, i32, i32, vp, vp,
                                     C.POINTER(Intrinsics), vp, C.POINTER(IntegrationOptions), vp]
    L.kfo_integrate.argtypes = [C.POINTER(Geometry), i32, i32, vp, vp, vp, i32, i32, vp, vp,
                                C.POINTER(Intrinsics), vp, C.POINTER(IntegrationOptions), vp]
    L.kfo_observed_count.argtypes = [C.POINTER(Geometry), i32, vp]
    L.kfo_observed_count.restype = i64
    L.kfo_bilateral.argtypes = [vp, i32, i32, f32, f32, f32, i32, f32, f32, vp]
    L.kfo_downsample.argtypes = [vp, i32, i32, f32, f32, vp]
    L.kfo_vertices.argtypes = [vp, i32, i32, C.POINTER(Intrinsics), f32, f32, f32, vp]
    L.kfo_normals.argtypes = [vp, i32, i32, f32, vp]
    L.kfo_raycast.argtypes = [C.POINTER(Geometry), i32, vp, vp, vp, vp, C.POINTER(Intrinsics), i32, i32,
                              f32, f32, f32, i32, i32, vp, vp, vp, vp, vp]
    L.kfo_icp_sums.argtypes = [vp, vp, i32, i32, vp, vp, i32, i32, C.POINTER(Intrinsics), vp, vp, f32, f32,
                               vp, vp]
    L.kfo_icp_sums.restype = i64
    L.kfo_sym_eig6.argtypes = [vp, vp]
    L.kfo_compose_increment.argtypes = [vp, vp]
    L.kfo_icp_track_level.argtypes = [vp, vp, i32, i32, vp, vp, i32, i32, C.POINTER(Intrinsics), vp, vp, i32,
                                      C.POINTER(IcpOptions), C.POINTER(IcpResult)]
    L.kfo_mesh_extract.argtypes = [C.POINTER(Geometry), i32, vp, vp, f32]
    L.kfo_mesh_extract.restype = vp
    L.kfo_mesh_extract_traversal.argtypes = [C.POINTER(Geometry), i32, vp, vp, f32, vp]
    L.kfo_mesh_extract_traversal.restype = vp
    L.kfo_band_walk.argtypes = [C.POINTER(Geometry), vp, i32, i32, C.POINTER(Intrinsics), vp,
                                C.POINTER(IntegrationOptions), f32, vp, i64]
    L.kfo_band_walk.restype = i64
    L.kfo_integrate_sparse.argtypes = [C.POINTER(Geometry), i32, i32, vp, vp, vp, vp, vp, i32, i32, vp, vp,
                                       C.POINTER(Intrinsics), vp, C.POINTER(IntegrationOptions), vp]
    L.kfo_mesh_sizes.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
    L.kfo_mesh_copy.argtypes = [vp, vp, vp, vp, vp]
    L.kfo_mesh_free.argtypes = [vp]
    L.kfo_q_encode.argtypes = [f32]
    L.kfo_q_encode.restype = C.c_int16
    L.kfo_q_decode.argtypes = [C.c_int16]
    L.kfo_q_decode.restype = f32
    L.kfo_bf_encode.argtypes = [f32]
    L.kfo_bf_encode.restype = C.c_uint16
    L.kfo_bf_decode.argtypes = [C.c_uint16]
    L.kfo_bf_decode.restype = f32
    L.kfo_w_encode.argtypes = [f32]
    L.kfo_w_encode.restype = C.c_uint16
    L.kfo_w_decode.argtypes = [C.c_uint16]
    L.kfo_w_decode.restype = f32


@dataclass
class Volume:
    """Dense host volume in the reference layout (x fastest)."""
    geometry: Geometry
    kind: int
    voxels: np.ndarray          # float32 (n, 2) or uint16 (n, 2)
    colors: np.ndarray | None   # float32 (n, 4) or None

    @staticmethod
    def create(res, voxel_size, origin, truncation, kind=FLOAT, color=True) -> "Volume":
        nx, ny, nz = (res, res, res) if np.isscalar(res) else res
        g = Geometry(nx, ny, nz, voxel_size, origin[0], origin[1], origin[2], truncation)
        n = nx * ny * nz
        vox = np.empty((n, 2), np.float32 if kind == FLOAT else np.uint16)
        col = np.empty((n, 4), np.float32) if color else None
        v = Volume(g, kind, vox, col)
        lib().kfo_volume_reset(C.byref(g), kind, _p(vox), _p(col))
        return v

    def decoded(self) -> tuple[np.ndarray, np.ndarray]:
        """(tsdf, weight) as float32 per voxel, through the store's codec."""
        if self.kind == FLOAT:
            return self.voxels[:, 0].copy(), self.voxels[:, 1].copy()
        w = self.voxels[:, 1].astype(np.float32) * np.float32(1.0 / 256.0)
        if self.kind == QUANTIZED:
            t = self.voxels[:, 0].view(np.int16).astype(np.float32) * np.float32(1.0 / 32767.0)
        else:
            t = (self.voxels[:, 0].astype(np.uint32) << 16).view(np.float32)
        return t, w


def pose12(m: np.ndarray) -> np.ndarray:
    """Row-major 3x4 float32 of a 4x4 (or 3x4) transform."""
    return np.ascontiguousarray(np.asarray(m, np.float64)[:3, :4].astype(np.float32).reshape(12))


def rigid_inverse(m: np.ndarray) -> np.ndarray:
    """[R^T, -R^T t] in FP64."""
    m = np.asarray(m, np.float64)
    out = np.eye(4)
    out[:3, :3] = m[:3, :3].T
    out[:3, 3] = -(out[:3, :3] @ m[:3, 3])
    return out


def integrate(vol: Volume, rule: int, depth: np.ndarray, intr, world_to_camera, opts=None,
              color: np.ndarray | None = None, normals: np.ndarray | None = None) -> tuple[int, int]:
    """One full-sweep integration; returns (voxels fused, colour voxels fused)."""
    h, w = depth.shape
    depth = np.ascontiguousarray(depth, np.uint16)
    if color is not None:
        color = np.ascontiguousarray(color, np.uint32)
    if normals is not None:
        normals = np.ascontiguousarray(normals[..., :3], np.float32)
    opts = opts or integration_options()
    k = Intrinsics(*intr)
    xf = pose12(world_to_camera)
    counts = np.zeros(2, np.int64)
    lib().kfo_integrate(C.byref(vol.geometry), vol.kind, rule, _p(vol.voxels), _p(vol.colors), _p(depth), w, h,
                        _p(color), _p(normals), C.byref(k), _p(xf), C.byref(opts), _p(counts))
    return int(counts[0]), int(counts[1])


class SparseVolume(Volume):
    """BlockRep<kHost> (block_rep.hpp:183-307): the slot grid (slots[block]
    = pool slot or -1, flat x-major block order), the allocation state
    [requested, capacity] and a pool of capacity x 512 voxels (`voxels`,
    `colors`), each slot's voxels in intra-block order
    (z % 8 * 8 + y % 8) * 8 + x % 8."""
    slots: np.ndarray
    state: np.ndarray

    @staticmethod
    def create_sparse(res, voxel_size, origin, truncation, kind=FLOAT, color=True, capacity=None) -> "SparseVolume":
        nx, ny, nz = (res, res, res) if np.isscalar(res) else res
        g = Geometry(nx, ny, nz, voxel_size, origin[0], origin[1], origin[2], truncation)
        nb = ((nx + 7) // 8) * ((ny + 7) // 8) * ((nz + 7) // 8)
        cap = capacity if capacity is not None else nb // 4
        n = cap * 512
        vox = np.empty((n, 2), np.float32 if kind == FLOAT else np.uint16)
        col = np.empty((n, 4), np.float32) if color else None
        lib().kfo_voxels_reset(kind, _p(vox), _p(col), n)
        v = SparseVolume(g, kind, vox, col)
        v.slots = np.full(nb, -1, np.int32)
        v.state = np.array([0, cap], np.int64)
        return v

    def blocks(self) -> tuple[np.ndarray, np.ndarray, np.ndarray | None]:
        """(keys increasing, voxels (n, 512, 2), colours (n, 512, 4) or None)
        of the allocated blocks."""
        keys = np.nonzero(self.slots >= 0)[0]
        slot = self.slots[keys].astype(np.int64)
        vox = self.voxels.reshape(-1, 512, 2)[slot]
        col = self.colors.reshape(-1, 512, 4)[slot] if self.colors is not None else None
        return keys.astype(np.uint32), vox, col

    def dense(self) -> Volume:
        """The dense expansion (BlockRep::host_dense_view): unallocated
        voxels fresh.  For small grids only."""
        g = self.geometry
        out = Volume.create((g.nx, g.ny, g.nz), g.voxel_size, (g.ox, g.oy, g.oz), g.truncation, self.kind,
                            self.colors is not None)
        idx = np.arange(g.nx * g.ny * g.nz)
        x, y, z = idx % g.nx, (idx // g.nx) % g.ny, idx // (g.nx * g.ny)
        bx, by = (g.nx + 7) // 8, (g.ny + 7) // 8
        slot = self.slots[((z // 8) * by + (y // 8)) * bx + (x // 8)].astype(np.int64)
        have = slot >= 0
        at = slot[have] * 512 + ((z[have] % 8) * 8 + (y[have] % 8)) * 8 + (x[have] % 8)
        out.voxels[have] = self.voxels[at]
        if out.colors is not None:
            out.colors[have] = self.colors[at]
        return out


def integrate_sparse(vol: SparseVolume, rule: int, depth: np.ndarray, intr, world_to_camera, opts=None,
                     color: np.ndarray | None = None, normals: np.ndarray | None = None) -> tuple[int, int, int]:
    """BlockRep::integrate: (voxels fused, colour voxels fused, blocks listed)."""
    h, w = depth.shape
    depth = np.ascontiguousarray(depth, np.uint16)
    if color is not None:
        color = np.ascontiguousarray(color, np.uint32)
    if normals is not None:
        normals = np.ascontiguousarray(normals[..., :3], np.float32)
    opts = opts or integration_options()
    k = Intrinsics(*intr)
    xf = pose12(world_to_camera)
    counts = np.zeros(3, np.int64)
    lib().kfo_integrate_sparse(C.byref(vol.geometry), vol.kind, rule, _p(vol.voxels), _p(vol.colors),
                               _p(vol.slots), _p(vol.state), _p(depth), w, h, _p(color), _p(normals), C.byref(k),
                               _p(xf), C.byref(opts), _p(counts))
    return int(counts[0]), int(counts[1]), int(counts[2])


def band_walk(geometry: Geometry, depth: np.ndarray, intr, world_to_camera, opts=None, step=None) -> np.ndarray:
    """TruncationBandWalk::visit over all pixels: visited block indices in order."""
    h, w = depth.shape
    depth = np.ascontiguousarray(depth, np.uint16)
    opts = opts or integration_options()
    step = step if step is not None else 0.5 * 8 * geometry.voxel_size
    k = Intrinsics(*intr)
    xf = pose12(world_to_camera)
    n = lib().kfo_band_walk(C.byref(geometry), _p(depth), w, h, C.byref(k), _p(xf), C.byref(opts), step, None, 0)
    out = np.empty(max(n, 1), np.int64)
    lib().kfo_band_walk(C.byref(geometry), _p(depth), w, h, C.byref(k), _p(xf), C.byref(opts), step, _p(out), n)
    return out[:n]


def integrate_band(vol: Volume, rule: int, depth: np.ndarray, intr, world_to_camera, opts=None,
                   color: np.ndarray | None = None, normals: np.ndarray | None = None) -> tuple[int, int, int]:
    """BandIntegrationSweep<kHost> (dense kBand mode): (voxels fused, colour
    voxels fused, blocks listed)."""
    h, w = depth.shape
    depth = np.ascontiguousarray(depth, np.uint16)
    if color is not None:
        color = np.ascontiguousarray(color, np.uint32)
    if normals is not None:
        normals = np.ascontiguousarray(normals[..., :3], np.float32)
    opts = opts or integration_options()
    counts = np.zeros(3, np.int64)
    lib().kfo_integrate_band(C.byref(vol.geometry), vol.kind, rule, _p(vol.voxels), _p(vol.colors), _p(depth), w, h,
                             _p(color), _p(normals), C.byref(Intrinsics(*intr)), _p(pose12(world_to_camera)),
                             C.byref(opts), _p(counts))
    return int(counts[0]), int(counts[1]), int(counts[2])


def observed_count(vol: Volume) -> int:
    if isinstance(vol, SparseVolume):
        keys, vox, _ = vol.blocks()
        w = vox[..., 1]
        return int(np.count_nonzero(w > 0) if vol.kind == FLOAT else np.count_nonzero(w != 0))
    return int(lib().kfo_observed_count(C.byref(vol.geometry), vol.kind, _p(vol.voxels)))


def bilateral(depth, scale, lo, hi, radius, sigma_s, sigma_r) -> np.ndarray:
    depth = np.ascontiguousarray(depth, np.uint16)
    out = np.empty_like(depth)
    lib().kfo_bilateral(_p(depth), depth.shape[1], depth.shape[0], scale, lo, hi, radius, sigma_s, sigma_r,
                        _p(out))
    return out


def downsample(depth, scale, max_jump=0.1) -> np.ndarray:
    depth = np.ascontiguousarray(depth, np.uint16)
    h, w = depth.shape
    out = np.empty((h // 2, w // 2), np.uint16)
    lib().kfo_downsample(_p(depth), w, h, scale, max_jump, _p(out))
    return out


def vertices(depth, intr, scale, lo=0.4, hi=8.0) -> np.ndarray:
    depth = np.ascontiguousarray(depth, np.uint16)
    h, w = depth.shape
    out = np.empty((h, w, 3), np.float32)
    lib().kfo_vertices(_p(depth), w, h, C.byref(Intrinsics(*intr)), scale, lo, hi, _p(out))
    return out


def normals(verts, max_jump=0.1) -> np.ndarray:
    verts = np.ascontiguousarray(verts[..., :3], np.float32)
    h, w = verts.shape[:2]
    out = np.empty((h, w, 3), np.float32)
    lib().kfo_normals(_p(verts), w, h, max_jump, _p(out))
    return out


def pyramid(depth, intr, scale, levels=3, lo=0.4, hi=8.0, radius=2, sigma_s=2.0, sigma_r=0.08,
            bilateral_filter=True, max_normal_jump=0.1, max_down_jump=0.1):
    """DepthProcessor::build_surface_pyramid (depth_processing.cpp:138-163)."""
    cur = bilateral(depth, scale, lo, hi, radius, sigma_s, sigma_r) if (bilateral_filter and radius > 0) \
        else np.array(depth, np.uint16)
    out = []
    for level in range(levels):
        s = 1.0 / (1 << level)
        k = (np.float32(intr[0]) * np.float32(s), np.float32(intr[1]) * np.float32(s),
             np.float32(intr[2]) * np.float32(s), np.float32(intr[3]) * np.float32(s))
        v = vertices(cur, k, scale, lo, hi)
        n = normals(v, max_normal_jump)
        out.append({"depth": cur, "intrinsics": k, "vertices": v, "normals": n})
        if level + 1 < levels:
            cur = downsample(cur, scale, max_down_jump)
    return out


def raycast(vol: Volume, cam_to_world, intr, width, height, min_depth=0.4, max_depth=8.0, step_scale=1.0,
            require_all=False, cell_normals=False, previous_points=None):
    """SurfaceRaycast over every pixel: (points, normals, colours, samples).
    previous_points ((H, W, 3) float32): seed_from_previous with these as the
    maps' previous surface points."""
    pts = np.empty((height, width, 3), np.float32)
    nrm = np.empty((height, width, 3), np.float32)
    col = np.empty((height, width), np.uint32)
    samples = np.zeros(1, np.int64)
    xf = pose12(cam_to_world)
    prev = None if previous_points is None else np.ascontiguousarray(previous_points[..., :3], np.float32)
    slots = vol.slots if isinstance(vol, SparseVolume) else None
    lib().kfo_raycast(C.byref(vol.geometry), vol.kind, _p(vol.voxels), _p(vol.colors), _p(slots), _p(xf),
                      C.byref(Intrinsics(*intr)), width, height, min_depth, max_depth, step_scale,
                      int(require_all), int(cell_normals), _p(prev), _p(pts), _p(nrm), _p(col), _p(samples))
    return pts, nrm, col, int(samples[0])


def icp_sums(live_v, live_n, model_v, model_n, intr, camera, model, max_dist=0.05, min_dot=0.9659258):
    """(sums64[29], sums32[29], count) of one correspondence sweep."""
    lv = np.ascontiguousarray(live_v[..., :3], np.float32)
    ln = np.ascontiguousarray(live_n[..., :3], np.float32)
    mv = np.ascontiguousarray(model_v[..., :3], np.float32)
    mn = np.ascontiguousarray(model_n[..., :3], np.float32)
    s64 = np.zeros(29, np.float64)
    s32 = np.zeros(29, np.float32)
    n = lib().kfo_icp_sums(_p(lv), _p(ln), lv.shape[1], lv.shape[0], _p(mv), _p(mn), mv.shape[1], mv.shape[0],
                           C.byref(Intrinsics(*intr)), _p(pose12(camera)), _p(pose12(model)), max_dist, min_dot,
                           _p(s64), _p(s32))
    return s64, s32, int(n)


def sym_eig6(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float64)
    ev = np.zeros(6)
    lib().kfo_sym_eig6(_p(a), _p(ev))
    return ev


def compose_increment(x, pose) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    p = np.ascontiguousarray(np.array(pose, np.float64).reshape(16))
    lib().kfo_compose_increment(_p(x), _p(p))
    return p.reshape(4, 4)


def icp_track_level(live_v, live_n, model_v, model_n, intr, model_pose, initial_pose, iterations,
                    opts: IcpOptions | None = None):
    """Device-loop ICP of one level; returns (pose 4x4 FP64, IcpResult)."""
    lv = np.ascontiguousarray(live_v[..., :3], np.float32)
    ln = np.ascontiguousarray(live_n[..., :3], np.float32)
    mv = np.ascontiguousarray(model_v[..., :3], np.float32)
    mn = np.ascontiguousarray(model_n[..., :3], np.float32)
    mp = np.ascontiguousarray(np.asarray(model_pose, np.float64).reshape(16))
    pose = np.ascontiguousarray(np.asarray(initial_pose, np.float64).reshape(16)).copy()
    res = IcpResult()
    lib().kfo_icp_track_level(_p(lv), _p(ln), lv.shape[1], lv.shape[0], _p(mv), _p(mn), mv.shape[1],
                              mv.shape[0], C.byref(Intrinsics(*intr)), _p(mp), _p(pose), iterations,
                              C.byref(opts or icp_options()), C.byref(res))
    return pose.reshape(4, 4), res


def mesh(vol: Volume, min_weight=0.0):
    """MarchingCubes::extract -> (positions, normals, colors, triangles);
    the sparse traversal for a SparseVolume."""
    L = lib()
    slots = vol.slots if isinstance(vol, SparseVolume) else None
    h = L.kfo_mesh_extract_traversal(C.byref(vol.geometry), vol.kind, _p(vol.voxels), _p(vol.colors), min_weight,
                                     _p(slots))
    nv, nt = C.c_int64(), C.c_int64()
    L.kfo_mesh_sizes(h, C.byref(nv), C.byref(nt))
    pos = np.empty((nv.value, 3), np.float32)
    nrm = np.empty((nv.value, 3), np.float32)
    col = np.empty((nv.value, 3), np.float32)
    tri = np.empty((nt.value, 3), np.uint32)
    L.kfo_mesh_copy(h, _p(pos), _p(nrm), _p(col), _p(tri))
    L.kfo_mesh_free(h)
    return pos, nrm, col, tri


def ate_umeyama(estimated: np.ndarray, ground: np.ndarray) -> dict:
    """Rigid (no scale) Umeyama alignment of estimated to ground-truth
    positions and the ATE statistics (AteEvaluator::evaluate,
    src/app/trajectory_eval.cpp:77-128)."""
    src = np.asarray(estimated, np.float64)
    dst = np.asarray(ground, np.float64)
    mu_s, mu_d = src.mean(0), dst.mean(0)
    cov = (dst - mu_d).T @ (src - mu_s) / len(src)
    u, _, vt = np.linalg.svd(cov)
    s = np.eye(3)
    if np.linalg.det(u) * np.linalg.det(vt) < 0:
        s[2, 2] = -1
    r = u @ s @ vt
    t = mu_d - r @ mu_s
    err = np.linalg.norm((src @ r.T + t) - dst, axis=1)
    return {"pairs": len(src), "rmse": float(np.sqrt(np.mean(err ** 2))), "mean": float(err.mean()),
            "median": float(np.median(err)), "max": float(err.max())}

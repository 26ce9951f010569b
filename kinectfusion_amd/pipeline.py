"""Host-side mirror of the reference's stage interfaces over libkf_hip.

Names, argument meaning and error behaviour follow the reference:
  PipelineConfig / Pipeline        include/kinectfusion/pipeline.hpp:29-161
  PyramidSource / PyramidLevel     include/kinectfusion/depth_processing.hpp:415-463
  ProjectiveIcpTracker, IcpOutcome include/kinectfusion/icp_optimizer.hpp:22-130
  MarchingCubes                    include/kinectfusion/marching_cubes.hpp:16-34
  Reconstruction (frame driver)    src/app/reconstruction.cpp:143-250

The only memory space is "hip".  Asking for "cpu" raises (the reference's
std::logic_error for an unbuilt backend); a missing GPU or library raises
too -- nothing falls back to the host.  Device buffers are torch tensors on
cuda:0; torch is plumbing here, every stage runs in libkf_hip.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise N.KinectFusionError(4, "no HIP device is available; the hip space has no host fallback")
    return torch


@dataclass
class CameraIntrinsics:
    """rgbd.hpp:28-71."""
    fx: float
    fy: float
    cx: float
    cy: float

    def scaled(self, level: int) -> "CameraIntrinsics":
        s = np.float32(1.0) / np.float32(1 << level)
        return CameraIntrinsics(float(np.float32(self.fx) * s), float(np.float32(self.fy) * s),
                                float(np.float32(self.cx) * s), float(np.float32(self.cy) * s))


@dataclass
class VolumeGeometry:
    """volume.hpp:265-287."""
    resolution: tuple[int, int, int] = (512, 512, 512)
    voxel_size: float = 0.01
    origin: tuple[float, float, float] = (-2.56, -2.56, -2.56)
    truncation_distance: float = 0.05

    @staticmethod
    def centered(resolution: int, voxel_size: float, camera_margin: float, truncation: float) -> "VolumeGeometry":
        """AppOptions::volume_origin (src/app/app_options.cpp:33-38)."""
        half = np.float32(0.5) * np.float32(resolution) * np.float32(voxel_size)
        return VolumeGeometry((resolution,) * 3, voxel_size, (-float(half), -float(half), -camera_margin),
                              truncation)


@dataclass
class TsdfIntegrationOptions:
    """tsdf_integration.hpp:31-50."""
    depth_scale: float = 5000.0
    observation_weight: float = 1.0
    max_weight: float = 196.0
    min_depth: float = 0.4
    max_depth: float = 8.0
    projective_distance: bool = True
    distance_scaled_truncation: bool = False
    truncation_distance_scale: float = 0.01
    mode: str = "full"                         # IntegrationMode (:29): full | band (dense storage)


@dataclass
class RaycastOptions:
    """raycasting.hpp:66-78."""
    min_depth: float = 0.4
    max_depth: float = 8.0
    step_scale: float = 1.0
    require_all_corners: bool = False
    cell_gradient_normals: bool = False
    seed_from_previous: bool = False


@dataclass
class DepthProcessingOptions:
    """depth_processing.hpp:31-46."""
    levels: int = 3
    max_normal_depth_jump: float = 0.1
    max_downsample_depth_jump: float = 0.1
    bilateral_filter: bool = True
    bilateral_radius: int = 2
    bilateral_spatial_sigma: float = 2.0
    bilateral_depth_sigma: float = 0.08


@dataclass
class PipelineConfig:
    """pipeline.hpp:29-46 plus the camera, pyramid and tracking settings the
    reference's app wires in (src/app/app_options.cpp)."""
    name: str = "baseline"
    space: str = "hip"
    tsdf_rule: str = "angle-weighted"          # --tsdf-variant
    voxel: str = "float"                       # --voxel float|quantized|bf16
    color: str = "float"                       # --color float|none
    storage: str = "dense"
    # RaycastBackend (raycasting.hpp:22): "march" and "bitmap-march" render
    # identical maps here (the march skips empty blocks exactly);
    # "band-march" (approximate in the reference) is not provided
    raycast_backend: str = "march"
    volume: VolumeGeometry = field(default_factory=VolumeGeometry)
    integration: TsdfIntegrationOptions = field(default_factory=TsdfIntegrationOptions)
    raycast: RaycastOptions = field(default_factory=RaycastOptions)
    depth: DepthProcessingOptions = field(default_factory=DepthProcessingOptions)
    intrinsics: CameraIntrinsics = field(default_factory=lambda: CameraIntrinsics(525.0, 525.0, 319.5, 239.5))
    width: int = 640
    height: int = 480
    init: str = "previous"                     # previous | constant-velocity | learned
    icp_iterations: tuple[int, ...] = (10, 5, 4)
    stage_timing: bool = False
    count_fused: bool = False
    # hashed storage (BlockRep, block_rep.hpp:183-307)
    hash_buckets: int = 1 << 20
    block_capacity: int = 500_000              # 0: a quarter of the grid's blocks (pipeline.hpp:40)
    # learned warm start: iterations per level from the network's proposal
    # (0 skips a level); the frame falls back to the constant-velocity seed
    # with icp_iterations when the finest level's final RMS (m) or its last
    # increment (m, rad: the run needed more iterations) exceeds these
    learned_iterations: tuple[int, ...] = (4, 0, 0)
    learned_max_rms: float = 0.01
    learned_max_update_translation: float = 1.0e-3
    learned_max_update_rotation: float = 1.0e-3

    def native(self) -> N.KfConfig:
        if self.space != "hip":
            raise NotImplementedError(f"memory space '{self.space}' is not built; this library provides 'hip' only")
        if self.raycast_backend not in ("march", "bitmap-march"):
            raise NotImplementedError(f"raycast backend '{self.raycast_backend}' is not provided; 'march' and "
                                      "'bitmap-march' render identical maps")
        # Pipeline::require_valid_storage (src/kinectfusion/pipeline.cpp:108-117)
        if self.storage == "hashed" and self.integration.mode != "band":
            raise ValueError("Sparse storage implies band integration (nothing is allocated far from surfaces); "
                             "set integration = 'band'")
        if self.storage == "hashed" and self.raycast_backend != "march":
            raise ValueError("Sparse storage supports the plain march raycast backend")
        c = N.KfConfig()
        N.hip().kf_config_defaults(C.byref(c))
        c.resolution[:] = list(self.volume.resolution)
        c.voxel_size = self.volume.voxel_size
        c.origin[:] = list(self.volume.origin)
        c.truncation_distance = self.volume.truncation_distance
        c.voxel_store = {"float": N.VOXEL_FLOAT, "quantized": N.VOXEL_QUANTIZED, "bf16": N.VOXEL_BF16}[self.voxel]
        c.color_store = {"float": N.COLOR_FLOAT, "none": N.COLOR_NONE}[self.color]
        c.storage = {"dense": N.STORAGE_DENSE, "hashed": N.STORAGE_HASHED}[self.storage]
        c.tsdf_rule = {"classic": N.RULE_CLASSIC, "angle-weighted": N.RULE_ANGLE_WEIGHTED}[self.tsdf_rule]
        i = self.integration
        c.depth_scale, c.observation_weight, c.max_weight = i.depth_scale, i.observation_weight, i.max_weight
        c.min_depth, c.max_depth = i.min_depth, i.max_depth
        c.projective_distance = int(i.projective_distance)
        c.distance_scaled_truncation = int(i.distance_scaled_truncation)
        c.truncation_distance_scale = i.truncation_distance_scale
        c.integration_mode = {"full": N.INTEGRATION_FULL, "band": N.INTEGRATION_BAND}[i.mode]
        r = self.raycast
        c.ray_min_depth, c.ray_max_depth, c.step_scale = r.min_depth, r.max_depth, r.step_scale
        c.require_all_corners, c.cell_normals = int(r.require_all_corners), int(r.cell_gradient_normals)
        c.seed_from_previous = int(r.seed_from_previous)
        d = self.depth
        c.levels = d.levels
        c.max_normal_depth_jump, c.max_downsample_depth_jump = d.max_normal_depth_jump, d.max_downsample_depth_jump
        c.bilateral_filter, c.bilateral_radius = int(d.bilateral_filter), d.bilateral_radius
        c.bilateral_spatial_sigma, c.bilateral_depth_sigma = d.bilateral_spatial_sigma, d.bilateral_depth_sigma
        c.width, c.height = self.width, self.height
        k = self.intrinsics
        c.fx, c.fy, c.cx, c.cy = k.fx, k.fy, k.cx, k.cy
        c.init_mode = {"previous": N.INIT_PREVIOUS, "constant-velocity": N.INIT_CONSTANT_VELOCITY,
                       "learned": N.INIT_LEARNED}[self.init]
        its = list(self.icp_iterations) + [4] * (4 - len(self.icp_iterations))
        c.icp_iterations[:] = its[:4]
        c.stage_timing = int(self.stage_timing)
        c.count_fused = int(self.count_fused)
        c.hash_buckets, c.block_capacity = self.hash_buckets, self.block_capacity
        lits = list(self.learned_iterations) + [0] * (4 - len(self.learned_iterations))
        c.learned_iterations[:] = lits[:4]
        c.learned_max_rms = self.learned_max_rms
        c.learned_max_update_translation = self.learned_max_update_translation
        c.learned_max_update_rotation = self.learned_max_update_rotation
        return c


@dataclass
class ProjectiveIcpOptions:
    """icp_optimizer.hpp:82-102."""
    min_correspondences: int = 64
    max_point_distance: float = 0.05
    min_normal_dot: float = 0.9659258
    min_update_translation: float = 1.0e-5
    min_update_rotation: float = 1.0e-5
    max_update_translation: float = 0.15
    max_update_rotation: float = 0.35
    min_system_eigenvalue: float = 1.0e-6
    max_condition_number: float = 1.0e6
    damping: str = "none"                 # none | identity | diagonal
    damping_lambda: float = 1.0e-3
    adaptive_damping: bool = False
    lambda_increase: float = 10.0
    lambda_decrease: float = 0.1
    min_lambda: float = 1.0e-9
    max_lambda: float = 1.0e6

    def native(self) -> N.KfIcpOptions:
        o = N.KfIcpOptions()
        for name, _ in N.KfIcpOptions._fields_:
            if name == "damping_mode":
                o.damping_mode = {"none": 0, "identity": 1, "diagonal": 2}[self.damping]
            elif name == "adaptive_damping":
                o.adaptive_damping = int(self.adaptive_damping)
            else:
                setattr(o, name, getattr(self, name))
        return o


FAILURES = {2: "kTooFewCorrespondences", 3: "kSolveFailed", 4: "kUpdateTooLarge", 5: "kUnconstrainedSystem"}


@dataclass
class IcpOutcome:
    """icp_optimizer.hpp:55-62: the pose is always present; `result` is
    "Converged", "MaxIterations" or the failure name."""
    pose: np.ndarray
    result: str
    diagnostics: N.KfIcpResult

    @property
    def ok(self) -> bool:
        return self.result in ("Converged", "MaxIterations")


def _outcome(pose: np.ndarray, r: N.KfIcpResult) -> IcpOutcome:
    name = {0: "MaxIterations", 1: "Converged"}.get(r.status, FAILURES.get(r.status, "skipped"))
    return IcpOutcome(pose, name, r)


def _check_image(t, name: str, dtypes, shape, optional: bool = False):
    """A device image the library reads through a raw pointer: a contiguous
    cuda tensor of one of `dtypes` and exactly `shape` (the reference's
    std::invalid_argument for an invalid frame, validation.hpp:9-13)."""
    if t is None:
        if optional:
            return None
        raise ValueError(f"{name}: missing")
    if not getattr(t, "is_cuda", False):
        raise ValueError(f"{name}: must be a device (cuda) tensor; the hip space reads device memory only")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be contiguous")
    if str(t.dtype).replace("torch.", "") not in dtypes:
        raise ValueError(f"{name}: dtype {t.dtype} is not one of {sorted(dtypes)}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)} differs from {tuple(shape)}")
    return t


_DEPTH = ("int16", "uint16")
_COLOR = ("int32", "uint32")
_FLOAT = ("float32",)


@dataclass
class SurfaceMaps:
    """volume.hpp:208-219 on the device: points/normals (H, W, 4) float32,
    colors (H, W) int32 RGBA8."""
    points: object
    normals: object
    colors: object


class Pipeline:
    """One reconstruction pipeline in the hip space (pipeline.hpp:64-161)."""

    def __init__(self, config: PipelineConfig):
        self.config = config
        native = config.native()
        self._torch = _torch()
        ctx = C.c_void_p()
        N.check(N.hip().kf_create(C.byref(native), C.byref(ctx)), None)
        self._ctx = ctx
        self._native = native
        # device inputs of enqueued work stay referenced until the next
        # synchronisation, so torch cannot recycle their memory early
        self._inflight: list = []

    @staticmethod
    def create(config: PipelineConfig) -> "Pipeline":
        return Pipeline(config)

    @property
    def ctx(self):
        return self._ctx

    def close(self) -> None:
        if getattr(self, "_ctx", None):
            N.hip().kf_destroy(self._ctx)  # synchronises the stream first
            self._ctx = None
            self._inflight = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, status: int) -> None:
        N.check(status, self._ctx)

    def hold(self, *tensors) -> None:
        self._inflight.extend(t for t in tensors if t is not None)

    def reset(self) -> None:
        self._check(N.hip().kf_reset(self._ctx))

    # ---- integrate (pipeline.hpp:78-80) ----------------------------------
    def integrate(self, depth, world_to_camera, intrinsics: CameraIntrinsics | None = None, color=None,
                  normals=None, count: bool = False):
        """depth/color/normals: device tensors ((H, W) uint16 as int16,
        (H, W) int32 RGBA8, (H, W, 4) float32)."""
        k = intrinsics or self.config.intrinsics
        if depth is None or len(getattr(depth, "shape", ())) != 2:
            raise ValueError("depth: must be an (H, W) device image")
        h, w = int(depth.shape[0]), int(depth.shape[1])
        if h > self.config.height or w > self.config.width:
            raise ValueError(f"depth: {w} x {h} exceeds the configured camera {self.config.width} x "
                             f"{self.config.height}")
        _check_image(depth, "depth", _DEPTH, (h, w))
        _check_image(color, "color", _COLOR, (h, w), optional=True)
        _check_image(normals, "normals", _FLOAT, (h, w, 4), optional=True)
        self._torch.cuda.synchronize()  # inputs come from torch's stream
        f = N.KfDepthFrame()
        f.depth = depth.data_ptr()
        f.color = color.data_ptr() if color is not None else None
        f.normals = normals.data_ptr() if normals is not None else None
        f.height, f.width = int(depth.shape[0]), int(depth.shape[1])
        f.fx, f.fy, f.cx, f.cy = k.fx, k.fy, k.cx, k.cy
        w2c = np.asarray(world_to_camera, np.float64)[:3, :4].astype(np.float32).reshape(12)
        f.world_to_camera[:] = list(w2c)
        counts = (C.c_int64 * 2)() if count else None
        self.hold(depth, color, normals)
        self._check(N.hip().kf_integrate(self._ctx, C.byref(f), counts))
        if count:
            self._inflight.clear()  # the counting call waited for the stream
        return (counts[0], counts[1]) if count else None

    # ---- raycast (pipeline.hpp:82) ----------------------------------------
    def raycast(self, camera_to_world, level: int = 0) -> SurfaceMaps:
        pose = (C.c_double * 16)(*np.asarray(camera_to_world, np.float64).reshape(16))
        mv = N.KfMapsView()
        self._check(N.hip().kf_raycast(self._ctx, pose, level, C.byref(mv)))
        return self._maps(mv)

    def _maps(self, mv: N.KfMapsView) -> SurfaceMaps:
        t = self._torch
        h, w = mv.height, mv.width
        self.synchronize()
        return SurfaceMaps(_wrap(t, mv.points, (h, w, 4), t.float32), _wrap(t, mv.normals, (h, w, 4), t.float32),
                           _wrap(t, mv.colors, (h, w), t.int32))

    # ---- volume access (pipeline.hpp:87-91) ----------------------------------
    def observed_voxel_count(self) -> int:
        n = C.c_int64()
        self._check(N.hip().kf_observed_voxel_count(self._ctx, C.byref(n)))
        return n.value

    def host_view(self) -> tuple[np.ndarray, np.ndarray | None]:
        """(voxels, colors) copied to the host in the native store layout."""
        vb, cb = C.c_size_t(), C.c_size_t()
        N.hip().kf_volume_bytes(self._ctx, C.byref(vb), C.byref(cb))
        n = int(np.prod(self.config.volume.resolution))
        vox = np.empty((n, 2), np.float32 if self.config.voxel == "float" else np.uint16)
        col = np.empty((n, 4), np.float32) if cb.value else None
        self._check(N.hip().kf_volume_download(self._ctx, vox.ctypes.data_as(C.c_void_p),
                                               col.ctypes.data_as(C.c_void_p) if col is not None else None))
        return vox, col

    def upload_volume(self, voxels: np.ndarray, colors: np.ndarray | None = None) -> None:
        voxels = np.ascontiguousarray(voxels)
        colors = np.ascontiguousarray(colors) if colors is not None else None
        self._check(N.hip().kf_volume_upload(self._ctx, voxels.ctypes.data_as(C.c_void_p),
                                             colors.ctypes.data_as(C.c_void_p) if colors is not None else None))

    # ---- meshing (pipeline.hpp:95-96) ---------------------------------------
    def extract_mesh(self, min_weight: float = 0.0):
        return MarchingCubes.extract(self, min_weight)

    # ---- whole frame (reconstruction.cpp:184-216) ----------------------------
    def process_frame(self, depth, color=None, icp: ProjectiveIcpOptions | None = None, report: bool = True,
                      sync_inputs: bool = True):
        """report=False only enqueues the frame; inputs produced on torch's
        stream must be complete first (sync_inputs)."""
        shape = (self.config.height, self.config.width)
        _check_image(depth, "depth", _DEPTH, shape)
        _check_image(color, "color", _COLOR, shape, optional=True)
        if sync_inputs:
            self._torch.cuda.synchronize()
        opts = (icp or ProjectiveIcpOptions()).native()
        res = N.KfFrameResult() if report else None
        self.hold(depth, color)
        try:
            self._check(N.hip().kf_process_frame(self._ctx, depth.data_ptr(),
                                                 color.data_ptr() if color is not None else None, C.byref(opts),
                                                 C.byref(res) if report else None))
        finally:
            if report:
                self._inflight.clear()  # the reporting call waited for the stream
        return res

    def trajectory(self) -> np.ndarray:
        n = C.c_int64()
        N.hip().kf_trajectory(self._ctx, None, 0, C.byref(n))
        out = np.zeros((max(n.value, 1), 4, 4), np.float64)
        self._check(N.hip().kf_trajectory(self._ctx, out.ctypes.data_as(C.POINTER(C.c_double)), n.value,
                                          C.byref(n)))
        return out[: n.value]

    def stage_times(self) -> N.KfStageTimes:
        t = N.KfStageTimes()
        self._check(N.hip().kf_last_stage_times(self._ctx, C.byref(t)))
        return t

    def set_sync_timeout(self, ticks: int) -> None:
        """Bound (device wall-clock ticks, 100 MHz) of the grid-wide wait of
        the learned fallback pass's one-launch levels; a diagnostic."""
        self._check(N.hip().kf_set_sync_timeout(self._ctx, int(ticks)))

    def attach_posenet(self, net) -> None:
        """Use a trained PoseNet for init="learned" frames (None detaches)."""
        self._check(N.hip().kf_attach_posenet(self._ctx, net.handle if net is not None else None))
        self._posenet = net

    def block_bitmaps(self) -> dict:
        """The per-block observed / skip / free-space bits as bool arrays in
        flat block order."""
        words = C.c_int64()
        self._check(N.hip().kf_block_bitmaps(self._ctx, None, None, None, C.byref(words)))
        maps = [np.zeros(words.value, np.uint32) for _ in range(3)]
        self._check(N.hip().kf_block_bitmaps(self._ctx, *[m.ctypes.data_as(C.c_void_p) for m in maps],
                                             C.byref(words)))
        r = self.config.volume.resolution
        nblocks = ((r[0] + 7) // 8) * ((r[1] + 7) // 8) * ((r[2] + 7) // 8)
        bits = [np.unpackbits(m.view(np.uint8), bitorder="little")[:nblocks].astype(bool) for m in maps]
        return {"observed": bits[0], "skip": bits[1], "free": bits[2]}

    def hashed_stats(self) -> dict:
        """Hashed volume: allocated blocks, dropped requests (overflow), hash
        entries, blocks integrated by the last frame."""
        out = (C.c_int64 * 4)()
        self._check(N.hip().kf_hashed_stats(self._ctx, out))
        return {"allocated": out[0], "overflow": out[1], "entries": out[2], "last_work": out[3]}

    def hashed_download(self):
        """(keys increasing, voxels (n, 512, 2), colours (n, 512, 4) or None)
        of the allocated blocks of a hashed volume."""
        n = C.c_int64()
        self._check(N.hip().kf_hashed_download(self._ctx, None, None, None, 0, C.byref(n)))
        m = n.value
        keys = np.empty(max(m, 1), np.uint32)
        vox = np.empty((max(m, 1), 512, 2), np.float32 if self.config.voxel == "float" else np.uint16)
        col = np.empty((max(m, 1), 512, 4), np.float32) if self.config.color == "float" else None
        self._check(N.hip().kf_hashed_download(self._ctx, keys.ctypes.data_as(C.c_void_p),
                                               vox.ctypes.data_as(C.c_void_p),
                                               col.ctypes.data_as(C.c_void_p) if col is not None else None, m,
                                               C.byref(n)))
        return keys[:m], vox[:m], (col[:m] if col is not None else None)

    def hashed_blocks(self) -> np.ndarray:
        """Sorted keys of the allocated blocks (flat x-major block indices)."""
        n = C.c_int64()
        self._check(N.hip().kf_hashed_blocks(self._ctx, None, 0, C.byref(n)))
        out = np.empty(max(n.value, 1), np.uint32)
        self._check(N.hip().kf_hashed_blocks(self._ctx, out.ctypes.data_as(C.c_void_p), n.value, C.byref(n)))
        return out[: n.value]

    def fused_counts(self) -> tuple[int, int]:
        out = (C.c_int64 * 2)()
        self._check(N.hip().kf_fused_counts(self._ctx, out))
        return out[0], out[1]

    def synchronize(self) -> None:
        self._check(N.hip().kf_synchronize(self._ctx))
        self._inflight.clear()


def _wrap(torch, ptr: int, shape, dtype):
    """Copy a device buffer owned by the library into a torch tensor."""
    n = int(np.prod(shape))
    out = torch.empty(shape, dtype=dtype, device="cuda")
    nbytes = n * torch.tensor([], dtype=dtype).element_size()
    torch.cuda.synchronize()
    _memcpy_d2d(out.data_ptr(), ptr, nbytes)
    return out


_hiprt = None


def _memcpy_d2d(dst: int, src: int, nbytes: int) -> None:
    global _hiprt
    if _hiprt is None:
        _hiprt = C.CDLL("libamdhip64.so")
        _hiprt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hiprt.hipMemcpy.restype = C.c_int
    err = _hiprt.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes), 3)  # DeviceToDevice
    if err != 0:
        raise N.KinectFusionError(3, f"hipMemcpy failed with {err}")


class PyramidSource:
    """PyramidSource over a pipeline's context (depth_processing.hpp:422-463)."""

    def __init__(self, pipeline: Pipeline):
        self.pipeline = pipeline

    def build(self, raw_depth) -> int:
        c = self.pipeline.config
        _check_image(raw_depth, "raw_depth", _DEPTH, (c.height, c.width))
        self.pipeline._torch.cuda.synchronize()
        self.pipeline.hold(raw_depth)
        n = C.c_int32()
        N.check(N.hip().kf_pyramid_build(self.pipeline.ctx, raw_depth.data_ptr(), C.byref(n)), self.pipeline.ctx)
        return n.value

    def level(self, index: int):
        """(vertices, normals, depth, intrinsics) of one level as device tensors."""
        lv = N.KfLevelView()
        N.check(N.hip().kf_pyramid_level(self.pipeline.ctx, index, C.byref(lv)), self.pipeline.ctx)
        self.pipeline.synchronize()
        t = self.pipeline._torch
        h, w = lv.height, lv.width
        return {
            "vertices": _wrap(t, lv.vertices, (h, w, 4), t.float32),
            "normals": _wrap(t, lv.normals, (h, w, 4), t.float32),
            "depth": _wrap(t, lv.depth, (h, w), t.int16),
            "intrinsics": CameraIntrinsics(lv.fx, lv.fy, lv.cx, lv.cy),
        }


def _check_maps(live_vertices, live_normals, model_points, model_normals):
    """(live h, w, model h, w) of float4 maps the ICP kernels read."""
    if len(getattr(live_vertices, "shape", ())) != 3 or len(getattr(model_points, "shape", ())) != 3:
        raise ValueError("maps: must be (H, W, 4) device images")
    lh, lw = int(live_vertices.shape[0]), int(live_vertices.shape[1])
    mh, mw = int(model_points.shape[0]), int(model_points.shape[1])
    _check_image(live_vertices, "live_vertices", _FLOAT, (lh, lw, 4))
    _check_image(live_normals, "live_normals", _FLOAT, (lh, lw, 4))
    _check_image(model_points, "model_points", _FLOAT, (mh, mw, 4))
    _check_image(model_normals, "model_normals", _FLOAT, (mh, mw, 4))
    return lh, lw, mh, mw


class ProjectiveIcpTracker:
    """icp_optimizer.hpp:107-130 in its whole-level device-loop form."""

    def __init__(self, pipeline: Pipeline, options: ProjectiveIcpOptions | None = None):
        self.pipeline = pipeline
        self.options = options or ProjectiveIcpOptions()

    def estimate_pose(self, live_vertices, live_normals, model_points, model_normals,
                      model_intrinsics: CameraIntrinsics, model_camera_to_world, initial_camera_to_world,
                      iterations: int) -> IcpOutcome:
        k = (C.c_float * 4)(model_intrinsics.fx, model_intrinsics.fy, model_intrinsics.cx, model_intrinsics.cy)
        mp = (C.c_double * 16)(*np.asarray(model_camera_to_world, np.float64).reshape(16))
        pose = (C.c_double * 16)(*np.asarray(initial_camera_to_world, np.float64).reshape(16))
        res = N.KfIcpResult()
        opts = self.options.native()
        lh, lw, mh, mw = _check_maps(live_vertices, live_normals, model_points, model_normals)
        self.pipeline._torch.cuda.synchronize()
        N.check(N.hip().kf_icp_track_maps(self.pipeline.ctx, live_vertices.data_ptr(), live_normals.data_ptr(),
                                          lw, lh, model_points.data_ptr(), model_normals.data_ptr(), mw, mh, k,
                                          mp, pose, iterations, C.byref(opts), C.byref(res)), self.pipeline.ctx)
        return _outcome(np.array(pose[:], np.float64).reshape(4, 4), res)

    def normal_equations(self, live_vertices, live_normals, model_points, model_normals,
                         model_intrinsics: CameraIntrinsics, camera12, model12):
        """One correspondence sweep: (sums[29] float64 of the FP32
        compensated device sums, count)."""
        k = (C.c_float * 4)(model_intrinsics.fx, model_intrinsics.fy, model_intrinsics.cx, model_intrinsics.cy)
        cam = (C.c_float * 12)(*np.asarray(camera12, np.float32).reshape(12))
        mod = (C.c_float * 12)(*np.asarray(model12, np.float32).reshape(12))
        sums = (C.c_double * 29)()
        count = C.c_int64()
        lh, lw, mh, mw = _check_maps(live_vertices, live_normals, model_points, model_normals)
        self.pipeline._torch.cuda.synchronize()
        N.check(N.hip().kf_icp_reduce(self.pipeline.ctx, live_vertices.data_ptr(), live_normals.data_ptr(), lw, lh,
                                      model_points.data_ptr(), model_normals.data_ptr(), mw, mh, k, cam, mod,
                                      self.options.max_point_distance, self.options.min_normal_dot, sums,
                                      C.byref(count)), self.pipeline.ctx)
        return np.array(sums[:], np.float64), count.value

    def track(self, camera_to_world) -> tuple[np.ndarray, list[IcpOutcome]]:
        """Coarse-to-fine tracking of the current pyramid
        (Reconstruction::track_pose, reconstruction.cpp:218-250)."""
        pose = (C.c_double * 16)(*np.asarray(camera_to_world, np.float64).reshape(16))
        res = (N.KfIcpResult * 4)()
        opts = self.options.native()
        N.check(N.hip().kf_track(self.pipeline.ctx, C.byref(opts), pose, res), self.pipeline.ctx)
        p = np.array(pose[:], np.float64).reshape(4, 4)
        return p, [_outcome(p, res[i]) for i in range(self.pipeline.config.depth.levels)]


@dataclass
class Mesh:
    """MarchingCubes::Mesh (marching_cubes.hpp:18-23)."""
    positions: np.ndarray
    normals: np.ndarray
    colors: np.ndarray
    triangles: np.ndarray


class MarchingCubes:
    @staticmethod
    def extract(pipeline: Pipeline, min_weight: float = 0.0) -> Mesh:
        L = N.hip()
        h = C.c_void_p()
        N.check(L.kf_mesh_extract(pipeline.ctx, min_weight, C.byref(h)), pipeline.ctx)
        try:
            nv, nt = C.c_int64(), C.c_int64()
            N.check(L.kf_mesh_size(h, C.byref(nv), C.byref(nt)), pipeline.ctx)
            pos = np.empty((nv.value, 3), np.float32)
            nrm = np.empty((nv.value, 3), np.float32)
            col = np.empty((nv.value, 3), np.float32)
            tri = np.empty((nt.value, 3), np.uint32)
            vp = C.c_void_p
            N.check(L.kf_mesh_copy(h, pos.ctypes.data_as(vp), nrm.ctypes.data_as(vp), col.ctypes.data_as(vp),
                                   tri.ctypes.data_as(vp)), pipeline.ctx)
        finally:
            L.kf_mesh_free(h)
        return Mesh(pos, nrm, col, tri)

"""Command-line options of the reconstruction app (src/app/app_options.hpp:
21-127, app_options.cpp:30-358): the same flags, defaults, derived option
structs and validation, bound to argparse.

Values that mirror library defaults are taken from the library's own option
types, so they stay in sync.  Differences from the reference:
* --space takes "hip" (the only space built here; the default) and refuses
  "cpu" / "cuda" when a pipeline is created (the reference's
  std::logic_error for a backend that is not built);
* --raycast accepts "march" and "bitmap-march", which render identical
  maps here (the march skips empty blocks exactly); "band-march" is refused;
* --icp-device-solve and --icp-capture-graph are accepted and change
  nothing: the HIP space always solves on the device and replays a captured
  frame graph;
* additions: --init (the tracker's seed: previous, constant-velocity or
  learned), --posenet (trained network parameters for --init learned),
  --intrinsics (a camera for datasets that are not TUM freiburg sequences)
  and --learned-iterations.
"""
from __future__ import annotations

import argparse
from dataclasses import dataclass, field, replace
from pathlib import Path

from ..pipeline import (CameraIntrinsics, DepthProcessingOptions, PipelineConfig, ProjectiveIcpOptions,
                        RaycastOptions, TsdfIntegrationOptions, VolumeGeometry)
from . import names

VERSION = "kinectfusion_amd 2.0"

# AppOptions defaults (app_options.hpp:21-27)
DEFAULT_MAX_FRAMES = 30
DEFAULT_VOLUME_RESOLUTION = 512
DEFAULT_VOXEL_SIZE = 0.01
DEFAULT_TRUNCATION_DISTANCE = 0.05
DEFAULT_VOLUME_CAMERA_MARGIN = 2.56
DEFAULT_MESH_MIN_WEIGHT = 2.0
# per-level ICP iterations when --icp-iterations is -1 (app_options.cpp:24-29)
DEFAULT_ICP_ITERATIONS = (10, 5, 4)

_TSDF = TsdfIntegrationOptions()
_DEPTH = DepthProcessingOptions()
_ICP = ProjectiveIcpOptions()


@dataclass
class AppOptions:
    """AppOptions (app_options.hpp:31-83)."""
    dataset_dir: Path = Path("../data/rgbd_dataset_freiburg1_xyz")
    max_frames: int = DEFAULT_MAX_FRAMES
    volume_resolution: int = DEFAULT_VOLUME_RESOLUTION
    voxel_size: float = DEFAULT_VOXEL_SIZE
    truncation_distance: float = DEFAULT_TRUNCATION_DISTANCE
    volume_camera_margin: float = DEFAULT_VOLUME_CAMERA_MARGIN
    depth_scale: float = _TSDF.depth_scale
    min_depth: float = _TSDF.min_depth
    max_depth: float = _TSDF.max_depth
    pyramid_levels: int = _DEPTH.levels
    icp_iterations: int = -1
    matching_distance: float = _ICP.max_point_distance
    min_normal_dot: float = _ICP.min_normal_dot
    max_pose_update_translation: float = _ICP.max_update_translation
    max_pose_update_rotation: float = _ICP.max_update_rotation
    min_icp_eigenvalue: float = _ICP.min_system_eigenvalue
    max_icp_condition_number: float = _ICP.max_condition_number
    icp_lambda: float = _ICP.damping_lambda
    icp_capture_graph: bool = False
    icp_device_solve: bool = False
    icp_adaptive_damping: bool = False
    icp_damping: str = "none"
    voxel_store: str = "float"
    color_store: str = "float"
    raycast_backend: str = "march"
    integration_mode: str = "full"
    storage_layout: str = "dense"
    sparse_capacity: int = 0
    projective_tsdf_distance: bool = True
    distance_scaled_truncation: bool = False
    truncation_distance_scale: float = _TSDF.truncation_distance_scale
    tsdf_variant: str = "angle-weighted"
    space: str = "hip"
    pipelines_config: Path | None = None
    compare_every_n_frames: int = 1
    bilateral_filter: bool = True
    bilateral_radius: int = _DEPTH.bilateral_radius
    bilateral_spatial_sigma: float = _DEPTH.bilateral_spatial_sigma
    bilateral_depth_sigma: float = _DEPTH.bilateral_depth_sigma
    interactive_relocalization: bool = False
    preload_frames: bool = False
    output_dir: Path = Path("kinectfusion_output")
    write_raycast_images: bool = True
    write_point_clouds: bool = True
    write_mesh: bool = True
    mesh_min_weight: float = DEFAULT_MESH_MIN_WEIGHT
    raycast_tsdf_from_valid_corners: bool = False
    cell_gradient_normals: bool = False
    raycast_seed_previous: bool = False
    # additions of this build
    init: str = "previous"
    posenet: Path | None = None
    intrinsics: tuple[float, float, float, float] | None = None
    learned_iterations: tuple[int, ...] = (4, 0, 0)

    # ---- derived option structs (app_options.cpp:33-178) -------------------
    def volume_origin(self) -> tuple[float, float, float]:
        g = VolumeGeometry.centered(self.volume_resolution, self.voxel_size, self.volume_camera_margin,
                                    self.truncation_distance)
        return g.origin

    def volume_geometry(self) -> VolumeGeometry:
        return VolumeGeometry.centered(self.volume_resolution, self.voxel_size, self.volume_camera_margin,
                                       self.truncation_distance)

    def tsdf_options(self) -> TsdfIntegrationOptions:
        return TsdfIntegrationOptions(depth_scale=self.depth_scale, min_depth=self.min_depth,
                                      max_depth=self.max_depth, projective_distance=self.projective_tsdf_distance,
                                      distance_scaled_truncation=self.distance_scaled_truncation,
                                      truncation_distance_scale=self.truncation_distance_scale,
                                      mode=names.integration_mode_from_name(self.integration_mode))

    def raycast_options(self) -> RaycastOptions:
        return RaycastOptions(min_depth=self.min_depth, max_depth=self.max_depth,
                              require_all_corners=self.raycast_tsdf_from_valid_corners,
                              cell_gradient_normals=self.cell_gradient_normals,
                              seed_from_previous=self.raycast_seed_previous)

    def depth_options(self) -> DepthProcessingOptions:
        return DepthProcessingOptions(levels=self.pyramid_levels, bilateral_filter=self.bilateral_filter,
                                      bilateral_radius=self.bilateral_radius,
                                      bilateral_spatial_sigma=self.bilateral_spatial_sigma,
                                      bilateral_depth_sigma=self.bilateral_depth_sigma)

    def icp_options(self) -> ProjectiveIcpOptions:
        return replace(_ICP, max_point_distance=self.matching_distance, min_normal_dot=self.min_normal_dot,
                       max_update_translation=self.max_pose_update_translation,
                       max_update_rotation=self.max_pose_update_rotation,
                       min_system_eigenvalue=self.min_icp_eigenvalue,
                       max_condition_number=self.max_icp_condition_number,
                       damping=names.icp_damping_mode_from_name(self.icp_damping), damping_lambda=self.icp_lambda,
                       adaptive_damping=self.icp_adaptive_damping)

    def icp_iterations_for_level(self, level: int) -> int:
        """A fixed override, or the default 10 / 5 / 4 schedule."""
        if self.icp_iterations > 0:
            return self.icp_iterations
        return DEFAULT_ICP_ITERATIONS[min(level, 2)]

    def pipeline_config(self, camera: CameraIntrinsics, width: int, height: int) -> PipelineConfig:
        """The single pipeline the CLI options describe (app_options.cpp:
        56-73), for a sensor of the given camera and size."""
        capacity = self.sparse_capacity
        if capacity == 0:
            r = self.volume_resolution
            capacity = (((r + 7) // 8) ** 3) // 4  # PipelineConfig::sparse_block_capacity default
        return PipelineConfig(
            name="baseline", space=names.memory_space_from_name(self.space),
            tsdf_rule=names.tsdf_rule_from_name(self.tsdf_variant),
            voxel=names.voxel_store_from_name(self.voxel_store),
            color=names.color_store_from_name(self.color_store),
            storage=names.storage_layout_from_name(self.storage_layout),
            raycast_backend=names.raycast_backend_from_name(self.raycast_backend),
            volume=self.volume_geometry(), integration=self.tsdf_options(), raycast=self.raycast_options(),
            depth=self.depth_options(), intrinsics=camera, width=width, height=height, init=self.init,
            icp_iterations=tuple(self.icp_iterations_for_level(level) for level in range(self.pyramid_levels)),
            hash_buckets=0, block_capacity=capacity, learned_iterations=tuple(self.learned_iterations))


def _flag(p: argparse.ArgumentParser, name: str, dest: str, default: bool, help_text: str, negatable: bool):
    if negatable:
        p.add_argument(name, dest=dest, action=argparse.BooleanOptionalAction, default=default, help=help_text)
    else:
        p.add_argument(name, dest=dest, action="store_true", default=default, help=help_text)


def build_parser() -> argparse.ArgumentParser:
    """configure_cli (app_options.cpp:160-318): every flag of the reference,
    then this build's additions."""
    d = AppOptions()
    p = argparse.ArgumentParser(prog="kinectfusion", description="KinectFusion reconstruction on an MI355X.")
    p.add_argument("--version", action="version", version=VERSION)
    p.add_argument("dataset", nargs="?", type=Path, default=d.dataset_dir,
                   help="TUM RGB-D dataset directory, or synthetic:<room|hall>[:seed] for a generated sequence.")
    p.add_argument("--frames", dest="max_frames", type=int, default=d.max_frames,
                   help="Maximum number of frames to process. Set -1 to process all frames.")
    p.add_argument("--volume-resolution", type=int, default=d.volume_resolution)
    p.add_argument("--voxel-size", type=float, default=d.voxel_size)
    p.add_argument("--truncation-distance", type=float, default=d.truncation_distance)
    p.add_argument("--volume-camera-margin", type=float, default=d.volume_camera_margin)
    p.add_argument("--depth-scale", type=float, default=d.depth_scale,
                   help="Raw depth units for one meter. The TUM default is 5000.")
    p.add_argument("--min-depth", type=float, default=d.min_depth)
    p.add_argument("--max-depth", type=float, default=d.max_depth)
    p.add_argument("--pyramid-levels", type=int, default=d.pyramid_levels)
    p.add_argument("--icp-iterations", type=int, default=d.icp_iterations,
                   help="Fixed number of ICP iterations for each pyramid level. Set -1 for the defaults [10, 5, 4].")
    p.add_argument("--matching-distance", type=float, default=d.matching_distance)
    p.add_argument("--min-normal-dot", type=float, default=d.min_normal_dot)
    p.add_argument("--max-pose-update-translation", type=float, default=d.max_pose_update_translation)
    p.add_argument("--max-pose-update-rotation", type=float, default=d.max_pose_update_rotation)
    p.add_argument("--min-icp-eigenvalue", type=float, default=d.min_icp_eigenvalue)
    p.add_argument("--max-icp-condition-number", type=float, default=d.max_icp_condition_number)
    p.add_argument("--voxel", dest="voxel_store", default=d.voxel_store,
                   help="Storage of the geometric TSDF voxel: float (8 bytes), quantized (int16) or bf16 (4 bytes).")
    p.add_argument("--color", dest="color_store", default=d.color_store, help="float or none.")
    _flag(p, "--cell-normals", "cell_gradient_normals", d.cell_gradient_normals,
          "Raycast normals from the 8 corners of the final sample.", False)
    _flag(p, "--raycast-seed-previous", "raycast_seed_previous", d.raycast_seed_previous,
          "Start each ray just in front of the surface from the last frame.", False)
    p.add_argument("--storage", dest="storage_layout", default=d.storage_layout,
                   help="dense, or sparse (8^3 voxel blocks along the truncation band; needs --integration band).")
    p.add_argument("--sparse-capacity", type=int, default=d.sparse_capacity,
                   help="Number of blocks in the sparse pool. Set 0 to use one quarter of the block count.")
    p.add_argument("--integration", dest="integration_mode", default=d.integration_mode, help="full or band.")
    p.add_argument("--raycast", dest="raycast_backend", default=d.raycast_backend,
                   help="march or bitmap-march (identical maps); band-march is not provided.")
    _flag(p, "--icp-device-solve", "icp_device_solve", d.icp_device_solve,
          "Accepted for compatibility: the HIP space always solves ICP on the device.", False)
    _flag(p, "--icp-capture-graph", "icp_capture_graph", d.icp_capture_graph,
          "Accepted for compatibility: the HIP space always replays a captured frame graph.", False)
    p.add_argument("--icp-damping", default=d.icp_damping, help="none, identity or diagonal.")
    p.add_argument("--icp-lambda", type=float, default=d.icp_lambda)
    _flag(p, "--icp-adaptive-damping", "icp_adaptive_damping", d.icp_adaptive_damping,
          "Adapt lambda per trial and roll back trials that raise the cost. Requires --icp-damping.", False)
    _flag(p, "--projective-tsdf-distance", "projective_tsdf_distance", d.projective_tsdf_distance,
          "Use the lambda-corrected projective TSDF distance.", True)
    _flag(p, "--distance-scaled-truncation", "distance_scaled_truncation", d.distance_scaled_truncation,
          "Widen the TSDF truncation band linearly with measured depth.", True)
    p.add_argument("--truncation-distance-scale", type=float, default=d.truncation_distance_scale)
    p.add_argument("--tsdf-variant", default=d.tsdf_variant, choices=["angle-weighted", "classic"])
    p.add_argument("--space", default=d.space, type=str.lower,
                   help="Memory space of the pipeline: hip (this build). cpu and cuda are not built here.")
    p.add_argument("--pipelines", dest="pipelines_config", type=Path, default=None,
                   help="TOML file that describes an ablation pipeline set ([[pipeline]] tables).")
    p.add_argument("--compare-every", dest="compare_every_n_frames", type=int, default=d.compare_every_n_frames)
    _flag(p, "--bilateral-filter", "bilateral_filter", d.bilateral_filter,
          "Filter the tracking depth with a bilateral filter.", True)
    p.add_argument("--bilateral-radius", type=int, default=d.bilateral_radius)
    p.add_argument("--bilateral-spatial-sigma", type=float, default=d.bilateral_spatial_sigma)
    p.add_argument("--bilateral-depth-sigma", type=float, default=d.bilateral_depth_sigma)
    _flag(p, "--interactive-relocalization", "interactive_relocalization", d.interactive_relocalization,
          "Pause after a tracking failure. Continue on user input.", False)
    _flag(p, "--preload", "preload_frames", d.preload_frames,
          "Decode the whole dataset into memory before the frame loop.", False)
    p.add_argument("--output-dir", type=Path, default=d.output_dir)
    _flag(p, "--write-raycast-images", "write_raycast_images", d.write_raycast_images,
          "Write the raycast color images as PNG files.", True)
    _flag(p, "--write-point-clouds", "write_point_clouds", d.write_point_clouds,
          "Write the raycast point clouds as binary PLY files.", True)
    _flag(p, "--write-mesh", "write_mesh", d.write_mesh, "Write the final TSDF surface as mesh.ply.", True)
    p.add_argument("--mesh-min-weight", type=float, default=d.mesh_min_weight)
    _flag(p, "--raycast-tsdf-from-valid-corners", "raycast_tsdf_from_valid_corners",
          d.raycast_tsdf_from_valid_corners, "Interpolate the raycast TSDF only where all 8 corners are valid.", True)
    # additions
    p.add_argument("--init", default=d.init, choices=["previous", "constant-velocity", "learned"],
                   help="Seed of the tracker: the previous pose (the reference), constant velocity, or the learned "
                        "pose proposal (needs --posenet).")
    p.add_argument("--posenet", type=Path, default=None,
                   help="Parameters of a trained pose network (.npy float32, kf_posenet_parameter_count values).")
    p.add_argument("--learned-iterations", type=lambda s: tuple(int(v) for v in s.split(",")),
                   default=d.learned_iterations, help="ICP iterations per level from the proposal, finest first.")
    p.add_argument("--intrinsics", type=lambda s: tuple(float(v) for v in s.split(",")), default=None,
                   help="fx,fy,cx,cy of the depth camera (default: by TUM sequence name).")
    return p


def parse_args(argv=None) -> AppOptions:
    ns = build_parser().parse_args(argv)
    opts = AppOptions(**{k: v for k, v in vars(ns).items() if k != "dataset"})
    opts.dataset_dir = ns.dataset
    return opts


def validate_options(o: AppOptions) -> None:
    """validate_options (app_options.cpp:320-356); ValueError is the
    reference's std::invalid_argument."""
    def require(ok: bool, msg: str) -> None:
        if not ok:
            raise ValueError(msg)

    require(o.volume_resolution > 0, "volume resolution must be positive")
    require(o.voxel_size > 0.0, "voxel size must be positive")
    extent = o.volume_resolution * o.voxel_size
    require(0.0 <= o.volume_camera_margin <= extent, "volume camera margin must be within [0, volume extent]")
    require(o.truncation_distance > 0.0, "truncation distance must be positive")
    require(o.depth_scale > 0.0, "depth scale must be positive")
    require(o.min_depth >= 0.0 and o.max_depth > o.min_depth, "depth range is invalid")
    require(o.pyramid_levels > 0, "pyramid levels must be positive")
    require(o.icp_iterations > 0 or o.icp_iterations == -1, "ICP iterations must be positive or -1")
    require(o.matching_distance > 0.0, "matching distance must be positive")
    require(-1.0 <= o.min_normal_dot <= 1.0, "min normal dot must be within [-1, 1]")
    require(o.truncation_distance_scale >= 0.0, "truncation distance scale must be non-negative")
    require(o.bilateral_radius >= 0, "bilateral radius must be non-negative")
    require(o.bilateral_spatial_sigma > 0.0 and o.bilateral_depth_sigma > 0.0, "bilateral sigmas must be positive")
    require(o.icp_lambda > 0.0, "ICP lambda must be positive")
    require(not o.icp_adaptive_damping or names.icp_damping_mode_from_name(o.icp_damping) != "none",
            "--icp-adaptive-damping requires --icp-damping")
    require(o.init != "learned" or o.posenet is not None, "--init learned requires --posenet")
    require(o.intrinsics is None or len(o.intrinsics) == 4, "--intrinsics takes fx,fy,cx,cy")

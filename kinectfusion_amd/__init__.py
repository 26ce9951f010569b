"""MI355X (gfx950) KinectFusion: HIP kernels behind the reference's stage
interfaces (maximsolodukhin/KinectFusion), reached through the C ABI of
libkf_hip.so (include/kf_hip.h).

Importing the package does not touch the GPU; creating a Pipeline does, and
fails loudly when the HIP library or a gfx950 device is missing.
"""
from ._native import KinectFusionError, hip, scene_lib  # noqa: F401
from .pipeline import (  # noqa: F401
    CameraIntrinsics,
    DepthProcessingOptions,
    IcpOutcome,
    MarchingCubes,
    Mesh,
    Pipeline,
    PipelineConfig,
    ProjectiveIcpOptions,
    ProjectiveIcpTracker,
    PyramidSource,
    RaycastOptions,
    SurfaceMaps,
    TsdfIntegrationOptions,
    VolumeGeometry,
)
from .scene import SyntheticScene  # noqa: F401

__all__ = [
    "CameraIntrinsics", "DepthProcessingOptions", "IcpOutcome", "KinectFusionError", "MarchingCubes", "Mesh",
    "Pipeline", "PipelineConfig", "ProjectiveIcpOptions", "ProjectiveIcpTracker", "PyramidSource",
    "RaycastOptions", "SurfaceMaps", "SyntheticScene", "TsdfIntegrationOptions", "VolumeGeometry",
]

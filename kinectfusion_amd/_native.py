"""ctypes binding of libkf_hip.so (include/kf_hip.h) and libkf_scene.so.

The structures mirror the C declarations field for field.  Loading fails
loudly: there is no host fallback for any stage.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parent
HIP_LIB = PKG / "libkf_hip.so"
SCENE_LIB = PKG / "libkf_scene.so"

KF_OK = 0
KF_ERR_INVALID_ARGUMENT = 1
STATUS_NAMES = {0: "KF_OK", 1: "KF_ERR_INVALID_ARGUMENT", 2: "KF_ERR_OUT_OF_MEMORY", 3: "KF_ERR_DEVICE",
                4: "KF_ERR_UNSUPPORTED", 5: "KF_ERR_STATE"}

VOXEL_FLOAT, VOXEL_QUANTIZED, VOXEL_BF16 = 0, 1, 2
COLOR_FLOAT, COLOR_NONE = 0, 1
STORAGE_DENSE, STORAGE_HASHED = 0, 1
RULE_CLASSIC, RULE_ANGLE_WEIGHTED = 0, 1
DAMPING_NONE, DAMPING_IDENTITY, DAMPING_DIAGONAL = 0, 1, 2
INIT_PREVIOUS, INIT_CONSTANT_VELOCITY, INIT_LEARNED = 0, 1, 2
INTEGRATION_FULL, INTEGRATION_BAND = 0, 1


class KfConfig(C.Structure):
    _fields_ = [
        ("resolution", C.c_int64 * 3), ("voxel_size", C.c_float), ("origin", C.c_float * 3),
        ("truncation_distance", C.c_float), ("voxel_store", C.c_int32), ("color_store", C.c_int32),
        ("storage", C.c_int32), ("tsdf_rule", C.c_int32),
        ("depth_scale", C.c_float), ("observation_weight", C.c_float), ("max_weight", C.c_float),
        ("min_depth", C.c_float), ("max_depth", C.c_float), ("projective_distance", C.c_int32),
        ("distance_scaled_truncation", C.c_int32), ("truncation_distance_scale", C.c_float),
        ("integration_mode", C.c_int32),
        ("ray_min_depth", C.c_float), ("ray_max_depth", C.c_float), ("step_scale", C.c_float),
        ("require_all_corners", C.c_int32), ("cell_normals", C.c_int32), ("seed_from_previous", C.c_int32),
        ("levels", C.c_int32), ("max_normal_depth_jump", C.c_float), ("max_downsample_depth_jump", C.c_float),
        ("bilateral_filter", C.c_int32), ("bilateral_radius", C.c_int32), ("bilateral_spatial_sigma", C.c_float),
        ("bilateral_depth_sigma", C.c_float),
        ("width", C.c_int32), ("height", C.c_int32), ("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float),
        ("cy", C.c_float),
        ("hash_buckets", C.c_int64), ("block_capacity", C.c_int64),
        ("init_mode", C.c_int32), ("icp_iterations", C.c_int32 * 4), ("stage_timing", C.c_int32),
        ("count_fused", C.c_int32), ("learned_iterations", C.c_int32 * 4), ("learned_max_rms", C.c_float),
        ("learned_max_update_translation", C.c_float), ("learned_max_update_rotation", C.c_float),
    ]


class KfIcpOptions(C.Structure):
    _fields_ = [
        ("min_correspondences", C.c_int64), ("max_point_distance", C.c_float), ("min_normal_dot", C.c_float),
        ("min_update_translation", C.c_float), ("min_update_rotation", C.c_float),
        ("max_update_translation", C.c_float), ("max_update_rotation", C.c_float),
        ("min_system_eigenvalue", C.c_float), ("max_condition_number", C.c_float),
        ("damping_mode", C.c_int32), ("damping_lambda", C.c_float), ("adaptive_damping", C.c_int32),
        ("lambda_increase", C.c_float), ("lambda_decrease", C.c_float), ("min_lambda", C.c_float),
        ("max_lambda", C.c_float),
    ]


class KfIcpResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("trials", C.c_int32), ("rejected", C.c_int32), ("pad", C.c_int32),
        ("correspondences", C.c_int64), ("mean_distance", C.c_double), ("mean_squared_residual", C.c_double),
        ("min_eigenvalue", C.c_double), ("condition_number", C.c_double), ("update_translation", C.c_double),
        ("update_rotation", C.c_double), ("lambda_", C.c_double),
    ]


class KfLevelView(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("fx", C.c_float), ("fy", C.c_float),
                ("cx", C.c_float), ("cy", C.c_float), ("depth", C.c_void_p), ("vertices", C.c_void_p),
                ("normals", C.c_void_p)]


class KfMapsView(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("points", C.c_void_p), ("normals", C.c_void_p),
                ("colors", C.c_void_p)]


class KfDepthFrame(C.Structure):
    _fields_ = [("depth", C.c_void_p), ("color", C.c_void_p), ("normals", C.c_void_p), ("width", C.c_int32),
                ("height", C.c_int32), ("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float),
                ("world_to_camera", C.c_float * 12)]


class KfStageTimes(C.Structure):
    _fields_ = [("pyramid_ms", C.c_float), ("raycast_ms", C.c_float), ("icp_ms", C.c_float),
                ("integrate_ms", C.c_float), ("integrate_kernel_ms", C.c_float), ("total_ms", C.c_float),
                ("proposal_ms", C.c_float)]


class KfFrameResult(C.Structure):
    _fields_ = [("tracked", C.c_int32), ("failed_level", C.c_int32), ("used_learned", C.c_int32),
                ("pad", C.c_int32), ("pose", C.c_double * 16), ("level", KfIcpResult * 4),
                ("proposal_translation_error", C.c_double), ("proposal_rotation_error", C.c_double),
                ("learned_rms", C.c_double), ("proposal_pose", C.c_double * 16)]


class KfPosenetConfig(C.Structure):
    _fields_ = [("batch_capacity", C.c_int32), ("pad", C.c_int32), ("seed", C.c_uint64),
                ("learning_rate", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("epsilon", C.c_float),
                ("depth_scale", C.c_float)]


# (name, restype, argtypes) of every exported symbol of include/kf_hip.h
_vp, _i32, _i64, _f32, _dp = C.c_void_p, C.c_int32, C.c_int64, C.c_float, C.POINTER(C.c_double)
HIP_SYMBOLS = [
    ("kf_config_defaults", None, [C.POINTER(KfConfig)]),
    ("kf_icp_options_defaults", None, [C.POINTER(KfIcpOptions)]),
    ("kf_create", _i32, [C.POINTER(KfConfig), C.POINTER(_vp)]),
    ("kf_destroy", None, [_vp]),
    ("kf_last_error", C.c_char_p, [_vp]),
    ("kf_synchronize", _i32, [_vp]),
    ("kf_set_sync_timeout", _i32, [_vp, _i64]),
    ("kf_reset", _i32, [_vp]),
    ("kf_pyramid_build", _i32, [_vp, _vp, C.POINTER(_i32)]),
    ("kf_pyramid_level", _i32, [_vp, _i32, C.POINTER(KfLevelView)]),
    ("kf_integrate", _i32, [_vp, C.POINTER(KfDepthFrame), C.POINTER(_i64)]),
    ("kf_raycast", _i32, [_vp, _dp, _i32, C.POINTER(KfMapsView)]),
    ("kf_icp_reduce", _i32, [_vp, _vp, _vp, _i32, _i32, _vp, _vp, _i32, _i32, C.POINTER(_f32), C.POINTER(_f32),
                             C.POINTER(_f32), _f32, _f32, _dp, C.POINTER(_i64)]),
    ("kf_icp_reduce_static", _i32, [_vp, _vp, _i32, _i32, _vp, _vp, _i32, _i32, C.POINTER(_f32), C.POINTER(_f32),
                                    C.POINTER(_f32), _f32, _f32, _vp, _vp, _vp]),
    ("kf_icp_track_maps", _i32, [_vp, _vp, _vp, _i32, _i32, _vp, _vp, _i32, _i32, C.POINTER(_f32), _dp, _dp, _i32,
                                 C.POINTER(KfIcpOptions), C.POINTER(KfIcpResult)]),
    ("kf_track", _i32, [_vp, C.POINTER(KfIcpOptions), _dp, C.POINTER(KfIcpResult)]),
    ("kf_process_frame", _i32, [_vp, _vp, _vp, C.POINTER(KfIcpOptions), C.POINTER(KfFrameResult)]),
    ("kf_trajectory", _i32, [_vp, _dp, _i64, C.POINTER(_i64)]),
    ("kf_last_stage_times", _i32, [_vp, C.POINTER(KfStageTimes)]),
    ("kf_fused_counts", _i32, [_vp, C.POINTER(_i64)]),
    ("kf_observed_voxel_count", _i32, [_vp, C.POINTER(_i64)]),
    ("kf_volume_bytes", _i32, [_vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    ("kf_volume_download", _i32, [_vp, _vp, _vp]),
    ("kf_volume_upload", _i32, [_vp, _vp, _vp]),
    ("kf_volume_device", _i32, [_vp, C.POINTER(_vp), C.POINTER(_vp)]),
    ("kf_mesh_extract", _i32, [_vp, _f32, C.POINTER(_vp)]),
    ("kf_mesh_size", _i32, [_vp, C.POINTER(_i64), C.POINTER(_i64)]),
    ("kf_mesh_copy", _i32, [_vp, _vp, _vp, _vp, _vp]),
    ("kf_mesh_free", None, [_vp]),
    ("kf_block_bitmaps", _i32, [_vp, _vp, _vp, _vp, C.POINTER(_i64)]),
    ("kf_hashed_stats", _i32, [_vp, C.POINTER(_i64)]),
    ("kf_hashed_blocks", _i32, [_vp, _vp, _i64, C.POINTER(_i64)]),
    ("kf_hashed_download", _i32, [_vp, _vp, _vp, _vp, _i64, C.POINTER(_i64)]),
    ("kf_posenet_config_defaults", None, [C.POINTER(KfPosenetConfig)]),
    ("kf_posenet_parameter_count", _i64, []),
    ("kf_posenet_create", _i32, [C.POINTER(KfPosenetConfig), C.POINTER(_vp)]),
    ("kf_posenet_destroy", None, [_vp]),
    ("kf_posenet_last_error", C.c_char_p, [_vp]),
    ("kf_posenet_get_parameters", _i32, [_vp, _vp]),
    ("kf_posenet_set_parameters", _i32, [_vp, _vp]),
    ("kf_posenet_forward", _i32, [_vp, _vp, _i32, _vp]),
    ("kf_posenet_gradients", _i32, [_vp, _vp, _vp, _i32, C.POINTER(_f32), _vp]),
    ("kf_posenet_train_step", _i32, [_vp, _vp, _vp, _i32, C.POINTER(_f32)]),
    ("kf_posenet_set_learning_rate", _i32, [_vp, _f32]),
    ("kf_posenet_synchronize", _i32, [_vp]),
    ("kf_attach_posenet", _i32, [_vp, _vp]),
]

SCENE_SYMBOLS = [
    ("kf_scene_create", _vp, [C.c_int, C.c_uint64, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                              C.c_double]),
    ("kf_scene_destroy", None, [_vp]),
    ("kf_scene_pose", None, [_vp, C.c_int, _dp]),
    ("kf_scene_render", None, [_vp, _dp, C.c_uint64, C.c_int, _vp, _vp, _vp]),
    ("kf_scene_description", _vp, [_vp, C.POINTER(C.c_size_t)]),
    ("kf_scene_render_device", C.c_int, [_vp, _dp, _vp, C.c_int, C.c_int, _vp, _vp, _vp]),
]

_hip: C.CDLL | None = None
_scene: C.CDLL | None = None


class KinectFusionError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def _bind(lib: C.CDLL, symbols) -> None:
    for name, res, args in symbols:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def hip() -> C.CDLL:
    """The HIP library; raises when it is not built (no fallback exists)."""
    global _hip
    if _hip is None:
        if not HIP_LIB.exists():
            raise ImportError(f"{HIP_LIB} is not built; run kinectfusion_amd/build.py (there is no CPU fallback)")
        lib = C.CDLL(str(HIP_LIB))
        _bind(lib, HIP_SYMBOLS)
        _hip = lib
    return _hip


def scene_lib() -> C.CDLL:
    global _scene
    if _scene is None:
        if not SCENE_LIB.exists():
            raise ImportError(f"{SCENE_LIB} is not built; run kinectfusion_amd/build.py")
        lib = C.CDLL(str(SCENE_LIB))
        _bind(lib, SCENE_SYMBOLS)
        _scene = lib
    return _scene


def check(status: int, ctx=None) -> None:
    if status != KF_OK:
        msg = hip().kf_last_error(ctx)
        raise KinectFusionError(status, msg.decode() if msg else "")

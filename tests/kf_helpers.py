"""Shared helpers for the parity tests: configurations of the synthetic
workloads and conversions between the oracle's host arrays and the
library's device tensors."""
import numpy as np

INTR = (525.0, 525.0, 319.5, 239.5)
W, H = 640, 480


def config1_geometry(res=128, voxel=0.01, margin=0.5, trunc=0.05):
    """Volume of the 5-frame small workload: res^3 voxels, origin
    (-res*v/2, -res*v/2, -margin) (src/app/app_options.cpp:33-38)."""
    half = np.float32(0.5) * np.float32(res) * np.float32(voxel)
    return res, voxel, (-float(half), -float(half), -margin), trunc


def pipeline_config(K, res=128, voxel=0.01, margin=0.5, trunc=0.05, store="float", rule="angle-weighted",
                    color="float", radius=3, sigma_s=4.5, sigma_r=0.03, levels=3, **kw):
    r, v, origin, t = config1_geometry(res, voxel, margin, trunc)
    # sparse storage integrates its band (Pipeline::require_valid_storage)
    mode = "band" if kw.get("storage") == "hashed" else "full"
    return K.PipelineConfig(
        voxel=store, tsdf_rule=rule, color=color,
        volume=K.VolumeGeometry((r, r, r), v, origin, t),
        integration=K.TsdfIntegrationOptions(depth_scale=1000.0, mode=mode),
        depth=K.DepthProcessingOptions(levels=levels, bilateral_radius=radius, bilateral_spatial_sigma=sigma_s,
                                       bilateral_depth_sigma=sigma_r),
        intrinsics=K.CameraIntrinsics(*INTR), width=W, height=H, **kw)


def oracle_volume(ko, res=128, voxel=0.01, margin=0.5, trunc=0.05, kind=0, color=True):
    r, v, origin, t = config1_geometry(res, voxel, margin, trunc)
    return ko.Volume.create(r, v, origin, t, kind, color)


STORE_KIND = {"float": 0, "quantized": 1, "bf16": 2}
RULE = {"classic": 0, "angle-weighted": 1}


def to_dev(torch, a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint16:
        return torch.from_numpy(a.view(np.int16)).cuda()
    if a.dtype == np.uint32:
        return torch.from_numpy(a.view(np.int32)).cuda()
    return torch.from_numpy(a).cuda()


def f4(a3):
    """(H, W, 3) float32 -> (H, W, 4) float32 with w = 0."""
    out = np.zeros(a3.shape[:2] + (4,), np.float32)
    out[..., :3] = a3
    return out


def host(t):
    return t.detach().cpu().numpy()


def same_bits(a, b):
    """Bitwise equality of float arrays (NaN payloads compared as bits)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def first_mismatch(a, b):
    a32 = np.ascontiguousarray(a).view(np.uint32).reshape(-1)
    b32 = np.ascontiguousarray(b).view(np.uint32).reshape(-1)
    idx = np.nonzero(a32 != b32)[0]
    return (len(idx), idx[:5], a.reshape(-1)[idx[:5]], b.reshape(-1)[idx[:5]]) if len(idx) else None

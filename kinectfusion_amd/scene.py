"""Seeded synthetic RGB-D scenes (libkf_scene.so, csrc/scene.cpp).

kind "room": box room 1.2 m wide with 6 spheres 0.5-1.2 m from the first
camera; kind "hall": 6 m x 3 m x 4 m room with 20 boxes and 20 spheres.
Frames are uint16 millimetre depth + RGBA8 colour + exact FP64 poses.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

KINDS = {"room": 0, "hall": 1}


class SyntheticScene:
    def __init__(self, kind: str = "room", seed: int = 0, width: int = 640, height: int = 480,
                 fx: float = 525.0, fy: float = 525.0, cx: float = 319.5, cy: float = 239.5):
        self.width, self.height = width, height
        self.intrinsics = (fx, fy, cx, cy)
        self._lib = N.scene_lib()
        self._h = self._lib.kf_scene_create(KINDS[kind], seed, width, height, fx, fy, cx, cy)

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.kf_scene_destroy(self._h)
            self._h = None

    def pose(self, frame: int) -> np.ndarray:
        """Ground-truth camera_to_world of frame `frame` (FP64 4x4)."""
        out = np.zeros(16, np.float64)
        self._lib.kf_scene_pose(self._h, frame, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out.reshape(4, 4)

    def render(self, pose: np.ndarray, noise_key: int = 0, noise: bool = True, true_depth: bool = False):
        """(depth_mm uint16 (H, W), rgba uint32 (H, W)[, exact depth float32])."""
        p = np.ascontiguousarray(np.asarray(pose, np.float64).reshape(16))
        depth = np.empty((self.height, self.width), np.uint16)
        rgba = np.empty((self.height, self.width), np.uint32)
        exact = np.empty((self.height, self.width), np.float32) if true_depth else None
        self._lib.kf_scene_render(self._h, p.ctypes.data_as(C.POINTER(C.c_double)), noise_key, int(noise),
                                  depth.ctypes.data_as(C.c_void_p), rgba.ctypes.data_as(C.c_void_p),
                                  exact.ctypes.data_as(C.c_void_p) if exact is not None else None)
        return (depth, rgba, exact) if true_depth else (depth, rgba)

    def render_device(self, poses: np.ndarray, keys, noise: bool = True):
        """Render len(poses) views on the GPU (same per-pixel renderer as
        render()); returns a (n, H, W) int16 cuda tensor of u16 depth."""
        import torch

        poses = np.ascontiguousarray(np.asarray(poses, np.float64).reshape(-1, 16))
        keys = np.ascontiguousarray(np.asarray(keys, np.uint64).reshape(-1))
        n = poses.shape[0]
        if keys.shape[0] != n:
            raise ValueError("one noise key per pose")
        torch.cuda.synchronize()
        out = torch.empty((n, self.height, self.width), dtype=torch.int16, device="cuda")
        rc = self._lib.kf_scene_render_device(self._h, poses.ctypes.data_as(C.POINTER(C.c_double)),
                                              keys.ctypes.data_as(C.c_void_p), n, int(noise), out.data_ptr(),
                                              None, None)
        if rc != 0:
            raise RuntimeError(f"device render failed with HIP error {rc}")
        return out

    def frame(self, index: int, noise: bool = True):
        """(depth, rgba, pose) of trajectory frame `index`."""
        pose = self.pose(index)
        d, c = self.render(pose, noise_key=index, noise=noise)
        return d, c, pose

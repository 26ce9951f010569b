"""Frame sources of the reconstruction app.

TumSensor: VirtualSensor over a TUM RGB-D sequence directory
(src/kinectfusion/virtual_sensor.cpp:41-194, include/kinectfusion/
virtual_sensor.hpp): depth.txt and rgb.txt list "timestamp path" after three
header lines; frame i pairs the i-th depth and the i-th colour file; depth
PNGs are 16-bit grayscale in depth-scale units (5000 per metre), colour PNGs
become RGBA8; the timestamps are the depth file's; the depth camera is the
TUM calibration of the sequence's family (freiburg2 / freiburg3 by
directory name, freiburg1 otherwise), unless the directory holds a
camera.txt ("fx fy cx cy", written for generated sequences) or the caller
passes one.  Frames are decoded ahead on worker threads (4 in flight), or
all up front with preload.

SyntheticSensor: the seeded generator (kinectfusion_amd.scene) as a
sequence: "synthetic:<room|hall>[:seed]"; depth in TUM units, stamps i / 30 s,
ground-truth poses known.

write_tum_dataset: a generated sequence written as a TUM directory (rgb/,
depth/, rgb.txt, depth.txt, groundtruth.txt, camera.txt).
"""
from __future__ import annotations

from collections import deque
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

from ..pipeline import CameraIntrinsics
from . import png

# TUM RGB-D calibrated depth intrinsics per camera family (virtual_sensor.cpp:18-26)
TUM_FREIBURG1 = CameraIntrinsics(517.3, 516.5, 318.6, 255.3)
TUM_FREIBURG2 = CameraIntrinsics(520.9, 521.0, 325.1, 249.7)
TUM_FREIBURG3 = CameraIntrinsics(535.4, 539.2, 320.1, 247.6)
TUM_DEPTH_SCALE = 5000.0
PREFETCH_DEPTH = 4


def tum_depth_intrinsics(dataset_dir) -> CameraIntrinsics:
    name = str(dataset_dir)
    if "freiburg2" in name:
        return TUM_FREIBURG2
    if "freiburg3" in name:
        return TUM_FREIBURG3
    return TUM_FREIBURG1


def read_index(path) -> tuple[list[float], list[str]] | None:
    """(timestamps, relative paths) of a TUM index file after its three
    header lines; None when the file cannot be read."""
    try:
        lines = Path(path).read_text().splitlines()
    except OSError:
        return None
    tokens = " ".join(lines[3:]).split()
    stamps, files = [], []
    for i in range(0, len(tokens) - 1, 2):
        try:
            stamps.append(float(tokens[i]))
        except ValueError:
            break
        files.append(tokens[i + 1])
    return stamps, files


class TumSensor:
    def __init__(self, dataset_dir, preload: bool = False, intrinsics: CameraIntrinsics | None = None):
        self.base = Path(dataset_dir)
        depth = read_index(self.base / "depth.txt")
        rgb = read_index(self.base / "rgb.txt")
        if depth is None or rgb is None or not depth[1] or not rgb[1]:
            raise FileNotFoundError(f"Failed to initialize dataset: {self.base}")
        self.depth_stamps, self.depth_files = depth
        self.color_files = rgb[1]
        self.count = min(len(self.depth_files), len(self.color_files))
        cam = self.base / "camera.txt"
        if intrinsics is not None:
            self.intrinsics = intrinsics
        elif cam.exists():
            self.intrinsics = CameraIntrinsics(*[float(v) for v in cam.read_text().split()[:4]])
        else:
            self.intrinsics = tum_depth_intrinsics(self.base)
        self.depth_scale = TUM_DEPTH_SCALE
        self.index = -1
        self._next = 0
        self._pool = ThreadPoolExecutor(max_workers=PREFETCH_DEPTH)
        self._prefetched: deque = deque()
        self._preloaded = None
        if preload:
            self._preloaded = list(self._pool.map(self.decode_frame, range(self.count)))
        else:
            self._schedule()
        self.depth = self.color = None

    def decode_frame(self, i: int):
        return (png.read_depth_png(self.base / self.depth_files[i]),
                png.read_color_png(self.base / self.color_files[i]))

    def _schedule(self) -> None:
        while len(self._prefetched) < PREFETCH_DEPTH and self._next < self.count:
            self._prefetched.append(self._pool.submit(self.decode_frame, self._next))
            self._next += 1

    def process_next_frame(self) -> bool:
        if self._preloaded is not None:
            if self.index + 1 >= len(self._preloaded):
                return False
            self.depth, self.color = self._preloaded[self.index + 1]
        else:
            if not self._prefetched:
                return False
            self.depth, self.color = self._prefetched.popleft().result()
            self._schedule()
        self.index += 1
        return True

    def timestamp(self) -> float:
        return self.depth_stamps[self.index] if 0 <= self.index < len(self.depth_stamps) else 0.0

    def groundtruth(self):
        """Path of groundtruth.txt, or None."""
        p = self.base / "groundtruth.txt"
        return p if p.exists() else None

    def close(self) -> None:
        self._pool.shutdown(wait=True, cancel_futures=True)


class SyntheticSensor:
    def __init__(self, spec: str, max_frames: int = 30):
        from ..scene import SyntheticScene

        parts = spec.split(":")
        kind = parts[1] if len(parts) > 1 else "room"
        seed = int(parts[2]) if len(parts) > 2 else 0
        self.scene = SyntheticScene(kind, seed=seed)
        self.intrinsics = CameraIntrinsics(*self.scene.intrinsics)
        self.depth_scale = TUM_DEPTH_SCALE
        self.count = max_frames if max_frames >= 0 else 300
        self.index = -1
        self.depth = self.color = None

    def process_next_frame(self) -> bool:
        if self.index + 1 >= self.count:
            return False
        self.index += 1
        d, c, _ = self.scene.frame(self.index)
        # millimetres -> TUM units (5 per millimetre), exact in uint16 below 13.1 m
        self.depth = (d.astype(np.uint32) * 5).astype(np.uint16)
        self.color = c
        return True

    def timestamp(self) -> float:
        return self.index / 30.0

    def groundtruth_poses(self, n: int) -> tuple[np.ndarray, np.ndarray]:
        return np.arange(n) / 30.0, np.stack([self.scene.pose(i) for i in range(n)])

    def close(self) -> None:
        pass


def write_tum_dataset(out_dir, kind: str = "room", seed: int = 0, frames: int = 30) -> Path:
    """A generated sequence as a TUM RGB-D directory."""
    from ..scene import SyntheticScene
    from ..trajectory import write_tum

    out = Path(out_dir)
    (out / "rgb").mkdir(parents=True, exist_ok=True)
    (out / "depth").mkdir(parents=True, exist_ok=True)
    scene = SyntheticScene(kind, seed=seed)
    stamps = 1305031102.0 + np.arange(frames) / 30.0
    header = "# generated by kinectfusion_amd: synthetic {} seed {}\n# file: '{}'\n# timestamp filename\n"
    dlines, clines, poses = [], [], []
    for i in range(frames):
        d, c, pose = scene.frame(i)
        name = f"{stamps[i]:.6f}.png"
        png.write_depth_png((d.astype(np.uint32) * 5).astype(np.uint16), out / "depth" / name)
        png.write_png(c, out / "rgb" / name)
        dlines.append(f"{stamps[i]:.6f} depth/{name}\n")
        clines.append(f"{stamps[i]:.6f} rgb/{name}\n")
        poses.append(pose)
    (out / "depth.txt").write_text(header.format(kind, seed, "depth.txt") + "".join(dlines))
    (out / "rgb.txt").write_text(header.format(kind, seed, "rgb.txt") + "".join(clines))
    write_tum(out / "groundtruth.txt", stamps, poses)
    k = scene.intrinsics
    (out / "camera.txt").write_text(f"{k[0]} {k[1]} {k[2]} {k[3]}\n")
    return out

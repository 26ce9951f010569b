"""Files the reconstruction app writes (src/app/frame_output.cpp:32-242):

* frame_NNNNNN_raycast.png            -- the raycast colour map, RGBA8
* frame_NNNNNN_raycast_point_cloud.ply -- binary little-endian PLY of the
  raycast pixels whose point and normal are finite: x y z nx ny nz (float)
  red green blue (uchar)
* mesh.ply                            -- the marching-cubes mesh, the same
  vertex properties plus faces as "list uchar uint vertex_indices"; colour
  channels are clamped to [0, 255] and truncated
* ablation_stats.csv                  -- one row per compared pipeline and
  frame, the reference's columns; the first write of a run truncates it

Frame files are encoded and written on background threads, at most 16 in
flight; an error surfaces at the next write_frame call or at close().
"""
from __future__ import annotations

from collections import deque
from concurrent.futures import Future, ThreadPoolExecutor
from pathlib import Path

import numpy as np

from . import png
from .pipeline_set import HostMaps, PipelineComparison

_VERTEX_PROPS = ("property float x\nproperty float y\nproperty float z\n"
                 "property float nx\nproperty float ny\nproperty float nz\n"
                 "property uchar red\nproperty uchar green\nproperty uchar blue\n")
_VERTEX = np.dtype([("p", "<f4", 3), ("n", "<f4", 3), ("c", "u1", 3)])
_FACE = np.dtype([("k", "u1"), ("v", "<u4", 3)])

CSV_HEADER = ("frame,pipeline,compared_voxels,volume_only_pipeline,volume_only_reference,max_distance_delta,"
              "mean_distance_delta,max_weight_delta,compared_pixels,surface_only_pipeline,surface_only_reference,"
              "max_point_distance,mean_point_distance,max_normal_angle,mean_normal_angle\n")


def frame_prefix(frame_index: int) -> str:
    return f"frame_{frame_index:06d}"


def write_point_cloud(maps: HostMaps, path) -> int:
    """Binary PLY of the valid raycast pixels; returns the vertex count."""
    ok = np.isfinite(maps.points).all(axis=-1) & np.isfinite(maps.normals).all(axis=-1)
    v = np.zeros(int(ok.sum()), _VERTEX)
    v["p"] = maps.points[ok]
    v["n"] = maps.normals[ok]
    rgba = maps.colors[ok].astype(np.uint32)
    v["c"] = np.stack([rgba & 0xFF, (rgba >> 8) & 0xFF, (rgba >> 16) & 0xFF], axis=1).astype(np.uint8)
    header = f"ply\nformat binary_little_endian 1.0\nelement vertex {len(v)}\n{_VERTEX_PROPS}end_header\n"
    with open(path, "wb") as f:
        f.write(header.encode())
        f.write(v.tobytes())
    return len(v)


def write_mesh_ply(mesh, path) -> None:
    """Binary PLY of a MarchingCubes mesh (positions, normals, colours
    (n, 3) float32; triangles (m, 3) uint32)."""
    v = np.zeros(len(mesh.positions), _VERTEX)
    v["p"] = mesh.positions
    v["n"] = mesh.normals
    v["c"] = np.clip(np.asarray(mesh.colors, np.float32), 0.0, 255.0).astype(np.uint8)
    f_ = np.zeros(len(mesh.triangles), _FACE)
    f_["k"] = 3
    f_["v"] = mesh.triangles
    header = (f"ply\nformat binary_little_endian 1.0\nelement vertex {len(v)}\n{_VERTEX_PROPS}"
              f"element face {len(f_)}\nproperty list uchar uint vertex_indices\nend_header\n")
    with open(path, "wb") as f:
        f.write(header.encode())
        f.write(v.tobytes())
        f.write(f_.tobytes())


def read_ply(path) -> tuple[dict, np.ndarray, np.ndarray | None]:
    """(header counts, vertices, faces or None) of a PLY written above."""
    data = Path(path).read_bytes()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    counts = {}
    for line in data[:end].decode().splitlines():
        if line.startswith("element "):
            _, name, n = line.split()
            counts[name] = int(n)
    nv = counts.get("vertex", 0)
    verts = np.frombuffer(data, _VERTEX, nv, end)
    faces = np.frombuffer(data, _FACE, counts["face"], end + nv * _VERTEX.itemsize) if "face" in counts else None
    return counts, verts, faces


def _fmt(v) -> str:
    """std::format("{}") of an integer or a float: the shortest form that
    reads back to the same float, fixed or scientific, fixed on a tie."""
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    x = np.float32(v)
    fixed = np.format_float_positional(x, unique=True, trim="-")
    sci = np.format_float_scientific(x, unique=True, trim="-", exp_digits=2)
    return sci if len(sci) < len(fixed) else fixed


def ablation_row(frame_index: int, c: PipelineComparison) -> str:
    vol = c.volume
    cols = [frame_index, c.name, vol.compared_voxels, vol.only_primary, vol.only_reference, vol.max_distance_delta,
            vol.mean_distance_delta, vol.max_weight_delta]
    s = c.surface
    if s is not None:
        cols += [s.compared_pixels, s.only_primary, s.only_reference, s.max_point_distance, s.mean_point_distance,
                 s.max_normal_angle, s.mean_normal_angle]
    else:
        cols += [""] * 7
    return ",".join(v if isinstance(v, str) else _fmt(v) for v in cols) + "\n"


class FrameOutput:
    MAX_PENDING = 16

    def __init__(self, output_dir, write_raycast_images: bool = True, write_point_clouds: bool = True):
        self.output_dir = Path(output_dir)
        self.write_raycast_images = write_raycast_images
        self.write_point_clouds = write_point_clouds
        self._stats_started = False
        self._pool = ThreadPoolExecutor(max_workers=4)
        self._pending: deque[Future] = deque()

    def writes_frames(self) -> bool:
        return self.write_raycast_images or self.write_point_clouds

    def write_frame(self, maps: HostMaps, frame_index: int, subdirectory: str = "") -> None:
        if not self.writes_frames():
            return
        while len(self._pending) >= self.MAX_PENDING:
            self._pending.popleft().result()
        d = self.output_dir / subdirectory if subdirectory else self.output_dir
        d.mkdir(parents=True, exist_ok=True)
        prefix = frame_prefix(frame_index)
        if self.write_raycast_images:
            self._pending.append(self._pool.submit(png.write_png, maps.colors, d / f"{prefix}_raycast.png"))
        if self.write_point_clouds:
            self._pending.append(self._pool.submit(write_point_cloud, maps, d / f"{prefix}_raycast_point_cloud.ply"))

    def append_ablation_stats(self, frame_index: int, comparisons: list[PipelineComparison]) -> None:
        if not comparisons:
            return
        self.output_dir.mkdir(parents=True, exist_ok=True)
        path = self.output_dir / "ablation_stats.csv"
        first = not self._stats_started
        with open(path, "w" if first else "a") as f:
            self._stats_started = True
            if first:
                f.write(CSV_HEADER)
            for c in comparisons:
                f.write(ablation_row(frame_index, c))

    def write_mesh(self, mesh, subdirectory: str = "") -> Path:
        d = self.output_dir / subdirectory if subdirectory else self.output_dir
        d.mkdir(parents=True, exist_ok=True)
        write_mesh_ply(mesh, d / "mesh.ply")
        return d / "mesh.ply"

    def close(self) -> None:
        try:
            while self._pending:
                self._pending.popleft().result()
        finally:
            self._pool.shutdown(wait=True)

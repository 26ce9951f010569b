"""Ablation harness: N pipelines on the same frames, compared with one of
them (the reference pipeline of the set).

PipelineSet       include/kinectfusion/pipeline_set.hpp:24-114,
                  src/kinectfusion/pipeline_set.cpp:24-178
Comparator        include/kinectfusion/comparison.hpp:11-90,
                  src/kinectfusion/comparison.cpp:12-81

A volume comparison reads every voxel of both pipelines through their
store's codec (tsdf(), weight_value(); volume.hpp:28-146): a voxel is
compared when both weights are > 0 and counted as one-sided otherwise; the
statistics are the maximum and mean |delta tsdf| and the maximum |delta
weight|.  A surface comparison takes the pixels whose point and normal are
both finite on both sides: maximum and mean point distance and normal angle
(atan2(|a x b|, a . b) of the normalised normals).  Deviations are float,
sums double, as in the reference.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ..pipeline import Pipeline, PipelineConfig


@dataclass
class VolumeComparison:
    compared_voxels: int = 0
    only_primary: int = 0
    only_reference: int = 0
    max_distance_delta: float = 0.0
    mean_distance_delta: float = 0.0
    max_weight_delta: float = 0.0


@dataclass
class SurfaceMapsComparison:
    compared_pixels: int = 0
    only_primary: int = 0
    only_reference: int = 0
    max_point_distance: float = 0.0
    mean_point_distance: float = 0.0
    max_normal_angle: float = 0.0
    mean_normal_angle: float = 0.0


@dataclass
class PipelineComparison:
    name: str
    volume: VolumeComparison
    surface: SurfaceMapsComparison | None = None


@dataclass
class HostMaps:
    """SurfaceMaps on the host: points and normals (H, W, 3) float32 with
    NaN where invalid, colours (H, W) uint32 RGBA8."""
    points: np.ndarray
    normals: np.ndarray
    colors: np.ndarray

    @staticmethod
    def from_device(maps) -> "HostMaps":
        return HostMaps(maps.points.cpu().numpy()[..., :3].copy(), maps.normals.cpu().numpy()[..., :3].copy(),
                        maps.colors.cpu().numpy().view(np.uint32).copy())


@dataclass
class PipelineOutput:
    name: str
    maps: HostMaps


def decode_voxels(voxels: np.ndarray, store: str) -> tuple[np.ndarray, np.ndarray]:
    """(tsdf, weight) float32 per voxel of a host volume in its store's
    layout: QuantizedTsdf raw * (1 / 32767), Bf16Tsdf raw << 16,
    QuantizedWeight raw * (1 / 256)."""
    if store == "float":
        return voxels[:, 0], voxels[:, 1]
    w = voxels[:, 1].astype(np.float32) * np.float32(1.0 / 256.0)
    if store == "quantized":
        t = voxels[:, 0].view(np.int16).astype(np.float32) * np.float32(1.0 / 32767.0)
    else:
        t = (voxels[:, 0].astype(np.uint32) << np.uint32(16)).view(np.float32)
    return t, w


def _mean(total: float, count: int) -> float:
    return float(np.float32(total / count)) if count else 0.0


class Comparator:
    @staticmethod
    def compare_volumes(primary: tuple[np.ndarray, np.ndarray], reference: tuple[np.ndarray, np.ndarray]
                        ) -> VolumeComparison:
        """Arguments: (tsdf, weight) float32 arrays of equal length."""
        (t0, w0), (t1, w1) = primary, reference
        if t0.shape != t1.shape:
            raise ValueError("Volume comparison requires matching geometry")
        p, r = w0 > 0.0, w1 > 0.0
        both = p & r
        dd = np.abs(t0[both] - t1[both]).astype(np.float32)
        dw = np.abs(w0[both] - w1[both]).astype(np.float32)
        n = int(both.sum())
        return VolumeComparison(compared_voxels=n, only_primary=int((p & ~r).sum()),
                                only_reference=int((r & ~p).sum()),
                                max_distance_delta=float(dd.max()) if n else 0.0,
                                mean_distance_delta=_mean(float(dd.astype(np.float64).sum()), n),
                                max_weight_delta=float(dw.max()) if n else 0.0)

    @staticmethod
    def compare_maps(primary: HostMaps, reference: HostMaps) -> SurfaceMapsComparison:
        if primary.points.shape != reference.points.shape or primary.normals.shape != primary.points.shape or \
                reference.normals.shape != reference.points.shape:
            raise ValueError("Surface map comparison requires matching dimensions")

        def valid(m: HostMaps):
            return np.isfinite(m.points).all(axis=-1) & np.isfinite(m.normals).all(axis=-1)

        p, r = valid(primary), valid(reference)
        both = p & r
        a, b = primary.points[both].astype(np.float32), reference.points[both].astype(np.float32)
        d = a - b
        dist = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])

        def unit(v):
            n = np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2])
            return v / n[:, None]

        na, nb = unit(primary.normals[both].astype(np.float32)), unit(reference.normals[both].astype(np.float32))
        cr = np.cross(na, nb)
        cn = np.sqrt((cr[:, 0] * cr[:, 0] + cr[:, 1] * cr[:, 1]) + cr[:, 2] * cr[:, 2])
        dot = (na[:, 0] * nb[:, 0] + na[:, 1] * nb[:, 1]) + na[:, 2] * nb[:, 2]
        ang = np.arctan2(cn, dot).astype(np.float32)
        n = int(both.sum())
        return SurfaceMapsComparison(compared_pixels=n, only_primary=int((p & ~r).sum()),
                                     only_reference=int((r & ~p).sum()),
                                     max_point_distance=float(dist.max()) if n else 0.0,
                                     mean_point_distance=_mean(float(dist.astype(np.float64).sum()), n),
                                     max_normal_angle=float(ang.max()) if n else 0.0,
                                     mean_normal_angle=_mean(float(ang.astype(np.float64).sum()), n))


class PipelineSet:
    """Pipelines sharing one volume geometry, one of them the reference that
    the others are compared with."""

    def __init__(self, members: list[Pipeline], reference_index: int, compare_every_n_frames: int):
        self.members = members
        self.reference_index = reference_index
        self.compare_every_n_frames = compare_every_n_frames

    @staticmethod
    def create(pipelines: list[PipelineConfig], reference: str = "", compare_every_n_frames: int = 1
               ) -> "PipelineSet":
        """ValueError on an empty set, duplicate names, an unknown reference
        or mismatched volume geometries (pipeline_set.cpp:33-79).  There is
        no fallback to another space."""
        if not pipelines:
            raise ValueError("Pipeline set requires at least one pipeline")
        names = [p.name for p in pipelines]
        if len(set(names)) != len(names):
            raise ValueError("Pipeline set requires unique pipeline names")
        if any(p.volume != pipelines[0].volume for p in pipelines):
            raise ValueError("Pipeline set requires matching volume geometries")
        index = 0
        if reference:
            if reference not in names:
                raise ValueError("Pipeline set reference must name a configured pipeline")
            index = names.index(reference)
        for p in pipelines:
            p.native()  # unsupported spaces and backends raise before any pipeline is created
        members = []
        try:
            for p in pipelines:
                members.append(Pipeline(p))
        except Exception:
            for m in members:
                m.close()
            raise
        return PipelineSet(members, index, compare_every_n_frames)

    def close(self) -> None:
        for m in self.members:
            m.close()

    def __len__(self) -> int:
        return len(self.members)

    @property
    def reference(self) -> Pipeline:
        return self.members[self.reference_index]

    def should_compare(self, frame_index: int) -> bool:
        return len(self.members) > 1 and self.compare_every_n_frames > 0 and \
            frame_index % self.compare_every_n_frames == 0

    def integrate(self, depth, world_to_camera, color=None, normals=None) -> None:
        """Every member fuses the same frame at the same pose."""
        for m in self.members:
            m.integrate(depth, world_to_camera, color=color, normals=normals)

    def raycast_all(self, camera_to_world, level: int = 0) -> list[PipelineOutput]:
        return [PipelineOutput(m.config.name, HostMaps.from_device(m.raycast(camera_to_world, level)))
                for m in self.members]

    def compare(self, outputs: list[PipelineOutput] | None = None) -> list[PipelineComparison]:
        if outputs is not None and len(outputs) != len(self.members):
            raise ValueError("Pipeline comparison requires one raycast output per pipeline")
        ref = self.reference
        ref_vox, _ = ref.host_view()
        ref_dec = decode_voxels(ref_vox, ref.config.voxel)
        out = []
        for i, m in enumerate(self.members):
            if i == self.reference_index:
                continue
            vox, _ = m.host_view()
            cmp = PipelineComparison(m.config.name, Comparator.compare_volumes(decode_voxels(vox, m.config.voxel),
                                                                               ref_dec))
            if outputs is not None:
                cmp.surface = Comparator.compare_maps(outputs[i].maps, outputs[self.reference_index].maps)
            out.append(cmp)
        return out

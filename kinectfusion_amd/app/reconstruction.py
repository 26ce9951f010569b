"""The reconstruction app's frame loop (src/app/reconstruction.cpp:75-440,
src/app/main.cpp): a pipeline set over one frame source.

* the first frame integrates into every pipeline at the identity;
* each later frame: every pipeline tracks against its own model from its own
  previous pose and integrates only when tracking succeeded (a rejected one
  keeps its last good pose and enters relocalization, pausing for input with
  --interactive-relocalization);
* per frame: the raycast at the reference pipeline's pose is written (PNG and
  point cloud) for every pipeline and, every --compare-every frames, the
  pipelines are compared with the reference into ablation_stats.csv;
* at the end: trajectory.txt per pipeline (TUM format), the ATE of each
  against groundtruth.txt when the dataset has one, mesh.ply per pipeline,
  and the frame rate of the loop.
Each pipeline's frame (pyramid, tracking, integration) runs in libkf_hip on
its own HIP stream; the reference pipeline writes into the output directory,
the others into a subdirectory named after them.
"""
from __future__ import annotations

import logging
import time
from dataclasses import dataclass, field

import numpy as np

from ..trajectory import AteEvaluator, write_tum
from .frame_output import FrameOutput
from .options import AppOptions
from .pipeline_config import PipelineSetConfig, parse_pipeline_set
from .pipeline_set import PipelineSet
from .sensor import SyntheticSensor, TumSensor

log = logging.getLogger("kinectfusion")


@dataclass
class Track:
    """Reconstruction::Track (reconstruction.hpp:31-40)."""
    icp: object
    trajectory: list = field(default_factory=list)
    relocalizing: bool = False
    relocalization_frames: int = 0


class Reconstruction:
    def __init__(self, options: AppOptions):
        self.options = options
        self.output = FrameOutput(options.output_dir, options.write_raycast_images, options.write_point_clouds)
        spec = str(options.dataset_dir)
        if spec.startswith("synthetic:"):
            self.sensor = SyntheticSensor(spec, options.max_frames)
        else:
            from ..pipeline import CameraIntrinsics

            k = CameraIntrinsics(*options.intrinsics) if options.intrinsics else None
            self.sensor = TumSensor(options.dataset_dir, options.preload_frames, k)
        self.pipelines: PipelineSet | None = None
        self.tracks: list[Track] = []
        self.processed_frames = 1
        self.net = None

    # ---- set-up -------------------------------------------------------------
    def _set_config(self, width: int, height: int) -> PipelineSetConfig:
        o = self.options
        base = o.pipeline_config(self.sensor.intrinsics, width, height)
        base.integration.depth_scale = self.sensor.depth_scale
        icp_default = (o.icp_options().damping, o.icp_lambda, o.icp_adaptive_damping)
        if o.pipelines_config is None:
            cfg = PipelineSetConfig([base], "", o.compare_every_n_frames, {base.name: icp_default})
        else:
            cfg = parse_pipeline_set(o.pipelines_config, base, o.compare_every_n_frames, icp_default)
        return cfg

    def _create(self) -> None:
        h, w = self.sensor.depth.shape
        cfg = self._set_config(w, h)
        self.pipelines = PipelineSet.create(cfg.pipelines, cfg.reference, cfg.compare_every_n_frames)
        from dataclasses import replace

        for p in cfg.pipelines:
            damping, lam, adaptive = cfg.icp_damping[p.name]
            self.tracks.append(Track(replace(self.options.icp_options(), damping=damping, damping_lambda=lam,
                                             adaptive_damping=adaptive)))
        if self.options.init == "learned":
            from .. import posenet as PN

            self.net = PN.PoseNet(PN.PoseNetConfig(depth_scale=self.sensor.depth_scale))
            self.net.set_parameters(np.load(self.options.posenet, allow_pickle=False))
            for m in self.pipelines.members:
                m.attach_posenet(self.net)
        for m in self.pipelines.members:
            log.info("Pipeline '%s' ready%s", m.config.name,
                     " (reference)" if len(self.pipelines) > 1 and m is self.pipelines.reference else "")

    # ---- frames ---------------------------------------------------------------
    def _device_frame(self):
        import torch

        d = torch.from_numpy(np.ascontiguousarray(self.sensor.depth).view(np.int16)).cuda()
        c = torch.from_numpy(np.ascontiguousarray(self.sensor.color).view(np.int32)).cuda()
        return d, c

    def _process(self, first: bool) -> None:
        depth, color = self._device_frame()
        stamp = self.sensor.timestamp()
        for i, (m, track) in enumerate(zip(self.pipelines.members, self.tracks)):
            res = m.process_frame(depth, color if m.config.color == "float" else None, icp=track.icp)
            pose = np.array(res.pose[:]).reshape(4, 4)
            if not first:
                if res.tracked:
                    if track.relocalizing:
                        log.info("Pipeline %d relocalized at frame %d after %d frame(s)", i, self.sensor.index,
                                 track.relocalization_frames)
                    track.relocalizing, track.relocalization_frames = False, 0
                else:
                    if not track.relocalizing:
                        log.warning("Pipeline %d entering relocalization mode at frame %d", i, self.sensor.index)
                    track.relocalizing = True
                    track.relocalization_frames += 1
                    log.warning("Frame %d pipeline %d tracking rejected at level %d", self.sensor.index, i,
                                res.failed_level)
                    if self.options.interactive_relocalization:
                        input("Relocalization paused. Align the live sensor with the last model prediction and "
                              "press Enter.\n")
            track.trajectory.append((stamp, pose))
        self._render_outputs()

    def _render_outputs(self) -> None:
        frame = self.sensor.index
        compare = self.pipelines.should_compare(frame)
        if not self.output.writes_frames() and not compare:
            return
        ref_pose = self.tracks[self.pipelines.reference_index].trajectory[-1][1]
        ref_name = self.pipelines.reference.config.name
        if not compare:
            from .pipeline_set import HostMaps

            self.output.write_frame(HostMaps.from_device(self.pipelines.reference.raycast(ref_pose, 0)), frame)
            return
        outputs = self.pipelines.raycast_all(ref_pose, 0)
        comparisons = self.pipelines.compare(outputs)
        for c in comparisons:
            log.info("Frame %d %s %s %s", frame, c.name, c.volume, c.surface)
        self.output.append_ablation_stats(frame, comparisons)
        for out in outputs:
            self.output.write_frame(out.maps, frame, "" if out.name == ref_name else out.name)

    # ---- end of run -------------------------------------------------------------
    def _member_dir(self, i: int):
        name = self.pipelines.members[i].config.name
        return self.options.output_dir if i == self.pipelines.reference_index else self.options.output_dir / name

    def write_trajectories(self) -> None:
        for i, track in enumerate(self.tracks):
            d = self._member_dir(i)
            d.mkdir(parents=True, exist_ok=True)
            write_tum(d / "trajectory.txt", [s for s, _ in track.trajectory], [p for _, p in track.trajectory])
            log.info("Wrote %d trajectory poses to %s", len(track.trajectory), d / "trajectory.txt")

    def report_ate(self) -> dict:
        if isinstance(self.sensor, SyntheticSensor):
            stamps, poses = self.sensor.groundtruth_poses(self.sensor.index + 1)
            evaluator = AteEvaluator(stamps=stamps, positions=poses[:, :3, 3])
        else:
            gt = self.sensor.groundtruth()
            evaluator = AteEvaluator(gt) if gt is not None else None
        out = {}
        if evaluator is None or not evaluator.has_groundtruth():
            log.info("No groundtruth.txt in %s; skipping ATE", self.options.dataset_dir)
            return out
        for m, track in zip(self.pipelines.members, self.tracks):
            s = evaluator.evaluate(track.trajectory)
            if s is None:
                print(f"ATE {m.config.name}: too few groundtruth associations")
                continue
            out[m.config.name] = s
            print(f"ATE {m.config.name}: pairs={s.pairs} rmse={s.rmse * 100:.3f} cm mean={s.mean * 100:.3f} cm "
                  f"median={s.median * 100:.3f} cm max={s.max_error * 100:.3f} cm")
        return out

    def write_meshes(self) -> None:
        ref = self.pipelines.reference.config.name
        for m in self.pipelines.members:
            mesh = m.extract_mesh(self.options.mesh_min_weight)
            self.output.write_mesh(mesh, "" if m.config.name == ref else m.config.name)
            log.info("Pipeline '%s' mesh: %d vertices, %d triangles", m.config.name, len(mesh.positions),
                     len(mesh.triangles))

    def run(self) -> int:
        try:
            if not self.sensor.process_next_frame():
                log.error("Dataset contains no frames")
                return 1
            self._create()
            self._process(first=True)
            start = time.perf_counter()
            while self.options.max_frames < 0 or self.processed_frames < self.options.max_frames:
                if not self.sensor.process_next_frame():
                    break
                self._process(first=False)
                self.processed_frames += 1
            seconds = time.perf_counter() - start
            self.output.close()
            self.write_trajectories()
            self.ate = self.report_ate()
            if self.options.write_mesh:
                self.write_meshes()
            log.info("Finished reconstruction: processed_frames=%d observed_voxels=%d", self.processed_frames,
                     self.pipelines.reference.observed_voxel_count())
            if seconds > 0.0:
                print(f"Frame loop: {self.processed_frames} frames in {seconds:.2f} s "
                      f"({self.processed_frames / seconds:.1f} fps)")
            return 0
        finally:
            self.output.close()
            if self.pipelines is not None:
                self.pipelines.close()
            if self.net is not None:
                self.net.close()
            self.sensor.close()

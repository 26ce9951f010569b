"""Learned ICP warm start: the pose-proposal network (kf_posenet_* in
include/kf_hip.h), its training data and its training loop.

The network maps the previous and current coarsest-level depth images
(160 x 120, the third level of the 640 x 480 pyramid) to the twist of the
camera motion between them, T_cur = T_prev [R(omega) | t].  Training pairs
come from the pipeline itself: seeded synthetic scenes rendered on the GPU,
the pipeline's own depth pyramid, and -- by default -- the converged ICP
estimate of each pair's motion as the label.  The reference has no learned
component; its tracker's initial pose is the hook
(src/app/reconstruction.cpp:218-250).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .pipeline import Pipeline, PipelineConfig, ProjectiveIcpOptions, VolumeGeometry, _torch
from .scene import SyntheticScene

IN_H, IN_W = 120, 160


# ---- twists ------------------------------------------------------------------------
def rotation_from_vector(w) -> np.ndarray:
    """Rodrigues: rotation of angle |w| about w / |w| (FP64)."""
    w = np.asarray(w, np.float64)
    th = float(np.linalg.norm(w))
    K = np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])
    if th < 1e-12:
        return np.eye(3) + K + 0.5 * K @ K
    return np.eye(3) + math.sin(th) / th * K + (1.0 - math.cos(th)) / (th * th) * K @ K


def vector_from_rotation(R) -> np.ndarray:
    """Inverse of rotation_from_vector for angles below pi."""
    R = np.asarray(R, np.float64)
    c = min(1.0, max(-1.0, 0.5 * (np.trace(R) - 1.0)))
    th = math.acos(c)
    v = np.array([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
    if th < 1e-9:
        return 0.5 * v
    return th / (2.0 * math.sin(th)) * v


def twist_to_pose(xi) -> np.ndarray:
    T = np.eye(4)
    T[:3, :3] = rotation_from_vector(xi[:3])
    T[:3, 3] = np.asarray(xi[3:6], np.float64)
    return T


def pose_to_twist(T) -> np.ndarray:
    T = np.asarray(T, np.float64)
    return np.concatenate([vector_from_rotation(T[:3, :3]), T[:3, 3]])


# ---- the network ----------------------------------------------------------------------
@dataclass
class PoseNetConfig:
    batch_capacity: int = 64
    seed: int = 0
    learning_rate: float = 1.0e-3
    beta1: float = 0.9
    beta2: float = 0.999
    epsilon: float = 1.0e-8
    depth_scale: float = 1000.0

    def native(self) -> N.KfPosenetConfig:
        c = N.KfPosenetConfig()
        N.hip().kf_posenet_config_defaults(C.byref(c))
        c.batch_capacity, c.seed = self.batch_capacity, self.seed
        c.learning_rate, c.beta1, c.beta2, c.epsilon = self.learning_rate, self.beta1, self.beta2, self.epsilon
        c.depth_scale = self.depth_scale
        return c


class PoseNet:
    """The BF16 MFMA pose network on cuda:0.  Inputs are device tensors:
    pairs (B, 2, 120, 160) int16 views of u16 raw depth, targets (B, 6)
    float32 twists."""

    def __init__(self, config: PoseNetConfig | None = None):
        self.config = config or PoseNetConfig()
        self._torch = _torch()
        net = C.c_void_p()
        native = self.config.native()
        status = N.hip().kf_posenet_create(C.byref(native), C.byref(net))
        if status != N.KF_OK:
            raise N.KinectFusionError(status, "kf_posenet_create failed")
        self._net = net
        self._held: list = []

    @property
    def handle(self):
        return self._net

    @staticmethod
    def parameter_count() -> int:
        return int(N.hip().kf_posenet_parameter_count())

    def close(self) -> None:
        if getattr(self, "_net", None):
            N.hip().kf_posenet_destroy(self._net)
            self._net = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, status: int) -> None:
        if status != N.KF_OK:
            msg = N.hip().kf_posenet_last_error(self._net)
            raise N.KinectFusionError(status, msg.decode() if msg else "")

    def _batch(self, pairs, targets=None):
        t = self._torch
        if pairs.dtype != t.int16 or tuple(pairs.shape[1:]) != (2, IN_H, IN_W) or not pairs.is_contiguous():
            raise ValueError("pairs must be a contiguous (B, 2, 120, 160) int16 device tensor")
        if targets is not None and (targets.dtype != t.float32 or tuple(targets.shape) != (pairs.shape[0], 6)
                                    or not targets.is_contiguous()):
            raise ValueError("targets must be a contiguous (B, 6) float32 device tensor")
        t.cuda.synchronize()  # inputs come from torch's stream
        return int(pairs.shape[0])

    def parameters(self) -> np.ndarray:
        out = np.empty(self.parameter_count(), np.float32)
        self._check(N.hip().kf_posenet_get_parameters(self._net, out.ctypes.data_as(C.c_void_p)))
        return out

    def set_parameters(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.parameter_count():
            raise ValueError("parameter vector has the wrong size")
        self._check(N.hip().kf_posenet_set_parameters(self._net, p.ctypes.data_as(C.c_void_p)))

    def forward(self, pairs):
        b = self._batch(pairs)
        out = self._torch.empty((b, 6), dtype=self._torch.float32, device="cuda")
        self._check(N.hip().kf_posenet_forward(self._net, pairs.data_ptr(), b, out.data_ptr()))
        return out

    def gradients(self, pairs, targets) -> tuple[float, np.ndarray]:
        b = self._batch(pairs, targets)
        g = np.empty(self.parameter_count(), np.float32)
        loss = C.c_float()
        self._check(N.hip().kf_posenet_gradients(self._net, pairs.data_ptr(), targets.data_ptr(), b, C.byref(loss),
                                                 g.ctypes.data_as(C.c_void_p)))
        return float(loss.value), g

    def train_step(self, pairs, targets, report: bool = False):
        """One Adam step; with report the loss is returned (and the call waits)."""
        b = self._batch(pairs, targets)
        loss = C.c_float() if report else None
        self._check(N.hip().kf_posenet_train_step(self._net, pairs.data_ptr(), targets.data_ptr(), b,
                                                  C.byref(loss) if report else None))
        self._held = [pairs, targets]
        return float(loss.value) if report else None

    def synchronize(self) -> None:
        self._check(N.hip().kf_posenet_synchronize(self._net))
        self._held = []


# ---- training data ---------------------------------------------------------------------
def random_motion(rng: np.random.Generator, max_translation: float, max_rotation_deg: float) -> np.ndarray:
    """Uniform direction; magnitude bound * u^2 (u uniform), so small motions,
    the common case between frames, are sampled densely."""
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    angle = math.radians(max_rotation_deg) * rng.uniform() ** 2
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    T = np.eye(4)
    T[:3, :3] = rotation_from_vector(axis * angle)
    T[:3, 3] = d * max_translation * rng.uniform() ** 2
    return T


def random_view(rng: np.random.Generator) -> np.ndarray:
    """A camera inside the synthetic room looking roughly down +z."""
    T = np.eye(4)
    T[:3, :3] = rotation_from_vector(np.radians(rng.uniform(-8.0, 8.0, 3)))
    T[:3, 3] = [rng.uniform(-0.15, 0.15), rng.uniform(-0.12, 0.12), rng.uniform(-0.25, 0.1)]
    return T


def generate_pairs(config: PipelineConfig, count: int, scene_seeds=tuple(range(1, 65)), seed: int = 0,
                   max_translation: float = 0.05, max_rotation_deg: float = 5.0, labels: str = "icp",
                   chunk: int = 128):
    """`count` (pairs, targets, stats) from seeded synthetic rooms, a new
    room (from scene_seeds, in turn) every `chunk` pairs.

    Each pair is a random view and the same view moved by up to
    max_translation / max_rotation_deg.  Both frames are rendered on the GPU,
    then go through the pipeline's own depth pyramid (config's depth options)
    to their coarsest level.  labels="icp": the target is the motion that
    level-0 ICP (kf_icp_track_maps, config's ICP defaults) converges to from
    the true motion, with the previous frame's maps as the model; pairs where
    ICP fails keep the true motion.  labels="truth": the true motion.
    """
    torch = _torch()
    if labels not in ("icp", "truth"):
        raise ValueError("labels must be 'icp' or 'truth'")
    rng = np.random.default_rng(seed)
    small = PipelineConfig(volume=VolumeGeometry((16, 16, 16), 0.01, (0.0, 0.0, 0.0), 0.05), depth=config.depth,
                           intrinsics=config.intrinsics, width=config.width, height=config.height,
                           integration=config.integration)
    prev_p, cur_p = Pipeline(small), Pipeline(small)
    k = config.intrinsics
    kvec = (C.c_float * 4)(k.fx, k.fy, k.cx, k.cy)
    opts = ProjectiveIcpOptions().native()
    pairs = torch.empty((count, 2, IN_H, IN_W), dtype=torch.int16, device="cuda")
    targets = np.empty((count, 6), np.float32)
    scenes = {s: SyntheticScene("room", seed=s, width=config.width, height=config.height, fx=k.fx, fy=k.fy,
                                cx=k.cx, cy=k.cy) for s in scene_seeds}
    icp_failed = 0
    diffs = []
    levels = config.depth.levels
    for start in range(0, count, chunk):
        n = min(chunk, count - start)
        sid = scene_seeds[(start // chunk) % len(scene_seeds)]
        base = [random_view(rng) for _ in range(n)]
        motion = [random_motion(rng, max_translation, max_rotation_deg) for _ in range(n)]
        cur = [b @ m for b, m in zip(base, motion)]
        keys = rng.integers(0, 2**62, size=2 * n, dtype=np.uint64)
        frames = scenes[sid].render_device(np.stack(base + cur), keys)
        for i in range(n):
            idx = start + i
            a, b = frames[i], frames[n + i]
            prev_p._check(N.hip().kf_pyramid_build(prev_p.ctx, a.data_ptr(), None))
            cur_p._check(N.hip().kf_pyramid_build(cur_p.ctx, b.data_ptr(), None))
            lp, lc = N.KfLevelView(), N.KfLevelView()
            prev_p._check(N.hip().kf_pyramid_level(prev_p.ctx, levels - 1, C.byref(lp)))
            cur_p._check(N.hip().kf_pyramid_level(cur_p.ctx, levels - 1, C.byref(lc)))
            prev_p.synchronize()
            cur_p.synchronize()
            nbytes = IN_H * IN_W * 2
            from .pipeline import _memcpy_d2d

            _memcpy_d2d(pairs[idx, 0].data_ptr(), lp.depth, nbytes)
            _memcpy_d2d(pairs[idx, 1].data_ptr(), lc.depth, nbytes)
            truth = motion[i]
            label = truth
            if labels == "icp":
                l0p, l0c = N.KfLevelView(), N.KfLevelView()
                prev_p._check(N.hip().kf_pyramid_level(prev_p.ctx, 0, C.byref(l0p)))
                cur_p._check(N.hip().kf_pyramid_level(cur_p.ctx, 0, C.byref(l0c)))
                model = (C.c_double * 16)(*np.eye(4).reshape(16))
                pose = (C.c_double * 16)(*truth.reshape(16))
                res = N.KfIcpResult()
                cur_p._check(N.hip().kf_icp_track_maps(cur_p.ctx, l0c.vertices, l0c.normals, l0c.width,
                                                       l0c.height, l0p.vertices, l0p.normals, l0p.width,
                                                       l0p.height, kvec, model, pose, 10, C.byref(opts),
                                                       C.byref(res)))
                if res.status in (0, 1):
                    label = np.array(pose[:], np.float64).reshape(4, 4)
                    diffs.append(np.linalg.norm(label[:3, 3] - truth[:3, 3]))
                else:
                    icp_failed += 1
            targets[idx] = pose_to_twist(label)
        del frames
    prev_p.close()
    cur_p.close()
    stats = {"icp_failed": icp_failed,
             "icp_vs_truth_m": float(np.median(diffs)) if diffs else None}
    return pairs, torch.from_numpy(targets).cuda(), stats


def train(net: PoseNet, pairs, targets, epochs: int, batch: int = 64, seed: int = 0, log=None,
          final_lr_fraction: float = 0.05):
    """Adam over shuffled batches with a cosine step-size decay from the
    configured rate to final_lr_fraction of it; returns the mean loss of each
    epoch."""
    torch = _torch()
    n = int(pairs.shape[0])
    rng = np.random.default_rng(seed)
    history = []
    steps_per_epoch = n // batch
    total = max(1, epochs * steps_per_epoch)
    lr0 = net.config.learning_rate
    step = 0
    for epoch in range(epochs):
        order = torch.from_numpy(rng.permutation(n)).cuda()
        losses = []
        for s in range(0, n - batch + 1, batch):
            lr = lr0 * (final_lr_fraction + (1.0 - final_lr_fraction) * 0.5 * (1.0 + math.cos(math.pi * step / total)))
            net._check(N.hip().kf_posenet_set_learning_rate(net.handle, lr))
            step += 1
            idx = order[s: s + batch]
            bp = pairs.index_select(0, idx).contiguous()
            bt = targets.index_select(0, idx).contiguous()
            report = (s // batch) % 16 == 0
            loss = net.train_step(bp, bt, report=report)
            if report:
                losses.append(loss)
        net.synchronize()
        history.append(float(np.mean(losses)))
        if log is not None:
            log(f"posenet epoch {epoch}: loss {history[-1]:.5f}")
    return history


def proposal_errors(net: PoseNet, pairs, targets) -> tuple[float, float]:
    """Median translation (m) and rotation (rad) error of the network's
    proposals against the targets."""
    torch = _torch()
    outs = []
    cap = net.config.batch_capacity
    for s in range(0, int(pairs.shape[0]), cap):
        outs.append(net.forward(pairs[s: s + cap].contiguous()).cpu().numpy())
    pred = np.concatenate(outs).astype(np.float64)
    tgt = targets.cpu().numpy().astype(np.float64)
    dt, dr = [], []
    for p, t in zip(pred, tgt):
        E = np.linalg.inv(twist_to_pose(p)) @ twist_to_pose(t)
        dt.append(np.linalg.norm(E[:3, 3]))
        dr.append(math.acos(min(1.0, max(-1.0, 0.5 * (np.trace(E[:3, :3]) - 1.0)))))
    del torch
    return float(np.median(dt)), float(np.median(dr))

"""Trajectory output and evaluation.

write_tum / read_tum: the TUM trajectory format the reference writes
(timestamp tx ty tz qx qy qz qw, nine decimals; src/app/reconstruction.cpp:
362-391).
AteEvaluator: absolute trajectory error against a TUM ground-truth file
(src/app/trajectory_eval.cpp:34-128): each estimated pose is associated with
the nearest ground-truth stamp within max_dt (0.02 s), fewer than three
associations give no result, the associated positions are aligned by the
closed-form rigid Umeyama solution (no scale), and the translational error
statistics are reported.
evaluate_ate_rpe: scripts/evaluate_ate.py: the same ATE, plus its RPE, the
frame-to-frame translational drift of consecutive associated frames with the
estimate's step rotated into the aligned frame (:96-105).
ate: ATE of two trajectories sampled at the same instants (pairs by index).
"""
from __future__ import annotations

import bisect
from dataclasses import dataclass

import numpy as np

DEFAULT_MAX_DT = 0.02     # AteEvaluator::kDefaultMaxDt
MIN_ASSOCIATIONS = 3      # AteEvaluator::kMinAssociations


def rotation_to_quaternion(r: np.ndarray) -> np.ndarray:
    """(qx, qy, qz, qw) of a rotation matrix (Shepperd's method)."""
    r = np.asarray(r, np.float64)
    t = np.trace(r)
    if t > 0.0:
        s = np.sqrt(t + 1.0) * 2.0
        q = [(r[2, 1] - r[1, 2]) / s, (r[0, 2] - r[2, 0]) / s, (r[1, 0] - r[0, 1]) / s, 0.25 * s]
    else:
        i = int(np.argmax(np.diag(r)))
        j, k = (i + 1) % 3, (i + 2) % 3
        s = np.sqrt(1.0 + r[i, i] - r[j, j] - r[k, k]) * 2.0
        q = [0.0, 0.0, 0.0, (r[k, j] - r[j, k]) / s]
        q[i] = 0.25 * s
        q[j] = (r[j, i] + r[i, j]) / s
        q[k] = (r[k, i] + r[i, k]) / s
    q = np.array(q)
    return q / np.linalg.norm(q)


def quaternion_to_rotation(q) -> np.ndarray:
    """Rotation matrix of (qx, qy, qz, qw) (scripts/evaluate_ate.py:32-39)."""
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def write_tum(path, stamps, poses) -> None:
    """One line per pose: stamp, position and unit quaternion, fixed nine
    decimals."""
    with open(path, "w") as f:
        for t, p in zip(stamps, poses):
            p = np.asarray(p, np.float64)
            q = rotation_to_quaternion(p[:3, :3])
            f.write(f"{t:.9f} {p[0, 3]:.9f} {p[1, 3]:.9f} {p[2, 3]:.9f} {q[0]:.9f} {q[1]:.9f} {q[2]:.9f} "
                    f"{q[3]:.9f}\n")


def read_tum(path) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(stamps, positions (n, 3), quaternions (n, 4)) of a TUM file in file
    order; comment lines and lines without eight numbers are skipped."""
    rows = []
    for line in open(path):
        line = line.strip()
        if not line or line.startswith("#"):
            continue
        parts = line.split()
        if len(parts) != 8:
            continue
        try:
            rows.append([float(v) for v in parts])
        except ValueError:
            continue
    a = np.array(rows, np.float64).reshape(-1, 8)
    return a[:, 0], a[:, 1:4], a[:, 4:8]


def umeyama_rigid(src: np.ndarray, dst: np.ndarray):
    """R, t minimising sum |R src_i + t - dst_i|^2 (Umeyama 1991 without
    scale; a reflection is turned into the nearest rotation)."""
    src = np.asarray(src, np.float64)
    dst = np.asarray(dst, np.float64)
    mu_s, mu_d = src.mean(0), dst.mean(0)
    sigma = (dst - mu_d).T @ (src - mu_s) / len(src)
    u, _, vt = np.linalg.svd(sigma)
    d = np.ones(3)
    if np.linalg.det(u) * np.linalg.det(vt) < 0.0:
        d[2] = -1.0
    r = (u * d) @ vt
    return r, mu_d - r @ mu_s


@dataclass
class AteStats:
    """Translational absolute trajectory error, metres
    (trajectory_eval.hpp:14-20)."""
    pairs: int
    rmse: float
    mean: float
    median: float
    max_error: float


def _stats(errors: np.ndarray) -> AteStats:
    return AteStats(pairs=int(len(errors)), rmse=float(np.sqrt(np.mean(errors ** 2))), mean=float(errors.mean()),
                    median=float(np.median(errors)), max_error=float(errors.max()))


class AteEvaluator:
    """AteEvaluator (trajectory_eval.cpp:34-128)."""

    def __init__(self, groundtruth_path=None, max_dt: float = DEFAULT_MAX_DT, stamps=None, positions=None):
        self.max_dt = max_dt
        if groundtruth_path is not None:
            try:
                stamps, positions, _ = read_tum(groundtruth_path)
            except OSError:
                stamps, positions = np.zeros(0), np.zeros((0, 3))
        order = np.argsort(np.asarray(stamps, np.float64), kind="stable")
        self.stamps = np.asarray(stamps, np.float64)[order]
        self.positions = np.asarray(positions, np.float64).reshape(-1, 3)[order]

    def has_groundtruth(self) -> bool:
        return len(self.stamps) > 0

    def associate(self, stamp: float) -> int | None:
        """Index of the nearest ground-truth sample within max_dt: of the
        first sample at or after `stamp` and the one before it, the closer
        one, the earlier one on a tie."""
        upper = bisect.bisect_left(self.stamps.tolist(), stamp)
        best, best_dt = None, self.max_dt
        for cand in (upper, upper if upper == 0 else upper - 1):
            if cand >= len(self.stamps):
                continue
            delta = abs(self.stamps[cand] - stamp)
            if delta <= best_dt:
                best, best_dt = cand, delta
        return best

    def evaluate(self, trajectory) -> AteStats | None:
        """trajectory: [(stamp, camera_to_world 4x4)]; None when fewer than
        three poses associate."""
        est, gt = [], []
        for stamp, pose in trajectory:
            j = self.associate(stamp)
            if j is not None:
                est.append(np.asarray(pose, np.float64)[:3, 3])
                gt.append(self.positions[j])
        if len(est) < MIN_ASSOCIATIONS:
            return None
        est, gt = np.array(est), np.array(gt)
        r, t = umeyama_rigid(est, gt)
        return _stats(np.linalg.norm(est @ r.T + t - gt, axis=1))


def evaluate_ate_rpe(groundtruth_path, trajectory_path, max_dt: float = DEFAULT_MAX_DT) -> dict | None:
    """scripts/evaluate_ate.py: ATE of the associated positions and the RPE
    of consecutive associated frames (metres, per frame); None below three
    associations."""
    g_st, g_pos, _ = read_tum(groundtruth_path)
    e_st, e_pos, _ = read_tum(trajectory_path)
    ground = dict(zip(g_st.tolist(), g_pos))
    estimate = dict(zip(e_st.tolist(), e_pos))
    g_sorted = np.array(sorted(ground))
    pairs = []
    for stamp in sorted(estimate):
        idx = int(np.searchsorted(g_sorted, stamp))
        best = None
        for cand in (idx - 1, idx):
            if 0 <= cand < len(g_sorted):
                dt = abs(g_sorted[cand] - stamp)
                if best is None or dt < best[0]:
                    best = (dt, g_sorted[cand])
        if best is not None and best[0] <= max_dt:
            pairs.append((best[1], stamp))
    if len(pairs) < MIN_ASSOCIATIONS:
        return None
    gt = np.array([ground[g] for g, _ in pairs])
    est = np.array([estimate[e] for _, e in pairs])
    r, t = umeyama_rigid(est, gt)
    errors = np.linalg.norm(est @ r.T + t - gt, axis=1)
    rel = np.array([np.linalg.norm(r @ (estimate[e1] - estimate[e0]) - (ground[g1] - ground[g0]))
                    for (g0, e0), (g1, e1) in zip(pairs, pairs[1:])])
    out = {"ate": _stats(errors), "pairs": len(pairs)}
    out["rpe_rmse"] = float(np.sqrt(np.mean(rel ** 2))) if len(rel) else 0.0
    out["rpe_mean"] = float(np.mean(rel)) if len(rel) else 0.0
    return out


def ate(estimated_poses: np.ndarray, ground_truth_poses: np.ndarray) -> dict:
    """ATE statistics (metres) of two trajectories sampled at the same
    instants, paired by index, after rigid alignment."""
    est = np.asarray(estimated_poses)[:, :3, 3]
    gt = np.asarray(ground_truth_poses)[:, :3, 3]
    r, t = umeyama_rigid(est, gt)
    s = _stats(np.linalg.norm(est @ r.T + t - gt, axis=1))
    return {"pairs": s.pairs, "rmse": s.rmse, "mean": s.mean, "median": s.median, "max": s.max_error}


def rpe(estimated_poses: np.ndarray, ground_truth_poses: np.ndarray) -> dict:
    """The reference script's RPE for trajectories paired by index: the
    frame-to-frame translational drift |R_align (p_e[i+1] - p_e[i]) -
    (p_g[i+1] - p_g[i])| (metres per frame), RMSE and mean."""
    est = np.asarray(estimated_poses)[:, :3, 3]
    gt = np.asarray(ground_truth_poses)[:, :3, 3]
    r, _ = umeyama_rigid(est, gt)
    rel = np.linalg.norm(np.diff(est, axis=0) @ r.T - np.diff(gt, axis=0), axis=1)
    return {"pairs": int(len(rel)), "rmse": float(np.sqrt(np.mean(rel ** 2))) if len(rel) else 0.0,
            "mean": float(np.mean(rel)) if len(rel) else 0.0}

"""Trajectory output and evaluation.

write_tum / read_tum: the TUM trajectory format the reference writes
(timestamp tx ty tz qx qy qz qw; src/app/reconstruction.cpp:362-391).
ate: absolute trajectory error after a rigid (no scale) Umeyama alignment of
the estimated positions onto the ground truth (AteEvaluator::evaluate,
src/app/trajectory_eval.cpp:77-128).
rpe: relative pose error over a fixed frame gap (scripts/evaluate_ate.py:96-105).
"""
from __future__ import annotations

import numpy as np


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


def write_tum(path, stamps, poses) -> None:
    with open(path, "w") as f:
        for t, p in zip(stamps, poses):
            q = rotation_to_quaternion(p[:3, :3])
            f.write(f"{t:.9f} {p[0, 3]:.9f} {p[1, 3]:.9f} {p[2, 3]:.9f} {q[0]:.9f} {q[1]:.9f} {q[2]:.9f} "
                    f"{q[3]:.9f}\n")


def read_tum(path):
    rows = [list(map(float, line.split())) for line in open(path) if line.strip() and not line.startswith("#")]
    a = np.array(rows)
    return a[:, 0], a[:, 1:4], a[:, 4:8]


def umeyama_rigid(src: np.ndarray, dst: np.ndarray):
    """R, t minimising |R src + t - dst| (Umeyama 1991, scale fixed at 1)."""
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


def ate(estimated_poses: np.ndarray, ground_truth_poses: np.ndarray) -> dict:
    """ATE statistics (metres) of camera positions after rigid alignment."""
    est = np.asarray(estimated_poses)[:, :3, 3]
    gt = np.asarray(ground_truth_poses)[:, :3, 3]
    r, t = umeyama_rigid(est, gt)
    err = np.linalg.norm(est @ r.T + t - gt, axis=1)
    return {"pairs": int(len(err)), "rmse": float(np.sqrt(np.mean(err ** 2))), "mean": float(err.mean()),
            "median": float(np.median(err)), "max": float(err.max())}


def rpe(estimated_poses: np.ndarray, ground_truth_poses: np.ndarray, delta: int = 1) -> dict:
    """Relative pose error over `delta` frames: translational RMSE (m) and
    mean rotational error (rad)."""
    est = np.asarray(estimated_poses)
    gt = np.asarray(ground_truth_poses)
    tr, rot = [], []
    for i in range(len(est) - delta):
        de = np.linalg.inv(est[i]) @ est[i + delta]
        dg = np.linalg.inv(gt[i]) @ gt[i + delta]
        e = np.linalg.inv(dg) @ de
        tr.append(np.linalg.norm(e[:3, 3]))
        rot.append(np.arccos(np.clip((np.trace(e[:3, :3]) - 1.0) / 2.0, -1.0, 1.0)))
    tr = np.array(tr)
    return {"pairs": int(len(tr)), "trans_rmse": float(np.sqrt(np.mean(tr ** 2))) if len(tr) else 0.0,
            "rot_mean": float(np.mean(rot)) if len(rot) else 0.0}

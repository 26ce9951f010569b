"""Trajectory evaluation and TUM I/O (src/app/trajectory_eval.cpp:34-128,
scripts/evaluate_ate.py, src/app/reconstruction.cpp:362-391) against the
oracle's numpy ATE and closed forms:

* a rigid copy of the ground truth has zero ATE and zero RPE (any rotation:
  an RPE that rotated the step the wrong way would not vanish);
* a copy scaled by 2 about the centroid aligns with the identity rotation,
  so its ATE is each point's distance to the centroid and its RPE each
  ground-truth step's length;
* association takes the nearest stamp within max_dt (the earlier one on a
  tie), drops the rest, and gives no result below three associations.
"""
import numpy as np
from scipy.spatial.transform import Rotation

from kinectfusion_amd import trajectory as T


def random_traj(n, seed):
    rng = np.random.default_rng(seed)
    out = np.tile(np.eye(4), (n, 1, 1))
    out[:, :3, :3] = Rotation.from_rotvec(rng.normal(scale=0.3, size=(n, 3))).as_matrix()
    out[:, :3, 3] = rng.normal(size=(n, 3))
    return out


def rigid(g, rotvec, t):
    m = np.eye(4)
    m[:3, :3] = Rotation.from_rotvec(rotvec).as_matrix()
    m[:3, 3] = t
    return np.einsum("ij,njk->nik", m, g)


def test_ate_matches_oracle_and_is_zero_for_rigid_copies(ko):
    gt = random_traj(40, 1)
    est = rigid(gt, [0.1, 0.5, -0.2], [1, -2, 0.5])
    assert T.ate(est, gt)["rmse"] < 1e-9
    noisy = est.copy()
    noisy[:, :3, 3] += np.random.default_rng(2).normal(scale=0.01, size=(40, 3))
    a = T.ate(noisy, gt)
    b = ko.ate_umeyama(noisy[:, :3, 3], gt[:, :3, 3])
    assert abs(a["rmse"] - b["rmse"]) < 1e-12 and a["pairs"] == 40


def test_rpe_is_the_scripts_frame_to_frame_drift():
    gt = random_traj(30, 3)
    est = rigid(gt, [0.4, -0.3, 0.9], [0.3, 0.1, -2.0])
    r = T.rpe(est, gt)
    assert r["pairs"] == 29 and r["rmse"] < 1e-9 and r["mean"] < 1e-9
    # scaled by 2 about the centroid: the alignment is the identity, so the
    # drift of each step is the step itself
    scaled = gt.copy()
    mu = gt[:, :3, 3].mean(0)
    scaled[:, :3, 3] = 2.0 * (gt[:, :3, 3] - mu) + mu
    steps = np.linalg.norm(np.diff(gt[:, :3, 3], axis=0), axis=1)
    r = T.rpe(scaled, gt)
    assert abs(r["rmse"] - np.sqrt(np.mean(steps ** 2))) < 1e-12 and abs(r["mean"] - steps.mean()) < 1e-12
    dist = np.linalg.norm(gt[:, :3, 3] - mu, axis=1)
    assert abs(T.ate(scaled, gt)["rmse"] - np.sqrt(np.mean(dist ** 2))) < 1e-12


def test_association_and_minimum_pairs(tmp_path):
    gt = random_traj(20, 5)
    stamps = 1000.0 + np.arange(20) / 32.0  # exact binary fractions: the midpoint below is a true tie
    T.write_tum(tmp_path / "gt.txt", stamps, gt)
    ev = T.AteEvaluator(tmp_path / "gt.txt")
    assert ev.has_groundtruth() and not T.AteEvaluator(tmp_path / "missing.txt").has_groundtruth()
    # nearest within 0.02 s; 0.03 s away associates with nothing
    assert ev.associate(stamps[4] + 0.01) == 4 and ev.associate(stamps[4] - 0.012) == 4
    assert ev.associate(stamps[0] - 0.03) is None and ev.associate(stamps[-1] + 0.03) is None
    # halfway between two samples (1/64 s from each): the earlier one
    assert ev.associate(0.5 * (stamps[6] + stamps[7])) == 6
    est = rigid(gt, [0.2, 0.1, 0.3], [0.5, 0.5, 0.5])
    traj = [(t + 0.004, p) for t, p in zip(stamps, est)]
    s = ev.evaluate(traj)
    assert s.pairs == 20 and s.rmse < 1e-6
    # only two stamps associate: no result
    far = [(t + 0.5 * (i >= 2) * 100, p) for i, (t, p) in enumerate(traj)]
    assert ev.evaluate(far) is None
    T.write_tum(tmp_path / "est.txt", [t for t, _ in traj], est)
    out = T.evaluate_ate_rpe(tmp_path / "gt.txt", tmp_path / "est.txt")
    assert out["pairs"] == 20 and out["ate"].rmse < 1e-6 and out["rpe_rmse"] < 1e-6


def test_tum_round_trip(tmp_path):
    poses = random_traj(5, 4)
    T.write_tum(tmp_path / "t.txt", np.arange(5) / 30.0, poses)
    stamps, pos, quat = T.read_tum(tmp_path / "t.txt")
    assert np.allclose(pos, poses[:, :3, 3], atol=1e-8)
    back = Rotation.from_quat(quat).as_matrix()
    assert np.allclose(back, poses[:, :3, :3], atol=1e-7)
    assert np.allclose(np.stack([T.quaternion_to_rotation(q) for q in quat]), back, atol=1e-12)

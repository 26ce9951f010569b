"""Trajectory evaluation and TUM I/O (src/app/trajectory_eval.cpp:77-128,
src/app/reconstruction.cpp:362-391) against the oracle's numpy ATE and
closed forms."""
import numpy as np
from scipy.spatial.transform import Rotation

from kinectfusion_amd import trajectory as T


def random_traj(n, seed):
    rng = np.random.default_rng(seed)
    out = np.tile(np.eye(4), (n, 1, 1))
    out[:, :3, :3] = Rotation.from_rotvec(rng.normal(scale=0.3, size=(n, 3))).as_matrix()
    out[:, :3, 3] = rng.normal(size=(n, 3))
    return out


def test_ate_matches_oracle_and_is_zero_for_rigid_copies(ko):
    gt = random_traj(40, 1)
    g = np.eye(4)
    g[:3, :3] = Rotation.from_rotvec([0.1, 0.5, -0.2]).as_matrix()
    g[:3, 3] = [1, -2, 0.5]
    est = np.einsum("ij,njk->nik", g, gt)
    assert T.ate(est, gt)["rmse"] < 1e-9
    noisy = est.copy()
    noisy[:, :3, 3] += np.random.default_rng(2).normal(scale=0.01, size=(40, 3))
    a = T.ate(noisy, gt)
    b = ko.ate_umeyama(noisy[:, :3, 3], gt[:, :3, 3])
    assert abs(a["rmse"] - b["rmse"]) < 1e-12 and a["pairs"] == 40


def test_rpe_of_identical_trajectories_is_zero():
    gt = random_traj(10, 3)
    r = T.rpe(gt, gt)
    assert r["trans_rmse"] < 1e-9 and r["rot_mean"] < 1e-6


def test_tum_round_trip(tmp_path):
    poses = random_traj(5, 4)
    T.write_tum(tmp_path / "t.txt", np.arange(5) / 30.0, poses)
    stamps, pos, quat = T.read_tum(tmp_path / "t.txt")
    assert np.allclose(pos, poses[:, :3, 3], atol=1e-8)
    back = Rotation.from_quat(quat).as_matrix()
    assert np.allclose(back, poses[:, :3, :3], atol=1e-7)

"""Parity at the benchmark workloads' full sizes, against the CPU oracle.

* 512^3 / 5 mm room (the precision workload): three frames integrated at the
  same poses into float, int16-quantized and bf16 stores give byte-identical
  volumes and colours and identical fused counts (tsdf_integration.hpp:
  183-252, volume.hpp:28-146); the plain FP32 pipeline tracks five frames
  within 1e-4 rad / 1e-4 m of the FP64 oracle's own run; the level-0 ICP
  sums (compensated FP32) stay within 1e-5 relative of FP64 sums over the
  same correspondences; with bf16 voxels and the learned warm start, every
  kept frame's pose equals the oracle's level-0 tracking (2 iterations) from
  the device's proposal on the identical volume.
* 256^3 / 1 cm room, 30 frames, classic and angle-weighted, float store, the
  full plain pipeline: each frame's pose within 1e-4 of the oracle tracking
  from the same previous pose against the same volume (the two run in lock
  step: the oracle integrates at the device's pose), the level-0 maps byte
  equal every 10 frames, the final volume byte equal, and the min-weight-2
  mesh identical (reconstruction.cpp:184-277, marching_cubes.cpp:367-507).
* 2048^3 / 5 mm hashed hall (2^20 buckets, 500 k blocks): ten frames spread
  over the sweep integrated at the same poses give the same allocated block
  set, block voxels and counts, and the same sparse-traversal mesh, as the
  oracle's slot-grid store (block_rep.hpp:183-307, marching_cubes.cpp:
  490-503); 300 frames of the full pipeline then never overflow, track
  within 5 mm ATE of the ground truth and repeat bit for bit.
"""
import numpy as np
import pytest

from kf_helpers import INTR, STORE_KIND, H, W, config1_geometry, f4, first_mismatch, host, pipeline_config, \
    same_bits, to_dev

pytestmark = pytest.mark.gpu

ITERS = {0: 10, 1: 5, 2: 4}


@pytest.fixture(scope="module")
def K():
    import kinectfusion_amd

    return kinectfusion_amd


def device_w2c(pose):
    """world_to_camera as the device forms it from a camera_to_world pose:
    [R^T, -R^T t] in FP64 with the products summed in row order, then
    rounded to float by the integration (the same IEEE operations, so the
    same float transform)."""
    a = np.asarray(pose, np.float64).reshape(16)
    out = np.eye(4)
    for r in range(3):
        for c in range(3):
            out[r, c] = a[c * 4 + r]
    for r in range(3):
        out[r, 3] = -((out[r, 0] * a[3] + out[r, 1] * a[7]) + out[r, 2] * a[11])
    return out


def pyramid(ko, depth):
    return ko.pyramid(depth, INTR, 1000.0, levels=3, radius=3, sigma_s=4.5, sigma_r=0.03)


def level_intr(level):
    s = np.float32(1.0 / (1 << level))
    return tuple(float(np.float32(v) * s) for v in INTR)


def pose_error(a, b):
    dr = a[:3, :3].T @ b[:3, :3]
    ang = float(np.arccos(np.clip((np.trace(dr) - 1) / 2, -1, 1)))
    return ang, float(np.linalg.norm(a[:3, 3] - b[:3, 3]))


def oracle_track(ko, vol, pyr, start, levels=(2, 1, 0), iters=ITERS, max_rms=None):
    """Reconstruction::track_pose on the oracle: per level, raycast the model
    at the running pose and run the level's device-loop ICP."""
    tracked = start.copy()
    for level in levels:
        if iters[level] == 0:
            continue
        k = level_intr(level)
        pts, nrm, _, _ = ko.raycast(vol, tracked, k, W >> level, H >> level)
        new_pose, r = ko.icp_track_level(pyr[level]["vertices"], pyr[level]["normals"], pts, nrm, k, tracked,
                                         tracked, iters[level])
        if r.status not in (0, 1):
            return None, r
        tracked = new_pose
    return tracked, r


def bench_config(K, store, res=512, voxel=0.005, **kw):
    return pipeline_config(K, res=res, voxel=voxel, store=store, **kw)


# --- 512^3 / 5 mm --------------------------------------------------------------------
@pytest.mark.parametrize("store", ["float", "quantized", "bf16"])
def test_512_integration_bit_exact(K, ko, torch_gpu, store):
    scene = K.SyntheticScene("room", seed=0)
    r, v, origin, t = config1_geometry(512, 0.005, 0.5, 0.05)
    vol = ko.Volume.create(r, v, origin, t, STORE_KIND[store], True)
    p = K.Pipeline(bench_config(K, store))
    for i in (0, 3, 6):
        depth, rgba, pose = scene.frame(i)
        nrm = pyramid(ko, depth)[0]["normals"]
        w2c = ko.rigid_inverse(pose)
        ref = ko.integrate(vol, 1, depth, INTR, w2c, ko.integration_options(1000.0), color=rgba, normals=nrm)
        got = p.integrate(to_dev(torch_gpu, depth), w2c, color=to_dev(torch_gpu, rgba),
                          normals=to_dev(torch_gpu, f4(nrm)), count=True)
        print(f"{store} frame {i}: fused {got} vs oracle {ref}", flush=True)
        assert tuple(got) == ref and ref[0] > 1_000_000
    vox, col = p.host_view()
    p.close()
    assert np.array_equal(vox.view(np.uint32), vol.voxels.view(np.uint32)), first_mismatch(vox, vol.voxels)
    assert same_bits(col, vol.colors), first_mismatch(col, vol.colors)


def test_512_plain_fp32_tracks_like_the_oracle(K, ko, torch_gpu):
    scene = K.SyntheticScene("room", seed=0)
    n = 6
    p = K.Pipeline(bench_config(K, "float"))
    for i in range(n):
        depth, rgba, _ = scene.frame(i)
        res = p.process_frame(to_dev(torch_gpu, depth), to_dev(torch_gpu, rgba))
        assert i == 0 or res.tracked == 1
    traj = p.trajectory()
    # the oracle's own run: frame 0 at the identity, then tracking from the previous pose
    r, v, origin, t = config1_geometry(512, 0.005, 0.5, 0.05)
    vol = ko.Volume.create(r, v, origin, t, ko.FLOAT, True)
    pose = np.eye(4)
    for i in range(n):
        depth, rgba, _ = scene.frame(i)
        pyr = pyramid(ko, depth)
        if i > 0:
            pose, _ = oracle_track(ko, vol, pyr, pose)
            assert pose is not None
        ko.integrate(vol, 1, depth, INTR, ko.rigid_inverse(pose), ko.integration_options(1000.0), color=rgba,
                     normals=pyr[0]["normals"])
        ang, dt = pose_error(traj[i], pose)
        print(f"frame {i}: device vs FP64 oracle {ang:.2e} rad {dt:.2e} m", flush=True)
        assert ang < 1e-4 and dt < 1e-4
    # level-0 normal equations of the next frame against the model at the last pose
    depth, _, _ = scene.frame(n)
    src = K.PyramidSource(p)
    src.build(to_dev(torch_gpu, depth))
    lv = src.level(0)
    maps = p.raycast(traj[-1], 0)
    tracker = K.ProjectiveIcpTracker(p)
    cam, mod = ko.pose12(traj[-1]), ko.pose12(ko.rigid_inverse(traj[-1]))
    sums, count = tracker.normal_equations(lv["vertices"], lv["normals"], maps.points, maps.normals,
                                           lv["intrinsics"], cam, mod)
    s64, s32, n_ref = ko.icp_sums(host(lv["vertices"]), host(lv["normals"]), host(maps.points), host(maps.normals),
                                  INTR, traj[-1], ko.rigid_inverse(traj[-1]))
    scale = np.abs(s64).max()
    err = np.abs(sums - s64).max() / scale
    err_seq = np.abs(s32.astype(np.float64) - s64).max() / scale
    print(f"level-0 ICP sums over {count} correspondences: compensated FP32 vs FP64 {err:.2e} "
          f"(sequential FP32 {err_seq:.2e})")
    assert count == n_ref and count > 100_000
    assert err < 1e-5
    p.close()


def test_512_bf16_learned_frames_match_oracle_tracking(K, ko, torch_gpu):
    from kinectfusion_amd import posenet as PN

    cfg = bench_config(K, "bf16", init="learned", learned_iterations=(2, 0, 0))
    pairs, targets, _ = PN.generate_pairs(cfg, 4096, seed=7)
    net = PN.PoseNet(PN.PoseNetConfig(seed=7))
    PN.train(net, pairs, targets, epochs=8, seed=7)
    del pairs, targets
    scene = K.SyntheticScene("room", seed=0)
    p = K.Pipeline(cfg)
    p.attach_posenet(net)
    r, v, origin, t = config1_geometry(512, 0.005, 0.5, 0.05)
    vol = ko.Volume.create(r, v, origin, t, ko.BF16, True)
    kept, n = 0, 8
    for i in range(n):
        depth, rgba, gt = scene.frame(i)
        res = p.process_frame(to_dev(torch_gpu, depth), to_dev(torch_gpu, rgba))
        pose = np.array(res.pose[:]).reshape(4, 4)
        pyr = pyramid(ko, depth)
        if i > 0:
            assert res.tracked == 1
            if res.used_learned:
                kept += 1
                proposal = np.array(res.proposal_pose[:]).reshape(4, 4)
                want, rr = oracle_track(ko, vol, pyr, proposal, levels=(0,), iters={0: 2})
                assert want is not None
                ang, dt = pose_error(pose, want)
                print(f"frame {i}: kept; device vs oracle from the proposal {ang:.2e} rad {dt:.2e} m; "
                      f"vs truth {np.linalg.norm(pose[:3, 3] - gt[:3, 3]) * 1000:.2f} mm", flush=True)
                assert ang < 1e-4 and dt < 1e-4
            else:
                print(f"frame {i}: fell back to the constant-velocity seed", flush=True)
        # the oracle integrates at the device's pose: identical volumes for the next frame
        ko.integrate(vol, 1, depth, INTR, device_w2c(pose), ko.integration_options(1000.0), color=rgba,
                     normals=pyr[0]["normals"])
    vox, _ = p.host_view()
    p.close()
    net.close()
    assert kept >= (n - 1) // 2
    assert np.array_equal(vox, vol.voxels), first_mismatch(vox.view(np.float32), vol.voxels.view(np.float32))


# --- 256^3 / 1 cm, 30 frames ---------------------------------------------------------
@pytest.mark.parametrize("rule", ["classic", "angle-weighted"])
def test_256_thirty_frames_lockstep(K, ko, torch_gpu, rule):
    rule_id = {"classic": 0, "angle-weighted": 1}[rule]
    scene = K.SyntheticScene("room", seed=0)
    p = K.Pipeline(pipeline_config(K, res=256, voxel=0.01, rule=rule))
    r, v, origin, t = config1_geometry(256, 0.01, 0.5, 0.05)
    vol = ko.Volume.create(r, v, origin, t, ko.FLOAT, True)
    src = K.PyramidSource(p)
    worst = (0.0, 0.0)
    for i in range(30):
        depth, rgba, gt = scene.frame(i)
        res = p.process_frame(to_dev(torch_gpu, depth), to_dev(torch_gpu, rgba))
        pose = np.array(res.pose[:]).reshape(4, 4)
        pyr = pyramid(ko, depth)
        assert same_bits(host(src.level(0)["normals"])[..., :3], pyr[0]["normals"]), f"frame {i}: pyramid differs"
        if i > 0:
            assert res.tracked == 1
            prev = p.trajectory()[i - 1]
            want, _ = oracle_track(ko, vol, pyr, prev)
            assert want is not None
            ang, dt = pose_error(pose, want)
            worst = (max(worst[0], ang), max(worst[1], dt))
            assert ang < 1e-4 and dt < 1e-4, (i, ang, dt)
            assert np.linalg.norm(pose[:3, 3] - gt[:3, 3]) < 0.01
        ko.integrate(vol, rule_id, depth, INTR, device_w2c(pose), ko.integration_options(1000.0), color=rgba,
                     normals=pyr[0]["normals"])
        if i % 10 == 9:
            maps = p.raycast(pose, 0)
            pts, nrm, col, _ = ko.raycast(vol, pose, INTR, W, H)
            assert same_bits(host(maps.points)[..., :3], pts), first_mismatch(host(maps.points)[..., :3], pts)
            assert same_bits(host(maps.normals)[..., :3], nrm)
            assert np.array_equal(host(maps.colors).view(np.uint32), col)
            print(f"{rule} frame {i}: level-0 maps byte-equal; worst pose gap so far {worst[0]:.2e} rad "
                  f"{worst[1]:.2e} m", flush=True)
    vox, colors = p.host_view()
    assert np.array_equal(vox.view(np.uint32), vol.voxels.view(np.uint32)), first_mismatch(vox, vol.voxels)
    assert same_bits(colors, vol.colors)
    gpu = p.extract_mesh(2.0)
    pos, nrm, col, tri = ko.mesh(vol, 2.0)
    print(f"{rule}: min-weight-2 mesh {len(pos)} vertices, {len(tri)} triangles")
    assert len(tri) > 10_000
    assert np.array_equal(gpu.triangles, tri)
    assert same_bits(gpu.positions, pos) and same_bits(gpu.normals, nrm) and same_bits(gpu.colors, col)
    p.close()


# --- 2048^3 hashed hall ---------------------------------------------------------------
HALL = dict(res=2048, voxel=0.005, margin=0.5, trunc=0.05)


def hall_config(K, **kw):
    return pipeline_config(K, res=HALL["res"], voxel=HALL["voxel"], store="float", color="none", storage="hashed",
                           hash_buckets=1 << 20, block_capacity=500_000, **kw)


def test_2048_hashed_hall_bit_exact(K, ko, torch_gpu):
    scene = K.SyntheticScene("hall", seed=0)
    r, v, origin, t = config1_geometry(HALL["res"], HALL["voxel"], HALL["margin"], HALL["trunc"])
    vol = ko.SparseVolume.create_sparse(r, v, origin, t, ko.FLOAT, False, capacity=500_000)
    p = K.Pipeline(hall_config(K))
    for i in range(0, 300, 30):
        depth, _, pose = scene.frame(i)
        nrm = pyramid(ko, depth)[0]["normals"]
        w2c = ko.rigid_inverse(pose)
        ref = ko.integrate_sparse(vol, 1, depth, INTR, w2c, ko.integration_options(1000.0), normals=nrm)
        got = p.integrate(to_dev(torch_gpu, depth), w2c, normals=to_dev(torch_gpu, f4(nrm)), count=True)
        stats = p.hashed_stats()
        print(f"frame {i}: fused {got} vs oracle {ref[:2]}; {stats['allocated']} blocks "
              f"(oracle {int(vol.state[0])}), {ref[2]} listed", flush=True)
        assert tuple(got) == ref[:2] and stats["overflow"] == 0
    keys, vox, _ = p.hashed_download()
    rkeys, rvox, _ = vol.blocks()
    assert np.array_equal(keys, rkeys) and len(keys) > 10_000
    assert np.array_equal(vox.view(np.uint32), rvox.view(np.uint32))
    # the raycast over metres of unobserved space (whole 32^3-voxel cells of
    # skip blocks, blocks without observed voxels next to observed ones)
    for i, level in ((135, 0), (135, 2), (290, 0)):
        pose = scene.pose(i)
        maps = p.raycast(pose, level)
        pts, nrm, col, _ = ko.raycast(vol, pose, level_intr(level), W >> level, H >> level)
        gp, gn = host(maps.points)[..., :3], host(maps.normals)[..., :3]
        print(f"raycast pose {i} level {level}: {np.isfinite(pts).all(-1).mean():.3f} of the rays hit", flush=True)
        assert np.isfinite(pts).all(-1).mean() > 0.9
        assert same_bits(gp, pts), first_mismatch(gp, pts)
        assert same_bits(gn, nrm), first_mismatch(gn, nrm)
        assert np.array_equal(host(maps.colors).view(np.uint32), col)
    gpu = p.extract_mesh(0.0)
    pos, nrm, col, tri = ko.mesh(vol, 0.0)
    print(f"hall mesh: {len(pos)} vertices, {len(tri)} triangles")
    assert len(tri) > 100_000
    assert np.array_equal(gpu.triangles, tri)
    assert same_bits(gpu.positions, pos) and same_bits(gpu.normals, nrm)
    p.close()


def test_2048_hashed_hall_300_frames(K, ko, torch_gpu):
    scene = K.SyntheticScene("hall", seed=0)
    n = 300
    poses = np.stack([scene.pose(i) for i in range(n)])
    depth = scene.render_device(poses, np.arange(n))

    def run():
        p = K.Pipeline(hall_config(K))
        failed = 0
        for i in range(n):
            res = p.process_frame(depth[i], None, report=(i % 50 == 49))
            if res is not None:
                failed += res.tracked == 0 and i > 0
        p.synchronize()
        traj, stats = p.trajectory(), p.hashed_stats()
        keys, vox, _ = p.hashed_download()
        p.close()
        return traj, stats, keys, vox, failed

    from kinectfusion_amd.trajectory import ate as ate_stats

    traj, stats, keys, vox, failed = run()
    ate = ate_stats(traj, poses)
    gt_err = max(np.linalg.norm(traj[i][:3, 3] - poses[i][:3, 3]) for i in range(n))
    print(f"hall: {stats['allocated']} blocks allocated, overflow {stats['overflow']}, ATE rmse "
          f"{ate['rmse'] * 1000:.2f} mm, max |t - t_gt| {gt_err * 1000:.2f} mm", flush=True)
    assert failed == 0 and stats["overflow"] == 0
    assert ate["rmse"] <= 0.005
    traj2, stats2, keys2, vox2, _ = run()
    assert np.array_equal(traj, traj2) and stats == stats2
    assert np.array_equal(keys, keys2) and np.array_equal(vox.view(np.uint32), vox2.view(np.uint32))

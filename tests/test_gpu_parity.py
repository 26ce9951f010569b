"""HIP library vs the CPU oracle on the same seeded synthetic frames.

Bars: integer and index work bit-exact; float TSDF, raycast maps and
vertex/normal maps bit-exact given identical inputs (|delta| <= 1e-5 is the
stated floor; these kernels meet it with zero difference); bilateral filter
bit-exact on the test frames (the device exponential is an FP64 one within
4e-16 rounded to float, the reference's is glibc expf); ICP sums within
1e-5 relative of the FP64 oracle; poses within 1e-4 rad / 1e-4 m.
"""
import numpy as np
import pytest

from kf_helpers import (INTR, RULE, STORE_KIND, H, W, f4, first_mismatch, host, oracle_volume, pipeline_config,
                        same_bits, to_dev)

pytestmark = pytest.mark.gpu

# 128^3 voxels of 2 cm span 2.56 m and hold the whole synthetic room; the
# 1 cm volume of the 5-frame workload only reaches z = 0.78 m.
ROOM_VOXEL = 0.02


@pytest.fixture(scope="module")
def K():
    import kinectfusion_amd

    return kinectfusion_amd


@pytest.fixture(scope="module")
def scene(K):
    return K.SyntheticScene("room", seed=0)


def frames(scene, n):
    return [scene.frame(i) for i in range(n)]


def oracle_pyramid(ko, depth, levels=3):
    return ko.pyramid(depth, INTR, 1000.0, levels=levels, radius=3, sigma_s=4.5, sigma_r=0.03)


# --- depth pyramid ---------------------------------------------------------------
@pytest.mark.parametrize("levels", [3, 2, 4])
def test_pyramid_parity(K, ko, scene, torch_gpu, levels):
    """3 levels: the one-launch pyramid kernel; 2 and 4: the per-level
    downsample and vertex/normal kernels."""
    depth, _, _ = scene.frame(3)
    p = K.Pipeline(pipeline_config(K, levels=levels))
    src = K.PyramidSource(p)
    assert src.build(to_dev(torch_gpu, depth)) == levels
    ref = oracle_pyramid(ko, depth, levels=levels)
    l0 = src.level(0)
    gpu_f = host(l0["depth"]).view(np.uint16)
    diff = np.abs(gpu_f.astype(np.int32) - ref[0]["depth"].astype(np.int32))
    # the FP64 table-driven exponential rounded to float reproduces the CPU's
    # float exponential on every weight of these frames: exact
    print(f"bilateral: {int((diff > 0).sum())} of {diff.size} pixels differ, max {diff.max()}")
    assert diff.max() == 0
    prev = gpu_f
    for level in range(levels):
        lv = src.level(level)
        d = host(lv["depth"]).view(np.uint16)
        if level > 0:
            assert np.array_equal(d, ko.downsample(prev, 1000.0))  # 2x2 rule, bit-exact
        k = ref[level]["intrinsics"]
        assert (lv["intrinsics"].fx, lv["intrinsics"].cx) == (float(k[0]), float(k[2]))
        v = ko.vertices(d, k, 1000.0)
        n = ko.normals(v)
        assert same_bits(host(lv["vertices"])[..., :3], v), first_mismatch(host(lv["vertices"])[..., :3], v)
        assert same_bits(host(lv["normals"])[..., :3], n), first_mismatch(host(lv["normals"])[..., :3], n)
        prev = d
    p.close()


@pytest.mark.parametrize("radius", [2, 5])
def test_bilateral_other_radii(K, ko, scene, torch_gpu, radius):
    """Radius 3 has an unrolled kernel instance; the others take the
    run-time-radius instance."""
    depth, _, _ = scene.frame(3)
    p = K.Pipeline(pipeline_config(K, radius=radius))
    src = K.PyramidSource(p)
    src.build(to_dev(torch_gpu, depth))
    ref = ko.pyramid(depth, INTR, 1000.0, levels=1, radius=radius, sigma_s=4.5, sigma_r=0.03)
    gpu_f = host(src.level(0)["depth"]).view(np.uint16)
    diff = np.abs(gpu_f.astype(np.int32) - ref[0]["depth"].astype(np.int32))
    print(f"bilateral r={radius}: {int((diff > 0).sum())} of {diff.size} pixels differ, max {diff.max()}")
    assert diff.max() == 0
    p.close()


def test_pyramid_without_filter_copies_raw(K, ko, scene, torch_gpu):
    depth, _, _ = scene.frame(1)
    p = K.Pipeline(pipeline_config(K, radius=0))
    src = K.PyramidSource(p)
    src.build(to_dev(torch_gpu, depth))
    assert np.array_equal(host(src.level(0)["depth"]).view(np.uint16), depth)
    p.close()


# --- integration ---------------------------------------------------------------------
@pytest.mark.parametrize("store", ["float", "quantized", "bf16"])
@pytest.mark.parametrize("rule", ["classic", "angle-weighted"])
@pytest.mark.parametrize("color", ["float", "none"])
def test_integration_bit_exact(K, ko, scene, torch_gpu, store, rule, color):
    p = K.Pipeline(pipeline_config(K, store=store, rule=rule, color=color))
    vol = oracle_volume(ko, kind=STORE_KIND[store], color=color == "float")
    for i, (depth, rgba, pose) in enumerate(frames(scene, 3)):
        ref = oracle_pyramid(ko, depth, levels=1)
        normals = ref[0]["normals"]
        w2c = ko.rigid_inverse(pose)
        n_ref = ko.integrate(vol, RULE[rule], depth, INTR, w2c, ko.integration_options(1000.0),
                             color=rgba, normals=normals)
        n_gpu = p.integrate(to_dev(torch_gpu, depth), w2c, color=to_dev(torch_gpu, rgba),
                            normals=to_dev(torch_gpu, f4(normals)), count=True)
        assert tuple(n_gpu) == n_ref, (i, n_gpu, n_ref)
    vox, col = p.host_view()
    assert np.array_equal(vox.view(np.uint8), vol.voxels.view(np.uint8)), first_mismatch(vox, vol.voxels)
    if color == "float":
        assert same_bits(col, vol.colors), first_mismatch(col, vol.colors)
    assert p.observed_voxel_count() == ko.observed_count(vol)
    p.close()


@pytest.mark.parametrize("store", ["float", "bf16"])
def test_integration_bit_exact_partial_bricks(K, ko, scene, torch_gpu, store):
    """A 100^3 grid: the last brick along x (16 voxels) and the last 8^3
    blocks along each axis are partial, so the cull, the per-wave brick
    walk and the part bits see grid edges inside a brick."""
    res, voxel = 100, 0.02
    p = K.Pipeline(pipeline_config(K, res=res, voxel=voxel, store=store))
    vol = oracle_volume(ko, res=res, voxel=voxel, kind=STORE_KIND[store])
    for i, (depth, rgba, pose) in enumerate(frames(scene, 3)):
        normals = oracle_pyramid(ko, depth, levels=1)[0]["normals"]
        w2c = ko.rigid_inverse(pose)
        n_ref = ko.integrate(vol, 1, depth, INTR, w2c, ko.integration_options(1000.0), color=rgba,
                             normals=normals)
        n_gpu = p.integrate(to_dev(torch_gpu, depth), w2c, color=to_dev(torch_gpu, rgba),
                            normals=to_dev(torch_gpu, f4(normals)), count=True)
        assert tuple(n_gpu) == n_ref, (i, n_gpu, n_ref)
    vox, col = p.host_view()
    assert np.array_equal(vox.view(np.uint8), vol.voxels.view(np.uint8)), first_mismatch(vox, vol.voxels)
    assert same_bits(col, vol.colors), first_mismatch(col, vol.colors)
    assert n_ref[0] > 0
    p.close()


def test_integration_distance_options(K, ko, scene, torch_gpu):
    """Non-projective distance and distance-scaled truncation take the same
    per-voxel path on both sides."""
    cfg = pipeline_config(K)
    cfg.integration.projective_distance = False
    cfg.integration.distance_scaled_truncation = True
    p = K.Pipeline(cfg)
    vol = oracle_volume(ko)
    opts = ko.integration_options(1000.0, projective_distance=0, distance_scaled_truncation=1)
    for depth, rgba, pose in frames(scene, 2):
        w2c = ko.rigid_inverse(pose)
        ko.integrate(vol, 1, depth, INTR, w2c, opts, color=rgba)
        p.integrate(to_dev(torch_gpu, depth), w2c, color=to_dev(torch_gpu, rgba))
    vox, _ = p.host_view()
    assert np.array_equal(vox.view(np.uint8), vol.voxels.view(np.uint8))
    p.close()


def test_integration_is_deterministic(K, scene, torch_gpu, ko):
    outs = []
    for _ in range(2):
        p = K.Pipeline(pipeline_config(K, store="bf16"))
        for depth, rgba, pose in frames(scene, 2):
            p.integrate(to_dev(torch_gpu, depth), ko.rigid_inverse(pose), color=to_dev(torch_gpu, rgba))
        outs.append(p.host_view())
        p.close()
    assert same_bits(outs[0][0], outs[1][0]) and same_bits(outs[0][1], outs[1][1])


# --- raycast -------------------------------------------------------------------------
@pytest.mark.parametrize("store,color,cell", [("float", "float", False), ("bf16", "none", False),
                                              ("quantized", "float", True)])
def test_raycast_bit_exact(K, ko, scene, torch_gpu, store, color, cell):
    cfg = pipeline_config(K, store=store, color=color, voxel=ROOM_VOXEL)
    cfg.raycast.cell_gradient_normals = cell
    p = K.Pipeline(cfg)
    vol = oracle_volume(ko, kind=STORE_KIND[store], color=color == "float", voxel=ROOM_VOXEL)
    for depth, rgba, pose in frames(scene, 2):
        w2c = ko.rigid_inverse(pose)
        ko.integrate(vol, 1, depth, INTR, w2c, ko.integration_options(1000.0), color=rgba)
        p.integrate(to_dev(torch_gpu, depth), w2c, color=to_dev(torch_gpu, rgba))
    pose = scene.pose(2)
    for level in range(3):
        s = 1 << level
        k = tuple(float(np.float32(v) * np.float32(1.0 / s)) for v in INTR)
        maps = p.raycast(pose, level)
        pts, nrm, col, _ = ko.raycast(vol, pose, k, W // s, H // s, cell_normals=cell)
        gp, gn = host(maps.points)[..., :3], host(maps.normals)[..., :3]
        valid = np.isfinite(pts).all(-1)
        assert valid.mean() > 0.5
        assert same_bits(gp, pts), first_mismatch(gp, pts)
        assert same_bits(gn, nrm), first_mismatch(gn, nrm)
        assert np.array_equal(host(maps.colors).view(np.uint32), col)
    p.close()


@pytest.mark.parametrize("store", ["float", "bf16"])
def test_raycast_through_holes_and_out_of_the_volume(K, ko, scene, torch_gpu, store):
    """The 1 cm volume ends at z = 0.78 m: most rays leave the box or cross
    unobserved blocks, the samples the raycaster steps over without reading
    the volume.  The maps must still equal the oracle's sample-by-sample march."""
    p = K.Pipeline(pipeline_config(K, store=store))
    vol = oracle_volume(ko, kind=STORE_KIND[store])
    for depth, rgba, pose in frames(scene, 2):
        w2c = ko.rigid_inverse(pose)
        ko.integrate(vol, 1, depth, INTR, w2c, ko.integration_options(1000.0), color=rgba)
        p.integrate(to_dev(torch_gpu, depth), w2c, color=to_dev(torch_gpu, rgba))
    for level, pose in ((2, scene.pose(2)), (1, scene.pose(1) @ np.diag([1.0, 1.0, 1.0, 1.0]))):
        s = 1 << level
        k = tuple(float(np.float32(v) * np.float32(1.0 / s)) for v in INTR)
        maps = p.raycast(pose, level)
        pts, nrm, col, samples = ko.raycast(vol, pose, k, W // s, H // s)
        valid = np.isfinite(pts).all(-1)
        print(f"level {level}: hit fraction {valid.mean():.3f}, oracle samples per ray {samples / valid.size:.0f}")
        assert 0.0 < valid.mean() < 0.9
        assert same_bits(host(maps.points)[..., :3], pts)
        assert same_bits(host(maps.normals)[..., :3], nrm)
        assert np.array_equal(host(maps.colors).view(np.uint32), col)
    p.close()


def test_raycast_after_volume_upload(K, ko, torch_gpu):
    """An uploaded volume rebuilds the occupancy bits the raycaster skips by."""
    import test_oracle_known_answers as T

    vol, _, _ = T.sphere_volume(ko)
    cfg = K.PipelineConfig(volume=K.VolumeGeometry((T.SYN_RES,) * 3, T.SYN_VOXEL, T.SYN_ORIGIN, T.SYN_TRUNC),
                           width=16, height=16, intrinsics=K.CameraIntrinsics(*T.SYN_INTR))
    p = K.Pipeline(cfg)
    p.upload_volume(vol.voxels, vol.colors)
    maps = p.raycast(np.eye(4), 0)
    pts, nrm, col, _ = ko.raycast(vol, np.eye(4), T.SYN_INTR, 16, 16)
    assert np.isfinite(pts).all(-1).any()
    assert same_bits(host(maps.points)[..., :3], pts) and same_bits(host(maps.normals)[..., :3], nrm)
    p.close()


# --- ICP -------------------------------------------------------------------------------
def tracking_setup(K, ko, scene, torch_gpu, n=2):
    p = K.Pipeline(pipeline_config(K, voxel=ROOM_VOXEL))
    vol = oracle_volume(ko, voxel=ROOM_VOXEL)
    for depth, rgba, pose in frames(scene, n):
        w2c = ko.rigid_inverse(pose)
        ko.integrate(vol, 1, depth, INTR, w2c, ko.integration_options(1000.0), color=rgba)
        p.integrate(to_dev(torch_gpu, depth), w2c, color=to_dev(torch_gpu, rgba))
    return p, vol


def test_icp_sums_against_fp64(K, ko, scene, torch_gpu):
    p, vol = tracking_setup(K, ko, scene, torch_gpu)
    depth, _, pose_true = scene.frame(2)
    src = K.PyramidSource(p)
    src.build(to_dev(torch_gpu, depth))
    model_pose = scene.pose(1)
    tracker = K.ProjectiveIcpTracker(p)
    for level in range(3):
        lv = src.level(level)
        k = lv["intrinsics"]
        maps = p.raycast(model_pose, level)
        cam = ko.pose12(model_pose)
        mod = ko.pose12(ko.rigid_inverse(model_pose))
        sums, count = tracker.normal_equations(lv["vertices"], lv["normals"], maps.points, maps.normals, k, cam, mod)
        s64, s32, n_ref = ko.icp_sums(host(lv["vertices"]), host(lv["normals"]), host(maps.points),
                                      host(maps.normals), (k.fx, k.fy, k.cx, k.cy), model_pose,
                                      ko.rigid_inverse(model_pose))
        assert count == n_ref and count > 1000
        scale = np.abs(s64).max()
        err_gpu = np.abs(sums - s64).max() / scale
        err_ref32 = np.abs(s32.astype(np.float64) - s64).max() / scale
        print(f"level {level}: n={count} compensated-FP32 rel err {err_gpu:.2e} "
              f"(sequential FP32 {err_ref32:.2e})")
        assert err_gpu < 1e-5
    p.close()


def test_icp_track_level_against_oracle(K, ko, scene, torch_gpu):
    p, vol = tracking_setup(K, ko, scene, torch_gpu)
    depth, _, pose_true = scene.frame(2)
    src = K.PyramidSource(p)
    src.build(to_dev(torch_gpu, depth))
    model_pose = scene.pose(1)
    tracker = K.ProjectiveIcpTracker(p)
    lv = src.level(1)
    k = lv["intrinsics"]
    maps = p.raycast(model_pose, 1)
    out = tracker.estimate_pose(lv["vertices"], lv["normals"], maps.points, maps.normals, k, model_pose,
                                model_pose, 5)
    ref_pose, r = ko.icp_track_level(host(lv["vertices"]), host(lv["normals"]), host(maps.points),
                                     host(maps.normals), (k.fx, k.fy, k.cx, k.cy), model_pose, model_pose, 5)
    assert out.ok and r.status in (0, 1)
    assert out.diagnostics.correspondences == r.correspondences
    assert np.abs(out.pose - ref_pose).max() < 1e-6
    # the stability spectrum: Householder tridiagonalisation and Sturm
    # multisection in one wave on the device, cyclic Jacobi in the oracle
    # (both FP64; the systems agree to float rounding)
    d = out.diagnostics
    assert r.min_eigenvalue > 0.0 and r.condition_number > 1.0
    assert abs(d.min_eigenvalue - r.min_eigenvalue) <= 1e-4 * r.min_eigenvalue
    assert abs(d.condition_number - r.condition_number) <= 1e-4 * r.condition_number
    p.close()


ICP_INTR = (100.0, 100.0, 1.0, 1.0)


def _patch(ko):
    import test_oracle_known_answers as T

    return T


@pytest.mark.parametrize("case", ["identity", "perturbation", "planar", "planar-damped", "empty"])
def test_reference_icp_cases_on_device(K, ko, torch_gpu, case):
    T = _patch(ko)
    p = K.Pipeline(pipeline_config(K, res=16))
    opts = K.ProjectiveIcpOptions(min_correspondences=6, max_point_distance=0.1, min_normal_dot=0.9,
                                  min_system_eigenvalue=1e-12, max_condition_number=1e12)
    init = np.eye(4)
    if case in ("identity", "perturbation"):
        v, n = T.identical_surface(ICP_INTR)
        if case == "perturbation":
            init = T.perturbed()
    elif case.startswith("planar"):
        v, n = T.planar_surface(ICP_INTR)
        opts.min_system_eigenvalue, opts.max_condition_number = 1e-6, 1e6
        if case == "planar-damped":
            opts.damping, opts.damping_lambda = "diagonal", 1e-3
    else:
        v = n = np.full((2, 2, 3), np.nan, np.float32)
        opts.min_correspondences = 1
    tv, tn = to_dev(torch_gpu, f4(v)), to_dev(torch_gpu, f4(n))
    tracker = K.ProjectiveIcpTracker(p, opts)
    out = tracker.estimate_pose(tv, tn, tv, tn, K.CameraIntrinsics(*ICP_INTR), np.eye(4), init, 15)
    if case == "identity":
        assert out.result == "Converged" and out.diagnostics.correspondences == 9
    elif case == "perturbation":
        assert out.result == "Converged" and np.linalg.norm(out.pose - np.eye(4)) <= 1e-3
        assert out.diagnostics.mean_distance <= 1e-4
    elif case == "planar":
        assert out.result == "kUnconstrainedSystem"
    elif case == "planar-damped":
        assert out.ok and out.diagnostics.correspondences == 9
    else:
        assert out.result == "kTooFewCorrespondences" and out.diagnostics.correspondences == 0
    p.close()


# --- whole frames -------------------------------------------------------------------
def oracle_run(ko, scene, n, store=0, res=128, voxel=ROOM_VOXEL, margin=0.5):
    """Reconstruction::initialize + process_frame on the oracle: integrate
    frame 0 at the identity, then track coarse-to-fine from the previous pose
    and integrate when every level succeeds."""
    vol = oracle_volume(ko, res=res, voxel=voxel, margin=margin, kind=store)
    iters = {0: 10, 1: 5, 2: 4}
    poses = []
    pose = np.eye(4)
    for i in range(n):
        depth, rgba, _ = scene.frame(i)
        pyr = oracle_pyramid(ko, depth)
        ok = True
        if i > 0:
            tracked = pose.copy()
            for level in (2, 1, 0):
                k = pyr[level]["intrinsics"]
                s = 1 << level
                pts, nrm, _, _ = ko.raycast(vol, tracked, k, W // s, H // s)
                new_pose, r = ko.icp_track_level(pyr[level]["vertices"], pyr[level]["normals"], pts, nrm, k,
                                                 tracked, tracked, iters[level])
                if r.status not in (0, 1):
                    ok = False
                    break
                tracked = new_pose
            pose = tracked
        if ok:
            ko.integrate(vol, 1, depth, INTR, ko.rigid_inverse(pose), ko.integration_options(1000.0), color=rgba,
                         normals=pyr[0]["normals"])
        poses.append(pose.copy())
    return np.array(poses), vol


def pose_error(a, b):
    dr = a[:3, :3].T @ b[:3, :3]
    ang = np.arccos(np.clip((np.trace(dr) - 1) / 2, -1, 1))
    return ang, np.linalg.norm(a[:3, 3] - b[:3, 3])


def test_frames_track_like_the_oracle(K, ko, scene, torch_gpu):
    """Five frames end to end on a 128^3 volume holding the room: per-frame
    pose within 1e-4 rad and 1e-4 m of the FP64 oracle, and within 1 cm of the
    ground truth."""
    n = 5
    p = K.Pipeline(pipeline_config(K, voxel=ROOM_VOXEL))
    for i in range(n):
        depth, rgba, _ = scene.frame(i)
        res = p.process_frame(to_dev(torch_gpu, depth), to_dev(torch_gpu, rgba))
        print(f"frame {i}: tracked={res.tracked} failed_level={res.failed_level} "
              + " ".join(f"L{l}:{res.level[l].status}/{res.level[l].trials}" for l in range(3)))
        if i > 0:
            assert res.tracked == 1
    traj = p.trajectory()
    ref, _ = oracle_run(ko, scene, n)
    for i in range(n):
        ang, dt = pose_error(traj[i], ref[i])
        gt_ang, gt_dt = pose_error(traj[i], scene.pose(i))
        print(f"frame {i}: vs oracle {ang:.2e} rad {dt:.2e} m; vs truth {gt_ang:.2e} rad {gt_dt:.2e} m")
        assert ang < 1e-4 and dt < 1e-4
        assert gt_dt < 0.01
    p.close()


# --- marching cubes -----------------------------------------------------------------
def compare_mesh(gpu, ref):
    pos, nrm, col, tri = ref
    assert gpu.triangles.shape == tri.shape and np.array_equal(gpu.triangles, tri)
    assert same_bits(gpu.positions, pos), first_mismatch(gpu.positions, pos)
    assert same_bits(gpu.normals, nrm), first_mismatch(gpu.normals, nrm)
    assert same_bits(gpu.colors, col), first_mismatch(gpu.colors, col)


@pytest.mark.parametrize("store", ["float", "quantized", "bf16"])
def test_mesh_bit_exact(K, ko, scene, torch_gpu, store):
    """Triangle topology, welded vertex numbering, positions, normals and
    colours identical to the CPU mesher (cell classification included)."""
    p = K.Pipeline(pipeline_config(K, store=store, voxel=ROOM_VOXEL))
    vol = oracle_volume(ko, kind=STORE_KIND[store], voxel=ROOM_VOXEL)
    for depth, rgba, pose in frames(scene, 3):
        w2c = ko.rigid_inverse(pose)
        ko.integrate(vol, 1, depth, INTR, w2c, ko.integration_options(1000.0), color=rgba)
        p.integrate(to_dev(torch_gpu, depth), w2c, color=to_dev(torch_gpu, rgba))
    for min_weight in (0.0, 2.0):
        ref = ko.mesh(vol, min_weight)
        gpu = p.extract_mesh(min_weight)
        print(f"{store} min_weight={min_weight}: {len(gpu.positions)} vertices, {len(gpu.triangles)} triangles")
        assert len(ref[3]) > 1000
        compare_mesh(gpu, ref)
    p.close()


def test_mesh_sphere_and_empty(K, ko, torch_gpu):
    import test_oracle_known_answers as T

    vol, _, _ = T.sphere_volume(ko)
    cfg = K.PipelineConfig(volume=K.VolumeGeometry((T.SYN_RES,) * 3, T.SYN_VOXEL, T.SYN_ORIGIN, T.SYN_TRUNC))
    p = K.Pipeline(cfg)
    assert len(p.extract_mesh().triangles) == 0  # unobserved volume
    p.upload_volume(vol.voxels, vol.colors)
    compare_mesh(p.extract_mesh(), ko.mesh(vol))
    p.close()


def test_small_workload_runs_and_is_deterministic(K, scene, torch_gpu):
    """The 5-frame 128^3 / 1 cm workload (its volume reaches z = 0.78 m only):
    every frame runs, and two runs give byte-identical poses and volumes."""
    outs = []
    for _ in range(2):
        p = K.Pipeline(pipeline_config(K))
        for i in range(5):
            depth, rgba, _ = scene.frame(i)
            p.process_frame(to_dev(torch_gpu, depth), to_dev(torch_gpu, rgba))
        outs.append((p.trajectory(), p.host_view()[0]))
        p.close()
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("radius", [3, 0])
def test_captured_frames_match_eager_frames(K, scene, torch_gpu, radius):
    """Tracking frames replay one captured HIP graph; with stage timing on,
    every frame is enqueued eagerly instead.  Both give byte-identical poses,
    voxels and colours, also across an ICP-option change (which re-captures)
    and a frame without colour, with the frames enqueued back to back, and
    the pyramid views after the last frame are that frame's in both.  With
    the bilateral filter on, its launch also stages the captured frame's
    inputs; without it (radius 0) a copy launch does."""
    def run(stage_timing):
        p = K.Pipeline(pipeline_config(K, voxel=ROOM_VOXEL, radius=radius, stage_timing=stage_timing))
        loose = K.ProjectiveIcpOptions(max_point_distance=0.2)
        frames = []
        for i in range(9):
            depth, rgba, _ = scene.frame(i)
            frames.append((to_dev(torch_gpu, depth), None if i == 4 else to_dev(torch_gpu, rgba)))
        torch_gpu.cuda.synchronize()
        for i, (depth, color) in enumerate(frames):
            p.process_frame(depth, color, icp=loose if i == 3 else None, report=False, sync_inputs=False)
        p.synchronize()
        vox, col = p.host_view()
        src = K.PyramidSource(p)
        levels = [{k: v.cpu().numpy().copy() for k, v in src.level(l).items() if k != "intrinsics"}
                  for l in range(3)]
        out = (p.trajectory(), vox.copy(), col.copy(), levels)
        p.close()
        return out

    eager, graph = run(True), run(False)
    assert np.array_equal(eager[0], graph[0])
    assert np.array_equal(eager[1], graph[1]) and np.array_equal(eager[2], graph[2])
    for a, b in zip(eager[3], graph[3]):
        for k in a:
            assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k


@pytest.mark.parametrize("res", [256, 200])
def test_skip_map_is_the_unobserved_neighbourhood(K, scene, torch_gpu, res):
    """The skip bit of a block is set exactly when no block among the 27
    around it (inside the grid) is observed, and a free block has all 27
    observed: at 256^3 a row of blocks fills whole bit words (the word-
    parallel map kernel), at 200^3 it does not (the per-block kernel)."""
    from scipy.ndimage import binary_dilation, binary_erosion

    p = K.Pipeline(pipeline_config(K, res=res, store="bf16", voxel=2.56 / res))
    for i in range(4):
        depth, rgba, _ = scene.frame(i)
        p.process_frame(to_dev(torch_gpu, depth), to_dev(torch_gpu, rgba))
    maps = p.block_bitmaps()
    p.close()
    nb = (res + 7) // 8
    cube = np.ones((3, 3, 3), bool)
    obs = maps["observed"].reshape(nb, nb, nb)
    skip = maps["skip"].reshape(nb, nb, nb)
    free = maps["free"].reshape(nb, nb, nb)
    assert 0 < obs.sum() < obs.size
    assert np.array_equal(skip, ~binary_dilation(obs, structure=cube))
    assert not np.any(free & ~binary_erosion(obs, structure=cube, border_value=0))
    print(f"{res}^3: observed {obs.sum()}, skip {skip.sum()}, free {free.sum()}")


@pytest.mark.parametrize("store", ["float", "bf16"])
def test_free_space_blocks_hold_only_ones(K, scene, torch_gpu, store):
    """A block the raycaster steps through as free space must hold, with its
    26 neighbours, only observed voxels reading exactly 1; and the marking
    finds most such blocks (256^3 of 1 cm, after 8 tracked frames)."""
    from scipy.ndimage import binary_erosion

    res = 256
    p = K.Pipeline(pipeline_config(K, res=res, store=store, voxel=0.01))
    for i in range(8):
        depth, rgba, _ = scene.frame(i)
        p.process_frame(to_dev(torch_gpu, depth), to_dev(torch_gpu, rgba))
    maps = p.block_bitmaps()
    vox, _ = p.host_view()
    nb = res // 8
    if store == "float":
        t, w = vox[:, 0], vox[:, 1]
    else:
        t = (vox[:, 0].astype(np.uint32) << 16).view(np.float32)
        w = vox[:, 1].astype(np.float32)
    ones = ((t == 1.0) & (w > 0)).reshape(res, res, res)  # z, y, x
    blocks_ones = ones.reshape(nb, 8, nb, 8, nb, 8).all(axis=(1, 3, 5))  # bz, by, bx
    possible = binary_erosion(blocks_ones, structure=np.ones((3, 3, 3), bool), border_value=0)
    free = maps["free"].reshape(nb, nb, nb)
    print(f"{store}: observed blocks {maps['observed'].sum()}, all-ones blocks {blocks_ones.sum()}, "
          f"free-able {possible.sum()}, free {free.sum()}")
    assert not np.any(free & ~possible)  # every free block really is free space
    assert possible.sum() > 0 and free.sum() >= 0.5 * possible.sum()
    p.close()

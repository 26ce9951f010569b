"""Pipeline options beyond the default path, against the CPU oracle:

* IntegrationMode::kBand on dense storage (tsdf_integration.hpp:29,444-497):
  the blocks a one-voxel band walk marks are integrated, every other voxel
  keeps its value; volumes and fused counts bit-identical to the oracle's
  BandIntegrationSweep for all three voxel stores;
* RaycastOptions::seed_from_previous (raycasting.hpp:122-131): rays start
  two truncation bands in front of the previous raycast's point at the same
  pixel; maps bit-identical to the oracle given the same previous maps;
* the learned fallback's bounded grid-wide wait: with its bound forced
  to zero from the host, a frame that times out is reported
  (KF_ERR_DEVICE), marked untracked and not integrated, and a reset context
  afterwards reproduces a fresh context's trajectory bit for bit (the
  ticket a timed-out launch leaves short is reset).
"""
import numpy as np
import pytest

from kf_helpers import INTR, STORE_KIND, H, W, config1_geometry, first_mismatch, host, pipeline_config, same_bits, \
    to_dev

pytestmark = pytest.mark.gpu

VOXEL = 0.02


@pytest.fixture(scope="module")
def K():
    import kinectfusion_amd

    return kinectfusion_amd


@pytest.fixture(scope="module")
def scene(K):
    return K.SyntheticScene("room", seed=0)


def oracle_dense(ko, store, color=True):
    r, v, origin, t = config1_geometry(128, VOXEL, 0.5, 0.05)
    return ko.Volume.create(r, v, origin, t, STORE_KIND[store], color)


@pytest.mark.parametrize("store", ["float", "quantized", "bf16"])
def test_band_integration_bit_exact(K, ko, scene, torch_gpu, store):
    cfg = pipeline_config(K, store=store, voxel=VOXEL)
    cfg.integration.mode = "band"
    p = K.Pipeline(cfg)
    vol = oracle_dense(ko, store)
    full = oracle_dense(ko, store)
    for i in range(3):
        depth, rgba, pose = scene.frame(i)
        w2c = ko.rigid_inverse(pose)
        ref = ko.integrate_band(vol, 1, depth, INTR, w2c, ko.integration_options(1000.0), color=rgba)
        ko.integrate(full, 1, depth, INTR, w2c, ko.integration_options(1000.0), color=rgba)
        got = p.integrate(to_dev(torch_gpu, depth), w2c, color=to_dev(torch_gpu, rgba), count=True)
        print(f"frame {i}: fused {got} vs oracle {ref[:2]}, {ref[2]} blocks listed")
        assert got == ref[:2] and ref[2] > 0
    vox, col = p.host_view()
    assert np.array_equal(vox.view(np.uint32), vol.voxels.view(np.uint32)), first_mismatch(vox, vol.voxels)
    assert same_bits(col, vol.colors)
    # the band is lossy (Niessner semantics): no free-space carving outside it
    assert not np.array_equal(vol.voxels, full.voxels)
    tb, wb = vol.decoded()
    tf, wf = full.decoded()
    shared = (wb > 0) & (wf > 0) & (np.abs(tf) < 1.0)
    assert shared.sum() > 0
    p.close()


def test_band_pipeline_tracks(K, scene, torch_gpu):
    cfg = pipeline_config(K, store="bf16", voxel=VOXEL)
    cfg.integration.mode = "band"
    p = K.Pipeline(cfg)
    for i in range(5):
        depth, rgba, _ = scene.frame(i)
        res = p.process_frame(to_dev(torch_gpu, depth), to_dev(torch_gpu, rgba))
        assert i == 0 or res.tracked == 1
    traj = p.trajectory()
    for i in range(5):
        assert np.linalg.norm(traj[i][:3, 3] - scene.pose(i)[:3, 3]) < 0.01
    p.close()


@pytest.mark.parametrize("store", ["float", "bf16"])
def test_seed_from_previous_raycast_bit_exact(K, ko, scene, torch_gpu, store):
    cfg = pipeline_config(K, store=store, voxel=VOXEL)
    cfg.raycast.seed_from_previous = True
    p = K.Pipeline(cfg)
    vol = oracle_dense(ko, store)
    for i in range(3):
        depth, rgba, pose = scene.frame(i)
        w2c = ko.rigid_inverse(pose)
        ko.integrate(vol, 1, depth, INTR, w2c, ko.integration_options(1000.0), color=rgba)
        p.integrate(to_dev(torch_gpu, depth), w2c, color=to_dev(torch_gpu, rgba))
    for level in range(3):
        s = 1 << level
        k = tuple(float(np.float32(v) * np.float32(1.0 / s)) for v in INTR)
        # the first raycast into the fresh (zero) maps starts at min_depth
        first = p.raycast(scene.pose(2), level)
        pts0, _, _, n_plain = ko.raycast(vol, scene.pose(2), k, W // s, H // s)
        assert same_bits(host(first.points)[..., :3], pts0)
        # the second, from the next pose, starts in front of those points
        maps = p.raycast(scene.pose(3), level)
        pts, nrm, col, n_seeded = ko.raycast(vol, scene.pose(3), k, W // s, H // s, previous_points=pts0)
        _, _, _, n_unseeded = ko.raycast(vol, scene.pose(3), k, W // s, H // s)
        print(f"level {level}: samples seeded {n_seeded} vs from min_depth {n_unseeded}")
        assert n_seeded < n_unseeded
        assert same_bits(host(maps.points)[..., :3], pts), first_mismatch(host(maps.points)[..., :3], pts)
        assert same_bits(host(maps.normals)[..., :3], nrm)
        assert np.array_equal(host(maps.colors).view(np.uint32), col)
    p.close()


def test_fallback_wait_timeout_is_reported_and_recovered(K, torch_gpu):
    from kinectfusion_amd import posenet as PN

    net = PN.PoseNet(PN.PoseNetConfig(seed=3))
    scene = K.SyntheticScene("room", seed=0)
    frames = [scene.frame(i * 2) for i in range(4)]
    cfg = pipeline_config(K, voxel=VOXEL, init="learned", learned_iterations=(4, 0, 0), learned_max_rms=1e-12)

    def run(p):
        out = []
        for depth, rgba, _ in frames:
            res = p.process_frame(to_dev(torch_gpu, depth), to_dev(torch_gpu, rgba))
            out.append(res.tracked)
        return out, p.trajectory(), p.host_view()[0]

    ref = K.Pipeline(cfg)
    ref.attach_posenet(net)
    want = run(ref)
    ref.close()

    p = K.Pipeline(cfg)
    p.attach_posenet(net)
    d0, c0, _ = frames[0]
    p.process_frame(to_dev(torch_gpu, d0), to_dev(torch_gpu, c0))
    before = p.host_view()[0]
    p.set_sync_timeout(0)  # every waiting workgroup gives up at once
    d1, c1, _ = frames[1]
    with pytest.raises(K.KinectFusionError) as err:
        p.process_frame(to_dev(torch_gpu, d1), to_dev(torch_gpu, c1))
    print(f"reported: {err.value}")
    assert err.value.status == 3 and "timed out" in str(err.value)
    assert np.array_equal(p.host_view()[0], before)  # the frame was not integrated
    p.synchronize()  # reported once: nothing pending
    p.set_sync_timeout(2_000_000)
    p.reset()
    got = run(p)
    assert got[0] == want[0] and all(t == 1 for t in got[0][1:])
    assert np.array_equal(got[1], want[1]) and np.array_equal(got[2], want[2])
    p.close()
    net.close()

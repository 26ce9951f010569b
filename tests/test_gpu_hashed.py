"""Hashed (sparse) volume against the oracle's sparse path (BlockRep,
block_rep.hpp:183-307; TruncationBandWalk, tsdf_integration.hpp:378-428).

Bars: the allocated block set, every voxel of the dense expansion, the fused
counts, the raycast maps and the mesh (traversal order, topology, welded
vertex numbering, positions, normals, colours) bit-identical to the oracle
on the same frames; with an exhausted pool, every allocated block equal to
the oracle's values for that block and nothing else written.
"""
import numpy as np
import pytest

from kf_helpers import INTR, STORE_KIND, H, W, config1_geometry, first_mismatch, host, pipeline_config, same_bits, \
    to_dev

pytestmark = pytest.mark.gpu

VOXEL = 0.02  # 128^3 of 2 cm hold the synthetic room


@pytest.fixture(scope="module")
def K():
    import kinectfusion_amd

    return kinectfusion_amd


@pytest.fixture(scope="module")
def scene(K):
    return K.SyntheticScene("room", seed=0)


def hashed_config(K, store="float", capacity=8192, buckets=1 << 16, **kw):
    return pipeline_config(K, store=store, voxel=VOXEL, storage="hashed", hash_buckets=buckets,
                           block_capacity=capacity, **kw)


def oracle_sparse(ko, store, capacity=8192, color=True):
    r, v, origin, t = config1_geometry(128, VOXEL, 0.5, 0.05)
    return ko.SparseVolume.create_sparse(r, v, origin, t, STORE_KIND[store], color, capacity=capacity)


def integrate_both(ko, p, vol, scene, torch_gpu, n, sparse=True):
    counts = []
    for i in range(n):
        depth, rgba, pose = scene.frame(i)
        w2c = ko.rigid_inverse(pose)
        ref = ko.integrate_sparse(vol, 1, depth, INTR, w2c, ko.integration_options(1000.0), color=rgba)
        got = p.integrate(to_dev(torch_gpu, depth), w2c, color=to_dev(torch_gpu, rgba), count=True)
        counts.append((ref, got))
    return counts


@pytest.mark.parametrize("store", ["float", "quantized", "bf16"])
def test_hashed_integration_bit_exact(K, ko, scene, torch_gpu, store):
    p = K.Pipeline(hashed_config(K, store))
    vol = oracle_sparse(ko, store)
    for ref, got in integrate_both(ko, p, vol, scene, torch_gpu, 3):
        print(f"fused {got} vs oracle {ref}")
        assert got == ref[:2]
    keys = p.hashed_blocks()
    assert np.array_equal(keys.astype(np.int64), np.nonzero(vol.slots >= 0)[0])
    st = p.hashed_stats()
    assert st["allocated"] == int(vol.state[0]) and st["overflow"] == 0
    vox, col = p.host_view()
    ex = vol.dense()
    assert np.array_equal(vox.view(np.uint32), ex.voxels.view(np.uint32)), first_mismatch(vox, ex.voxels)
    assert same_bits(col, ex.colors)
    assert p.observed_voxel_count() == ko.observed_count(vol)
    gk, gv, gc = p.hashed_download()
    rk, rv, rc = vol.blocks()
    assert np.array_equal(gk, rk) and same_bits(gv, rv) and same_bits(gc, rc)
    p.close()


def test_hashed_raycast_and_mesh_bit_exact(K, ko, scene, torch_gpu):
    p = K.Pipeline(hashed_config(K, "bf16"))
    vol = oracle_sparse(ko, "bf16")
    integrate_both(ko, p, vol, scene, torch_gpu, 3)
    pose = scene.pose(3)
    for level in range(3):
        s = 1 << level
        k = tuple(float(np.float32(v) * np.float32(1.0 / s)) for v in INTR)
        maps = p.raycast(pose, level)
        pts, nrm, col, _ = ko.raycast(vol, pose, k, W // s, H // s)
        gp, gn = host(maps.points)[..., :3], host(maps.normals)[..., :3]
        assert np.isfinite(pts).all(-1).mean() > 0.5
        assert same_bits(gp, pts), first_mismatch(gp, pts)
        assert same_bits(gn, nrm), first_mismatch(gn, nrm)
        assert np.array_equal(host(maps.colors).view(np.uint32), col)
    for min_weight in (0.0, 2.0):
        rpos, rnrm, rcol, rtri = ko.mesh(vol, min_weight)
        m = p.extract_mesh(min_weight)
        print(f"min_weight {min_weight}: {len(m.positions)} vertices, {len(m.triangles)} triangles")
        assert len(rtri) > 1000
        assert np.array_equal(m.triangles, rtri)
        assert same_bits(m.positions, rpos) and same_bits(m.normals, rnrm) and same_bits(m.colors, rcol)
    p.close()


def test_exhausted_pool_writes_only_allocated_blocks(K, ko, scene, torch_gpu):
    """capacity 64 (BlockRep overflow, tests.cpp:747-770): dropped blocks stay
    fresh; each allocated block holds exactly the oracle's values for it."""
    p = K.Pipeline(hashed_config(K, "float", capacity=64, buckets=1 << 12))
    full = oracle_sparse(ko, "float", capacity=1 << 20)
    depth, rgba, pose = scene.frame(0)
    w2c = ko.rigid_inverse(pose)
    ko.integrate_sparse(full, 1, depth, INTR, w2c, ko.integration_options(1000.0), color=rgba)
    p.integrate(to_dev(torch_gpu, depth), w2c, color=to_dev(torch_gpu, rgba))
    st = p.hashed_stats()
    assert st["allocated"] == 64 and st["overflow"] > 0
    keys = p.hashed_blocks()
    assert np.all(full.slots[keys] >= 0)  # only blocks the band walk visits
    vox, _ = p.host_view()
    g = full.geometry
    bx, by = (g.nx + 7) // 8, (g.ny + 7) // 8
    idx = np.arange(g.nx * g.ny * g.nz)
    x, y, z = idx % g.nx, (idx // g.nx) % g.ny, idx // (g.nx * g.ny)
    block = ((z // 8) * by + (y // 8)) * bx + (x // 8)
    inside = np.isin(block, keys)
    assert np.array_equal(vox[inside], full.dense().voxels[inside])
    assert np.all(vox[~inside, 1] == 0.0) and np.all(vox[~inside, 0] == 1.0)
    p.close()


def test_hashed_frames_track(K, ko, scene, torch_gpu):
    """Five frames through process_frame on the hashed store: every frame
    tracks, within 1 cm of the ground truth, and runs are deterministic."""
    outs = []
    for _ in range(2):
        p = K.Pipeline(hashed_config(K, "bf16"))
        for i in range(5):
            depth, rgba, _ = scene.frame(i)
            res = p.process_frame(to_dev(torch_gpu, depth), to_dev(torch_gpu, rgba))
            if i > 0:
                assert res.tracked == 1
        traj = p.trajectory()
        for i in range(5):
            assert np.linalg.norm(traj[i][:3, 3] - scene.pose(i)[:3, 3]) < 0.01
        outs.append((traj, p.host_view()[0]))
        p.close()
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])

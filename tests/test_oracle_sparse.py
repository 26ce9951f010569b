"""The oracle's sparse path pinned by the reference's own sparse tests and by
what the mathematics fixes: on its allocated blocks a single frame's sparse
integration equals the dense sweep (tests.cpp:722-745 compares sparse and
dense band); an exhausted pool degrades to missing blocks, never wrong
voxels (tests.cpp:747-770); the sparse mesh is the dense mesh of the same
voxels in another traversal order (marching_cubes.cpp:490-503); every band
sample of a fronto-parallel plane lies within the band of the plane."""
import numpy as np

from kf_helpers import INTR, config1_geometry

VOXEL = 0.02


def plane_depth(d_mm=600, w=160, h=120):
    return np.full((h, w), d_mm, np.uint16)


def small_intr(w=160, h=120):
    return (131.25, 131.25, (w - 1) / 2.0, (h - 1) / 2.0)


def sparse_and_dense(ko, capacity=None):
    r, v, origin, t = config1_geometry(64, VOXEL, 0.5, 0.05)
    sp = ko.SparseVolume.create_sparse(r, v, origin, t, ko.FLOAT, True, capacity=capacity)
    dn = ko.Volume.create(r, v, origin, t, ko.FLOAT, True)
    return sp, dn


def blocks_of(vol):
    g = vol.geometry
    bx, by = (g.nx + 7) // 8, (g.ny + 7) // 8
    idx = np.arange(g.nx * g.ny * g.nz)
    x, y, z = idx % g.nx, (idx // g.nx) % g.ny, idx // (g.nx * g.ny)
    return ((z // 8) * by + (y // 8)) * bx + (x // 8)


def test_sparse_equals_dense_on_allocated_blocks(ko):
    sp, dn = sparse_and_dense(ko)
    depth = plane_depth()
    opts = ko.integration_options(1000.0)
    pose = np.eye(4)
    ko.integrate(dn, 0, depth, small_intr(), pose, opts)
    fused, _, listed = ko.integrate_sparse(sp, 0, depth, small_intr(), pose, opts)
    assert listed > 0 and fused > 0
    alloc = np.isin(blocks_of(sp), np.nonzero(sp.slots >= 0)[0])
    ex = sp.dense()
    assert np.array_equal(ex.voxels[alloc], dn.voxels[alloc])
    assert np.all(ex.voxels[~alloc, 1] == 0.0)
    # the band (|sd| < truncation) lies in the allocated blocks, up to the
    # block corners that half-block steps can pass between
    t, w = dn.decoded()
    band = (w > 0) & (np.abs(t) < 1.0)
    assert band.sum() > 0 and alloc[band].mean() > 0.99


def test_exhausted_pool_drops_blocks_in_walk_order(ko):
    sp, dn = sparse_and_dense(ko, capacity=1)
    depth = plane_depth()
    opts = ko.integration_options(1000.0)
    ko.integrate(dn, 0, depth, small_intr(), np.eye(4), opts)
    ko.integrate_sparse(sp, 0, depth, small_intr(), np.eye(4), opts)
    visits = ko.band_walk(sp.geometry, depth, small_intr(), np.eye(4), opts)
    assert np.count_nonzero(sp.slots >= 0) == 1 and sp.slots[visits[0]] == 0
    assert sp.state[0] > sp.state[1]  # overflow recorded
    alloc = blocks_of(sp) == visits[0]
    ex = sp.dense()
    assert np.array_equal(ex.voxels[alloc], dn.voxels[alloc])
    assert np.all(ex.voxels[~alloc, 1] == 0.0)


def test_band_samples_stay_within_the_band_of_a_plane(ko):
    r, v, origin, t = config1_geometry(64, VOXEL, 0.5, 0.05)
    sp = ko.SparseVolume.create_sparse(r, v, origin, t, ko.FLOAT, False)
    depth = plane_depth(600)
    visits = np.unique(ko.band_walk(sp.geometry, depth, small_intr(), np.eye(4), ko.integration_options(1000.0)))
    g = sp.geometry
    bx, by = (g.nx + 7) // 8, (g.ny + 7) // 8
    bz = visits // (bx * by)
    z_lo = g.oz + bz * 8 * g.voxel_size
    z_hi = z_lo + 8 * g.voxel_size
    # samples span depths [d - band, d + band + step]
    assert len(visits) > 0
    assert np.all(z_hi >= 0.6 - 0.05 - 1e-4) and np.all(z_lo <= 0.6 + 0.05 + 0.08 + 1e-4)


def test_sparse_mesh_is_the_dense_mesh_reordered(ko, scene_room):
    r, v, origin, t = config1_geometry(128, VOXEL, 0.5, 0.05)
    vol = ko.SparseVolume.create_sparse(r, v, origin, t, ko.FLOAT, True)
    for i in range(2):
        depth, rgba, pose = scene_room.frame(i)
        ko.integrate_sparse(vol, 1, depth, INTR, ko.rigid_inverse(pose), ko.integration_options(1000.0), color=rgba)
    dense = vol.dense()

    def tri_set(m):
        pos, _, _, tri = m
        return sorted(map(tuple, pos[tri].reshape(-1, 9).tolist()))

    ms, md = ko.mesh(vol), ko.mesh(dense)
    assert len(ms[3]) == len(md[3]) > 1000
    assert not np.array_equal(ms[3], md[3])  # the order differs
    assert tri_set(ms) == tri_set(md)

"""The oracle's interpolated geometry pinned on TSDF fields set directly:
a linear field is reproduced exactly by trilinear blending and by the
marching-cubes edge interpolation, so on a tilted plane

* the raycast's zero crossing (raycasting.hpp:188-193, the linear
  interpolation between the last positive and first non-positive sample)
  lies on the plane to float rounding -- without the interpolation it would
  be off by up to a step;
* the central-difference normal (trilinear.hpp:246-269) and the cell
  gradient (trilinear.hpp:214-243) are the field's gradient direction -n --
  a wrong sign, axis or stencil fails it;
* marching-cubes vertices (marching_cubes.cpp:437-441, -v_near / (v_far -
  v_near) along the edge between cell centres) lie on the plane -- midpoints
  would not;
* uniform colour voxels give the truncated RGBA8 of that colour
  (pixel_from_color, rgbd.hpp:84-93), and without colour the headlight
  shade is 255 max(0, n . view) of the geometry (raycasting.hpp:233-243)."""
import numpy as np
import pytest

RES, VOXEL, ORIGIN, TRUNC = 48, 0.02, (-0.48, -0.48, 0.3), 0.1
INTR, W, H = (40.0, 40.0, 15.5, 15.5), 32, 32
COLOR = (10.7, 200.2, 33.9)


def plane():
    n = np.array([0.21, -0.13, 1.0])
    n /= np.linalg.norm(n)
    return n, 0.92  # points p with n . p = d0, facing the camera at the origin


def linear_volume(ko, color=True, cut=1.0):
    """Float voxels: tsdf = clamp((d0 - n . c) / trunc) at every voxel
    centre c, weight 1."""
    vol = ko.Volume.create(RES, VOXEL, ORIGIN, TRUNC, ko.FLOAT, color)
    n, d0 = plane()
    i = np.arange(RES ** 3)
    x, y, z = i % RES, (i // RES) % RES, i // (RES * RES)
    c = np.stack([ORIGIN[0] + (x + 0.5) * VOXEL, ORIGIN[1] + (y + 0.5) * VOXEL, ORIGIN[2] + (z + 0.5) * VOXEL], 1)
    t = np.clip((d0 - c @ n) / TRUNC, -cut, cut).astype(np.float32)
    vol.voxels[:, 0] = t
    vol.voxels[:, 1] = 1.0
    if color:
        vol.colors[:, :3] = np.array(COLOR, np.float32)
        vol.colors[:, 3] = 1.0
    return vol


@pytest.fixture(scope="module")
def rays(ko):
    vol = linear_volume(ko)
    return vol, ko.raycast(vol, np.eye(4), INTR, W, H)


def test_raycast_hits_lie_on_the_plane(rays):
    n, d0 = plane()
    _, (pts, nrm, col, _) = rays
    valid = np.isfinite(pts).all(-1)
    assert valid.mean() > 0.9
    off = np.abs(pts[valid].astype(np.float64) @ n - d0)
    print(f"max distance from the plane {off.max():.2e} m")
    assert off.max() < 1e-5  # a voxel is 2e-2; no interpolation would leave ~1e-2


def test_raycast_normals_are_the_gradient_direction(ko, rays):
    n, _ = plane()
    vol, (pts, nrm, _, _) = rays
    valid = np.isfinite(pts).all(-1)
    err = np.abs(nrm[valid].astype(np.float64) + n).max()
    print(f"max normal error {err:.2e}")
    assert err < 1e-6
    _, cnrm, _, _ = ko.raycast(vol, np.eye(4), INTR, W, H, cell_normals=True)
    assert np.abs(cnrm[valid].astype(np.float64) + n).max() < 1e-6


def test_raycast_colour_of_uniform_voxels(rays):
    _, (pts, _, col, _) = rays
    valid = np.isfinite(pts).all(-1)
    want = 10 | (200 << 8) | (33 << 16) | (0xFF << 24)
    assert (col[valid] == want).all() and (col[~valid] == 0).all()


def test_headlight_shading_without_colour(ko):
    n, _ = plane()
    vol = linear_volume(ko, color=False)
    pts, nrm, col, _ = ko.raycast(vol, np.eye(4), INTR, W, H)
    valid = np.isfinite(pts).all(-1)
    p = pts[valid].astype(np.float64)
    view = -p / np.linalg.norm(p, axis=1, keepdims=True)
    shade = 255.0 * np.maximum(0.0, view @ -n)
    got = col[valid]
    r, g, b, a = got & 0xFF, (got >> 8) & 0xFF, (got >> 16) & 0xFF, got >> 24
    assert (a == 0xFF).all() and (r == g).all() and (g == b).all()
    assert np.abs(r - np.floor(shade)).max() <= 1 and (r == np.floor(shade)).mean() > 0.95


def test_mesh_vertices_lie_on_the_plane(ko):
    n, d0 = plane()
    vol = linear_volume(ko)
    pos, nrm, col, tri = ko.mesh(vol)
    assert len(tri) > 1000
    off = np.abs(pos.astype(np.float64) @ n - d0)
    print(f"mesh: {len(pos)} vertices, max distance from the plane {off.max():.2e} m")
    assert off.max() < 1e-6
    # central-difference normals where all six samples are inside the grid;
    # at its faces the face-average fallback (marching_cubes.cpp:452-467)
    lo = np.array(ORIGIN) + 2.5 * VOXEL
    hi = np.array(ORIGIN) + (RES - 2.5) * VOXEL
    inner = ((pos > lo) & (pos < hi)).all(1)
    assert inner.mean() > 0.8
    assert np.abs(nrm[inner].astype(np.float64) + n).max() < 1e-5
    assert (nrm[~inner].astype(np.float64) @ -n > 0.99).all()
    assert np.allclose(col, np.array(COLOR, np.float32), atol=1e-4)
    # winding: faces point along the field's gradient direction -n (outward from the solid side)
    face = np.cross(pos[tri[:, 1]] - pos[tri[:, 0]], pos[tri[:, 2]] - pos[tri[:, 0]])
    assert ((face @ -n) > 0).all()


def test_clamped_field_band_edges(ko):
    """With the field clamped at +-0.5 (a band of half the truncation) the
    crossing still lies on the plane: the samples that bracket it are inside
    the linear band."""
    n, d0 = plane()
    vol = linear_volume(ko, cut=0.5)
    pts, _, _, _ = ko.raycast(vol, np.eye(4), INTR, W, H)
    valid = np.isfinite(pts).all(-1)
    assert valid.mean() > 0.9
    assert np.abs(pts[valid].astype(np.float64) @ n - d0).max() < 1e-5

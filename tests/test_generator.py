"""The seeded synthetic RGB-D generator (no reference counterpart) pinned by
its analytic ground truth and the noise model it states
(sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2 m, 1 % dropped pixels, millimetre
depth; per-frame steps <= 1 cm and <= 0.5 deg; 6 spheres of radius
U(0.05, 0.2) m in a 1.2 m wide room).

The analytic depth here is computed independently of the renderer: every
pixel ray is intersected in numpy with the room box, the spheres and the
boxes read from the scene description, the nearest hit point is moved into
the camera frame, and its z is the depth (the convention of
CameraIntrinsics::back_project, rgbd.hpp:53-57)."""
import ctypes as C

import numpy as np
import pytest


class D3(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class Sphere(C.Structure):
    _fields_ = [("c", D3), ("r", C.c_double), ("rgba", C.c_uint32)]


class Box(C.Structure):
    _fields_ = [("lo", D3), ("hi", D3), ("rgba", C.c_uint32)]


class SceneDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("seed", C.c_uint64), ("width", C.c_int), ("height", C.c_int),
                ("fx", C.c_double), ("fy", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
                ("room_lo", D3), ("room_hi", D3), ("wall_rgba", C.c_uint32 * 6), ("n_spheres", C.c_int),
                ("n_boxes", C.c_int), ("spheres", Sphere * 64), ("boxes", Box * 64),
                ("t_amp", C.c_double * 3), ("t_freq", C.c_double * 3), ("t_phase", C.c_double * 3),
                ("r_amp", C.c_double * 3), ("r_freq", C.c_double * 3), ("r_phase", C.c_double * 3),
                ("pan_amp", C.c_double), ("walk_amp", C.c_double), ("sweep_period", C.c_double)]


def description(scene) -> SceneDesc:
    size = C.c_size_t()
    ptr = scene._lib.kf_scene_description(scene._h, C.byref(size))
    assert size.value == C.sizeof(SceneDesc)
    return SceneDesc.from_buffer_copy(C.string_at(ptr, size.value))


def v3(d):
    return np.array([d.x, d.y, d.z])


def analytic(desc: SceneDesc, pose: np.ndarray, w: int, h: int):
    """(camera-frame z of the nearest hit, colour) per pixel."""
    u, v = np.meshgrid(np.arange(w, dtype=np.float64), np.arange(h, dtype=np.float64))
    dc = np.stack([(u - desc.cx) / desc.fx, (v - desc.cy) / desc.fy, np.ones_like(u)], -1)
    R, o = pose[:3, :3], pose[:3, 3]
    d = dc @ R.T
    best = np.full((h, w), np.inf)
    col = np.zeros((h, w), np.uint32)
    # room: exit through the nearest wall (walls 2a / 2a + 1 = low / high side of axis a)
    lo, hi = v3(desc.room_lo), v3(desc.room_hi)
    for a in range(3):
        with np.errstate(divide="ignore", invalid="ignore"):
            s_hi = np.where(d[..., a] > 1e-15, (hi[a] - o[a]) / d[..., a], np.inf)
            s_lo = np.where(d[..., a] < -1e-15, (lo[a] - o[a]) / d[..., a], np.inf)
        for s, wall in ((s_hi, 2 * a + 1), (s_lo, 2 * a)):
            take = (s > 1e-9) & (s < best)
            best = np.where(take, s, best)
            col = np.where(take, desc.wall_rgba[wall], col)
    for i in range(desc.n_spheres):
        sp = desc.spheres[i]
        oc = o - v3(sp.c)
        a2 = (d * d).sum(-1)
        b = (d * oc).sum(-1)
        c = oc @ oc - sp.r ** 2
        disc = b * b - a2 * c
        ok = disc >= 0
        q = np.sqrt(np.where(ok, disc, 0.0))
        s0 = (-b - q) / a2
        s0 = np.where(s0 <= 1e-9, (-b + q) / a2, s0)
        take = ok & (s0 > 1e-9) & (s0 < best)
        best = np.where(take, s0, best)
        col = np.where(take, sp.rgba, col)
    for i in range(desc.n_boxes):
        bx = desc.boxes[i]
        blo, bhi = v3(bx.lo), v3(bx.hi)
        tmin, tmax = np.full((h, w), -np.inf), np.full((h, w), np.inf)
        for a in range(3):
            with np.errstate(divide="ignore", invalid="ignore"):
                t0 = (blo[a] - o[a]) / d[..., a]
                t1 = (bhi[a] - o[a]) / d[..., a]
            tmin = np.maximum(tmin, np.minimum(t0, t1))
            tmax = np.minimum(tmax, np.maximum(t0, t1))
        take = (tmax >= tmin) & (tmin > 1e-9) & (tmin < best)
        best = np.where(take, tmin, best)
        col = np.where(take, bx.rgba, col)
    hit = o + best[..., None] * d
    z = ((hit - o) @ R)[..., 2]  # the hit point in the camera frame
    return np.where(np.isfinite(best), z, 0.0), col


@pytest.fixture(scope="module")
def K():
    import kinectfusion_amd

    return kinectfusion_amd


@pytest.mark.parametrize("kind,frames", [("room", (0, 17, 60)), ("hall", (0, 150, 299))])
def test_noise_free_depth_is_the_analytic_depth(K, kind, frames):
    scene = K.SyntheticScene(kind, seed=0, width=160, height=120, fx=131.25, fy=131.25, cx=79.5, cy=59.5)
    desc = description(scene)
    for f in frames:
        pose = scene.pose(f)
        depth, rgba, exact = scene.render(pose, noise=False, true_depth=True)
        z, col = analytic(desc, pose, 160, 120)
        assert (z > 0).all()
        assert np.abs(exact - z).max() <= 2e-7 * z.max()
        # millimetre quantisation of the exact depth
        assert np.array_equal(depth, np.floor(z * 1000.0 + 0.5).astype(np.uint16))
        assert (col == rgba).mean() > 0.999  # ties at silhouettes aside
        # hits lie on the surfaces: a ray through a sphere's centre pixel ends on it
        for i in range(desc.n_spheres):
            sp = desc.spheres[i]
            cam = np.linalg.inv(pose) @ np.append(v3(sp.c), 1.0)
            if cam[2] <= sp.r + 0.05:
                continue
            px = int(round(desc.fx * cam[0] / cam[2] + desc.cx)), int(round(desc.fy * cam[1] / cam[2] + desc.cy))
            if not (0 <= px[0] < 160 and 0 <= px[1] < 120) or rgba[px[1], px[0]] != sp.rgba:
                continue
            dc = np.array([(px[0] - desc.cx) / desc.fx, (px[1] - desc.cy) / desc.fy, 1.0])
            p = pose[:3, :3] @ (dc * exact[px[1], px[0]]) + pose[:3, 3]
            assert abs(np.linalg.norm(p - v3(sp.c)) - sp.r) < 1e-5


def test_room_layout(K):
    for seed in range(4):
        desc = description(K.SyntheticScene("room", seed=seed))
        assert desc.n_spheres == 6 and desc.n_boxes == 0
        assert abs((desc.room_hi.x - desc.room_lo.x) - 1.2) < 1e-12
        for i in range(6):
            sp = desc.spheres[i]
            assert 0.05 <= sp.r < 0.2
            c = v3(sp.c)
            # inside the room with a gap to the side walls
            assert desc.room_lo.x + sp.r < c[0] < desc.room_hi.x - sp.r
            assert desc.room_lo.y + sp.r < c[1] < desc.room_hi.y - sp.r
            # in front of the first camera, its surface 0.5-1.2 m away before the wall clamp
            assert c[2] > 0.4 and 0.4 < np.linalg.norm(c) - sp.r < 1.25
    hall = description(K.SyntheticScene("hall", seed=0))
    assert hall.n_spheres + hall.n_boxes == 40
    ext = v3(hall.room_hi) - v3(hall.room_lo)
    assert sorted(ext.tolist()) == [3.0, 4.0, 6.0]


def test_noise_model_and_dropout(K):
    """Residuals of the noisy depth against the exact depth, binned by z:
    their standard deviation is sigma(z) within 5 % (after the millimetre
    rounding's 1/12 mm^2), their mean is zero within the bin's standard error,
    and 1 % +- 0.2 % of the pixels are dropped."""
    scene = K.SyntheticScene("room", seed=0)
    res, zs = [], []
    dropped = total = 0
    for f in range(6):
        pose = scene.pose(f)
        noisy, _ = scene.render(pose, noise_key=f, noise=True)
        _, _, exact = scene.render(pose, noise=False, true_depth=True)
        valid = exact > 0
        total += int(valid.sum())
        dropped += int((valid & (noisy == 0)).sum())
        keep = valid & (noisy > 0)
        res.append(noisy[keep] / 1000.0 - exact[keep].astype(np.float64))
        zs.append(exact[keep].astype(np.float64))
    res, zs = np.concatenate(res), np.concatenate(zs)
    assert abs(dropped / total - 0.01) < 0.002
    edges = np.quantile(zs, np.linspace(0, 1, 9))
    for lo, hi in zip(edges[:-1], edges[1:]):
        m = (zs >= lo) & (zs < hi)
        if m.sum() < 20000:
            continue
        sigma = 0.0012 + 0.0019 * (zs[m] - 0.4) ** 2
        expect = np.sqrt(np.mean(sigma ** 2) + 1e-6 / 12.0)
        got = res[m].std()
        assert abs(got / expect - 1.0) < 0.05, (lo, hi, got, expect)
        assert abs(res[m].mean()) < 4 * got / np.sqrt(m.sum())


@pytest.mark.parametrize("kind,frames", [("room", 100), ("hall", 300)])
def test_trajectory_steps_are_bounded(K, kind, frames):
    scene = K.SyntheticScene(kind, seed=0)
    assert np.allclose(scene.pose(0), np.eye(4))
    for f in range(1, frames):
        a, b = scene.pose(f - 1), scene.pose(f)
        rel = np.linalg.inv(a) @ b
        ang = np.degrees(np.arccos(np.clip((np.trace(rel[:3, :3]) - 1) / 2, -1, 1)))
        assert np.linalg.norm(rel[:3, 3]) <= 0.01 and ang <= 0.5, (f, rel[:3, 3], ang)
        assert np.allclose(b[:3, :3] @ b[:3, :3].T, np.eye(3), atol=1e-12)

"""The CPU oracle against the reference's own known-answer tests
(test/tests.cpp) and against closed forms that a plausible mistake (dropped
term, wrong sign or index, transposed operand) would fail."""
import numpy as np
import pytest

# Fixtures of test/tests.cpp:42-83
DEPTH_1M, DEPTH_1P4M = 5000, 7000
SYN_RES, SYN_VOXEL, SYN_ORIGIN, SYN_TRUNC = 32, 0.05, (-0.8, -0.8, 0.2), 0.05
SYN_INTR = (20.0, 20.0, 7.5, 7.5)


def syn_volume(ko, kind=0, color=True):
    return ko.Volume.create(SYN_RES, SYN_VOXEL, SYN_ORIGIN, SYN_TRUNC, kind, color)


def flat(raw, n=16):
    return np.full((n, n), raw, np.uint16)


# --- depth processing -------------------------------------------------------
def test_back_projection_flat_image(ko):
    # tests.cpp:187-210
    depth = flat(DEPTH_1M, 3)
    v = ko.vertices(depth, (1.0, 1.0, 1.0, 1.0), 5000.0)
    assert np.allclose(v[1, 1], [0, 0, 1]) and np.allclose(v[1, 2], [1, 0, 1])
    n = ko.normals(v)
    assert np.allclose(n[1, 1], [0, 0, -1])
    assert np.isnan(n[0, 0]).all()  # border


def test_downsample_rejects_mixed_depth(ko):
    # tests.cpp:212-225
    d = np.array([[DEPTH_1M, DEPTH_1P4M], [DEPTH_1M, DEPTH_1M]], np.uint16)
    assert ko.downsample(d, 5000.0)[0, 0] == 0


def test_downsample_rounds_mean_and_ignores_zeros(ko):
    d = np.array([[1000, 1001], [0, 1002]], np.uint16)
    # (sum + count/2) / count over non-zero samples
    assert ko.downsample(d, 1000.0)[0, 0] == (1000 + 1001 + 1002 + 1) // 3
    assert ko.downsample(np.zeros((2, 2), np.uint16), 1000.0)[0, 0] == 0


def test_bilateral_closed_forms(ko):
    rng = np.random.default_rng(1)
    # constant image is a fixed point
    assert (ko.bilateral(flat(1200, 12), 1000.0, 0.4, 8.0, 3, 4.5, 0.03) == 1200).all()
    # a step far beyond 3 sigma_r passes through unchanged
    d = np.full((10, 10), 1000, np.uint16)
    d[:, 5:] = 2000
    assert (ko.bilateral(d, 1000.0, 0.4, 8.0, 3, 4.5, 0.03) == d).all()
    # zeros stay zero and are ignored by neighbours; out-of-range centres give 0
    d = rng.integers(900, 1100, (20, 24)).astype(np.uint16)
    d[5, 5] = 0
    d[7, 9] = 100  # 0.1 m < min depth
    out = ko.bilateral(d, 1000.0, 0.4, 8.0, 3, 4.5, 0.03)
    assert out[5, 5] == 0 and out[7, 9] == 0
    # brute force in float64 (clamped window) agrees to one raw unit
    ss, sr = -0.5 / 4.5 ** 2, -0.5 / 0.03 ** 2
    for (r, c) in [(0, 0), (10, 11), (19, 23), (6, 5)]:
        cm = d[r, c] / 1000.0
        num = den = 0.0
        for y in range(max(r - 3, 0), min(r + 3, 19) + 1):
            for x in range(max(c - 3, 0), min(c + 3, 23) + 1):
                m = d[y, x] / 1000.0
                if d[y, x] == 0 or m < 0.4 or m > 8.0:
                    continue
                wgt = np.exp(((x - c) ** 2 + (y - r) ** 2) * ss + (m - cm) ** 2 * sr)
                num += wgt * d[y, x]
                den += wgt
        assert abs(int(out[r, c]) - int(np.floor(num / den + 0.5))) <= 1


# --- voxel codecs (volume.hpp:28-100) ---------------------------------------
def test_codec_known_answers(ko):
    L = ko.lib()
    assert L.kfo_bf_encode(1.0) == 0x3F80
    assert L.kfo_bf_decode(0x3F80) == 1.0
    assert L.kfo_q_encode(1.0) == 32767 and L.kfo_q_encode(-1.0) == -32767
    assert L.kfo_q_encode(0.5) == 16383  # trunc(16383.5)
    assert L.kfo_q_encode(-0.5) == -16383
    assert L.kfo_q_encode(7.0) == 32767  # clamped
    assert L.kfo_w_encode(1.0) == 256 and L.kfo_w_encode(0.00390) == 0 and L.kfo_w_encode(1e9) == 65535
    # bf16 rounds ties to even: 0.5 (1 + 2^-8) -> 0x3F00, 0.5 (1 + 3 2^-8) -> 0x3F02
    assert L.kfo_bf_encode(np.float32(0.5 * (1.0 + 2 ** -8))) == 0x3F00
    assert L.kfo_bf_encode(np.float32(0.5 * (1.0 + 3 * 2 ** -8))) == 0x3F02
    assert L.kfo_bf_encode(np.float32(1.5)) == 0x3F80  # clamped to 1 first


# --- integration -------------------------------------------------------------
def test_flat_plane_classic_closed_form(ko):
    """One frame of a fronto-parallel plane at 1 m (tests.cpp:454-474 setup):
    every fused voxel holds clamp((d - |p|/lambda)/trunc) with weight 1."""
    vol = syn_volume(ko)
    fused, _ = ko.integrate(vol, ko.CLASSIC, flat(DEPTH_1M), SYN_INTR, np.eye(4))
    t, w = vol.decoded()
    assert fused > 0 and fused == int((w > 0).sum()) and set(np.unique(w)) <= {0.0, 1.0}
    n = SYN_RES
    idx = np.nonzero(w > 0)[0]
    x, y, z = idx % n, (idx // n) % n, idx // (n * n)
    p = np.stack([SYN_ORIGIN[0] + (x + 0.5) * SYN_VOXEL, SYN_ORIGIN[1] + (y + 0.5) * SYN_VOXEL,
                  SYN_ORIGIN[2] + (z + 0.5) * SYN_VOXEL], 1)
    u = np.floor(20.0 * p[:, 0] / p[:, 2] + 7.5 + 0.5)
    v = np.floor(20.0 * p[:, 1] / p[:, 2] + 7.5 + 0.5)
    lam = np.sqrt(((u - 7.5) / 20.0) ** 2 + ((v - 7.5) / 20.0) ** 2 + 1.0)
    expect = np.clip((1.0 - np.linalg.norm(p, axis=1) / lam) / SYN_TRUNC, -1, 1)
    assert np.abs(t[idx] - expect).max() < 1e-5
    # behind the band nothing is fused; in front, free space is carved to +1
    assert (expect >= -1.0).all() and (t[idx][expect > 1.0 - 1e-9] == 1.0).all()


def test_storage_tolerances(ko):
    # tests.cpp:667-696: int16 differs by < 1e-4, bf16 by < 5e-3, same voxels
    vols = {k: syn_volume(ko, k, color=False) for k in (ko.FLOAT, ko.QUANTIZED, ko.BF16)}
    for vol in vols.values():
        ko.integrate(vol, ko.ANGLE, flat(DEPTH_1M), SYN_INTR, np.eye(4))
    t0, w0 = vols[ko.FLOAT].decoded()
    for kind, tol in ((ko.QUANTIZED, 1e-4), (ko.BF16, 5e-3)):
        t, w = vols[kind].decoded()
        assert ((w > 0) == (w0 > 0)).all()
        d = np.abs(t - t0)[w0 > 0].max()
        assert 0.0 < d < tol


def test_rules_differ_and_repeat_exactly(ko):
    # tests.cpp:505-558
    normals = np.zeros((16, 16, 3), np.float32)
    normals[..., 2] = -1.0

    def run(rule):
        vol = syn_volume(ko)
        for raw in (DEPTH_1M, DEPTH_1P4M):
            ko.integrate(vol, rule, flat(raw), SYN_INTR, np.eye(4), normals=normals)
        return vol.decoded()

    (tc, wc), (ta, wa), (tr, wr) = run(ko.CLASSIC), run(ko.ANGLE), run(ko.ANGLE)
    both = (wc > 0) & (wa > 0)
    assert ((wc > 0) == (wa > 0)).all() and both.any()
    assert np.abs(wc - wa)[both].max() > 0 and np.abs(tc - ta)[both].max() > 0
    assert np.array_equal(ta, tr) and np.array_equal(wa, wr)


def test_angle_weight_is_cos_over_depth(ko):
    """Head-on normal at depth d: every first-frame weight is cos/d with
    cos = -n.r_hat of the voxel's pixel ray."""
    vol = syn_volume(ko, color=False)
    normals = np.zeros((16, 16, 3), np.float32)
    normals[..., 2] = -1.0
    ko.integrate(vol, ko.ANGLE, flat(DEPTH_1M), SYN_INTR, np.eye(4), normals=normals)
    _, w = vol.decoded()
    vals = np.unique(w[w > 0])
    # rays: cos = 1/lambda, d = 1 -> weights in (1/lambda_max, 1]
    lam_max = np.sqrt((7.5 / 20) ** 2 * 2 + 1)
    assert vals.max() <= 1.0 and vals.min() >= 1.0 / lam_max - 1e-6


# --- raycast -------------------------------------------------------------------
def test_raycast_plane(ko):
    # tests.cpp:454-474 plus the closed form: the hit lies on the plane z = 1
    vol = syn_volume(ko)
    ko.integrate(vol, ko.ANGLE, flat(DEPTH_1M), SYN_INTR, np.eye(4))
    pts, nrm, col, samples = ko.raycast(vol, np.eye(4), SYN_INTR, 16, 16)
    valid = np.isfinite(pts).all(-1) & np.isfinite(nrm).all(-1)
    assert valid.sum() > 0 and samples > 0
    assert np.abs(pts[valid][:, 2] - 1.0).max() < SYN_VOXEL
    assert (nrm[valid][:, 2] < -0.9).all()  # facing the camera


# --- marching cubes ---------------------------------------------------------
def sphere_volume(ko):
    vol = syn_volume(ko)
    n = SYN_RES
    i = np.arange(n ** 3)
    x, y, z = i % n, (i // n) % n, i // (n * n)
    c = np.array([(SYN_ORIGIN[k] + 0.5 * n * SYN_VOXEL) for k in range(3)], np.float32)
    p = np.stack([SYN_ORIGIN[0] + (x + 0.5) * SYN_VOXEL, SYN_ORIGIN[1] + (y + 0.5) * SYN_VOXEL,
                  SYN_ORIGIN[2] + (z + 0.5) * SYN_VOXEL], 1)
    r = 0.3 * n * SYN_VOXEL
    vol.voxels[:, 0] = np.clip((np.linalg.norm(p - c, axis=1) - r) / SYN_TRUNC, -1, 1)
    vol.voxels[:, 1] = 1.0
    return vol, c, r


def test_mesh_sphere_is_closed_outward_manifold(ko):
    # tests.cpp:320-385
    vol, c, r = sphere_volume(ko)
    pos, nrm, col, tri = ko.mesh(vol)
    assert len(pos) and len(tri)
    radial = pos - c
    assert (np.abs(np.linalg.norm(radial, axis=1) - r) < SYN_VOXEL).all()
    assert ((nrm * radial).sum(1) > 0).all()
    face = np.cross(pos[tri[:, 1]] - pos[tri[:, 0]], pos[tri[:, 2]] - pos[tri[:, 0]])
    cen = pos[tri].mean(1)
    assert ((face * (cen - c)).sum(1) > 0).all()
    edges = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), 1)
    _, uses = np.unique(edges, axis=0, return_counts=True)
    assert (uses == 2).all()
    assert len(pos) - len(uses) + len(tri) == 2  # Euler characteristic
    assert (col == 255.0).all()  # no colour observed: default white


def test_mesh_plane_empty_and_min_weight(ko):
    # tests.cpp:387-407 and :439-452
    vol = syn_volume(ko)
    ko.integrate(vol, ko.ANGLE, flat(DEPTH_1M), SYN_INTR, np.eye(4))
    pos, _, _, tri = ko.mesh(vol)
    assert len(tri) and (np.abs(pos[:, 2] - 1.0) < SYN_VOXEL).all()
    assert len(ko.mesh(vol, 2.0)[3]) == 0
    classic = syn_volume(ko)
    ko.integrate(classic, ko.CLASSIC, flat(DEPTH_1M), SYN_INTR, np.eye(4))
    assert len(ko.mesh(classic, 1.0)[3]) > 0
    assert len(ko.mesh(syn_volume(ko))[0]) == 0


def test_triangle_table_crosses_exactly_the_sign_change_edges():
    """Every case lists exactly the edges whose endpoints differ in sign, in
    whole triangles (the property kEdgeTable encodes)."""
    import re
    from pathlib import Path

    src = (Path(__file__).resolve().parents[1] / "oracle" / "mc_cases.inc").read_text()
    cases = re.findall(r'"([0-9a-f]*)"', src)
    assert len(cases) == 256
    ends = [(0, 1), (1, 2), (3, 2), (0, 3), (4, 5), (5, 6), (7, 6), (4, 7), (0, 4), (1, 5), (2, 6), (3, 7)]
    for case, s in enumerate(cases):
        assert len(s) % 3 == 0
        used = {int(ch, 16) for ch in s}
        crossed = {e for e, (a, b) in enumerate(ends) if ((case >> a) & 1) != ((case >> b) & 1)}
        assert used == crossed
        # complementary cases have the same number of triangles except the ambiguous ones
    assert cases[0] == "" and cases[255] == ""


# --- ICP (tests.cpp:880-1074) ---------------------------------------------------
def identical_surface(intr, seeds=None):
    seeds = seeds or [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, -1, 0.5),
                      (-1, 0.5, 1), (0.5, 1, -1)]
    v = np.zeros((3, 3, 3), np.float32)
    n = np.zeros((3, 3, 3), np.float32)
    fx, fy, cx, cy = intr
    for y in range(3):
        for x in range(3):
            z = np.float32(1.0) + np.float32(0.05) * np.float32(x + y)
            v[y, x] = [(x - cx) * z / fx, (y - cy) * z / fy, z]
            s = np.array(seeds[y * 3 + x], np.float32)
            n[y, x] = s / np.float32(np.sqrt(np.float32(s @ s)))
    return v, n


def planar_surface(intr):
    fx, fy, cx, cy = intr
    v = np.array([[[(x - cx) / fx, (y - cy) / fy, 1.0] for x in range(3)] for y in range(3)], np.float32)
    n = np.zeros_like(v)
    n[..., 2] = -1
    return v, n


ICP_INTR = (100.0, 100.0, 1.0, 1.0)


def perturbed():
    a = 0.01
    p = np.eye(4)
    p[:3, :3] = [[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]]
    p[:3, 3] = [0.002, -0.001, 0.0015]
    return p


def test_icp_identity_converges(ko):
    v, n = identical_surface(ICP_INTR)
    opts = ko.icp_options(min_correspondences=6, max_point_distance=0.1, min_normal_dot=0.99,
                          min_system_eigenvalue=1e-12, max_condition_number=1e12)
    pose, r = ko.icp_track_level(v, n, v, n, ICP_INTR, np.eye(4), np.eye(4), 3, opts)
    assert r.status == 1 and r.correspondences == 9
    assert r.update_translation <= 1e-5 and r.update_rotation <= 1e-5


def test_icp_recovers_perturbation(ko):
    v, n = identical_surface(ICP_INTR)
    opts = ko.icp_options(min_correspondences=6, max_point_distance=0.1, min_normal_dot=0.9,
                          min_system_eigenvalue=1e-12, max_condition_number=1e12)
    pose, r = ko.icp_track_level(v, n, v, n, ICP_INTR, np.eye(4), perturbed(), 15, opts)
    assert r.status == 1 and r.correspondences == 9
    assert np.linalg.norm(pose - np.eye(4)) <= 1e-3
    assert r.mean_distance <= 1e-4


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("adaptive", [0, 1])
def test_icp_damped_recovers_perturbation(ko, mode, adaptive):
    v, n = identical_surface(ICP_INTR)
    opts = ko.icp_options(min_correspondences=6, max_point_distance=0.1, min_normal_dot=0.9,
                          min_system_eigenvalue=1e-12, max_condition_number=1e12, damping_mode=mode,
                          damping_lambda=1e-4, adaptive_damping=adaptive)
    pose, r = ko.icp_track_level(v, n, v, n, ICP_INTR, np.eye(4), perturbed(), 25, opts)
    assert r.status in (0, 1) and r.correspondences == 9 and r.lambda_ > 0
    assert np.linalg.norm(pose - np.eye(4)) <= 1e-3


def test_icp_planar_vetoed_then_damped(ko):
    v, n = planar_surface(ICP_INTR)
    opts = ko.icp_options(min_correspondences=6, max_point_distance=0.1, min_normal_dot=0.9)
    _, r = ko.icp_track_level(v, n, v, n, ICP_INTR, np.eye(4), np.eye(4), 3, opts)
    assert r.status == 5  # kUnconstrainedSystem
    opts.damping_mode, opts.damping_lambda = 2, 1e-3
    _, r = ko.icp_track_level(v, n, v, n, ICP_INTR, np.eye(4), np.eye(4), 3, opts)
    assert r.status in (0, 1) and r.correspondences == 9


def test_icp_adaptive_shrinks_lambda(ko):
    v, n = identical_surface(ICP_INTR)
    opts = ko.icp_options(min_correspondences=6, max_point_distance=0.1, min_normal_dot=0.9,
                          min_system_eigenvalue=1e-12, max_condition_number=1e12, damping_mode=1,
                          damping_lambda=1e-2, adaptive_damping=1)
    init = np.eye(4)
    init[:3, 3] = [0.002, -0.001, 0.0015]
    pose, r = ko.icp_track_level(v, n, v, n, ICP_INTR, np.eye(4), init, 25, opts)
    assert r.status in (0, 1) and r.rejected == 0 and r.lambda_ < 1e-2
    assert np.linalg.norm(pose - np.eye(4)) <= 1e-3


def test_icp_empty_is_too_few(ko):
    nan = np.full((2, 2, 3), np.nan, np.float32)
    _, r = ko.icp_track_level(nan, nan, nan, nan, ICP_INTR, np.eye(4), np.eye(4), 10,
                              ko.icp_options(min_correspondences=1))
    assert r.status == 2 and r.correspondences == 0


def test_icp_sums_match_brute_force(ko):
    """Normal equations against an independent numpy evaluation of J = [s x n, n],
    r = n.(v - s) on a random surface."""
    rng = np.random.default_rng(3)
    h, w = 12, 16
    intr = (20.0, 20.0, 7.5, 5.5)
    z = rng.uniform(0.9, 1.1, (h, w)).astype(np.float32)
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float32)
    v = np.stack([(xs - intr[2]) * z / intr[0], (ys - intr[3]) * z / intr[1], z], -1).astype(np.float32)
    nrm = rng.normal(size=(h, w, 3)).astype(np.float32)
    nrm[..., 2] = -np.abs(nrm[..., 2]) - 1.0
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    s64, s32, count = ko.icp_sums(v, nrm, v, nrm, intr, np.eye(4), np.eye(4), 0.05, 0.5)
    assert count == h * w
    J = np.concatenate([np.cross(v.reshape(-1, 3), nrm.reshape(-1, 3)), nrm.reshape(-1, 3)], 1).astype(np.float64)
    JtJ = J.T @ J
    iu = np.triu_indices(6)
    assert np.allclose(s64[:21], JtJ[iu], rtol=1e-5, atol=1e-6)
    assert np.allclose(s64[21:27], 0.0, atol=1e-6)  # identical maps: zero residual
    assert np.allclose(s32[:21], s64[:21], rtol=1e-4, atol=1e-5)


def test_sym_eig6_and_compose(ko):
    rng = np.random.default_rng(5)
    a = rng.normal(size=(6, 6))
    a = a @ a.T + np.eye(6)
    assert np.allclose(np.sort(ko.sym_eig6(a)), np.linalg.eigvalsh(a), rtol=1e-10)
    from scipy.spatial.transform import Rotation

    x = np.array([0.01, -0.02, 0.03, 0.1, 0.2, -0.3])
    p0 = np.eye(4)
    p0[:3, :3] = Rotation.from_rotvec([0.3, 0.1, -0.2]).as_matrix()
    p0[:3, 3] = [1, 2, 3]
    p = ko.compose_increment(x, p0)
    inc = np.eye(4)
    inc[:3, :3] = Rotation.from_rotvec(x[:3]).as_matrix()
    inc[:3, 3] = x[3:]
    assert np.allclose(p, inc @ p0, atol=1e-12)


def test_ate_umeyama_recovers_rigid_motion(ko):
    from scipy.spatial.transform import Rotation

    rng = np.random.default_rng(9)
    gt = rng.normal(size=(50, 3))
    r = Rotation.from_rotvec([0.2, -0.4, 0.1]).as_matrix()
    est = (gt - [1, 2, 3]) @ r  # est = R^T (gt - t)
    assert ko.ate_umeyama(est, gt)["rmse"] < 1e-9
    noisy = est + rng.normal(scale=0.01, size=est.shape)
    assert 0.005 < ko.ate_umeyama(noisy, gt)["rmse"] < 0.02

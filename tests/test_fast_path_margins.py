"""The integration kernel's division-free decisions (kf_integrate.hip,
fuse_voxel) against the exact IEEE float computation, on the CPU.

The kernel computes the projection q' = (fx x) * rcp(z) and the distance
s' = sqrt(x.x) * rcp(lambda) with the hardware reciprocal and square root
(each within 1 ulp, so q' and s' are within a relative 2^-21 of the exact
quotients), and uses them only when they are farther than a margin from a
decision boundary.  Here the approximations are pushed to the worst case
(every relative error in [-2^-21, 2^-21], and 4x that), and every decision
the margins accept must equal the exact one: the rounded pixel index where it
lies in the image (outside it, both paths skip the voxel), and
sd >= truncation / sd < -truncation."""
import numpy as np

F = np.float32
REL = 2.0 ** -21


def lround_half_away(v):
    return np.where(v >= 0, np.floor(v + 0.5), -np.floor(-v + 0.5))


def test_pixel_rounding_margin():
    rng = np.random.default_rng(0)
    n = 400_000
    fx, cx = F(525.0), F(319.5)
    x = rng.uniform(-3, 3, n).astype(F)
    z = rng.uniform(0.05, 5, n).astype(F)
    a = (fx * x).astype(F)
    q = (a / z).astype(F)
    u = (q + cx).astype(F)
    exact = lround_half_away(u.astype(np.float64))
    width = 640
    for scale in (1.0, 4.0):
        for e in (-REL * scale, REL * scale, 0.0):
            qa = (q.astype(np.float64) * (1.0 + e)).astype(F)
            ua = (qa + cx).astype(F)
            nu = np.floor((ua + F(0.5)).astype(F))
            with np.errstate(invalid="ignore"):
                safe = np.abs((ua - nu).astype(F)) < (F(0.5) - F(3e-4)) - F(2e-6) * np.abs(qa)
            # the decision that matters: the pixel index when it lies in the
            # image, or "outside" (both paths then skip the voxel)
            fast_pix = np.where((nu >= 0) & (nu < width), nu, -1)
            exact_pix = np.where((exact >= 0) & (exact < width), exact, -1)
            assert np.array_equal(fast_pix[safe], exact_pix[safe])
            # almost every voxel that lands near the image takes the fast path
            near = np.abs(ua) < 1000
            assert safe[near].mean() > 0.997


def test_pixel_rounding_margin_near_half_integers():
    """Projections aimed within 1e-3 of a half-integer, across the image and
    its borders (u near -0.5 and 639.5), with the worst-case approximation
    errors."""
    rng = np.random.default_rng(5)
    n = 400_000
    fx, cx, width = F(525.0), F(319.5), 640
    z = rng.uniform(0.05, 5, n).astype(F)
    target = rng.integers(-3, width + 3, n) + 0.5 + rng.uniform(-1e-3, 1e-3, n)
    x = ((target - float(cx)) * z / float(fx)).astype(F)
    q = ((fx * x).astype(F) / z).astype(F)
    u = (q + cx).astype(F)
    exact = lround_half_away(u.astype(np.float64))
    exact_pix = np.where((exact >= 0) & (exact < width), exact, -1)
    taken = 0
    for e in (-4 * REL, -REL, 0.0, REL, 4 * REL):
        qa = (q.astype(np.float64) * (1.0 + e)).astype(F)
        ua = (qa + cx).astype(F)
        nu = np.floor((ua + F(0.5)).astype(F))
        safe = np.abs((ua - nu).astype(F)) < (F(0.5) - F(3e-4)) - F(2e-6) * np.abs(qa)
        fast_pix = np.where((nu >= 0) & (nu < width), nu, -1)
        assert np.array_equal(fast_pix[safe], exact_pix[safe])
        taken += int(safe.sum())
    assert taken > 0


def test_signed_distance_margins():
    rng = np.random.default_rng(1)
    n = 400_000
    p = rng.uniform(-2, 2, (n, 3)).astype(F)
    p[:, 2] = rng.uniform(0.3, 4, n).astype(F)
    d = rng.uniform(0.4, 4, n).astype(F)
    lam = rng.uniform(1.0, 1.5, n).astype(F)
    trunc = F(0.05)
    dot = ((p[:, 0] * p[:, 0] + p[:, 1] * p[:, 1]).astype(F) + p[:, 2] * p[:, 2]).astype(F)
    s = (np.sqrt(dot).astype(F) / lam).astype(F)
    sd = (d - s).astype(F)
    for e in (-4 * REL, 4 * REL, 0.0):
        sa = (s.astype(np.float64) * (1.0 + e)).astype(F)
        sda = (d - sa).astype(F)
        margin = F(1e-6) + F(4e-6) * sa
        reject = sda < -trunc - margin
        free = sda >= trunc + margin
        assert np.all(sd[reject] < -trunc)
        assert np.all(sd[free] >= trunc)
        # free space: tsdf = clamp(sd / trunc) is exactly 1
        assert np.all(np.clip((sd[free] / trunc).astype(F), -1, 1) == 1.0)


def test_unit_update_is_exact():
    """((1 w0) + (1 w)) / (w0 + w) is exactly 1 in float: the numerator and
    the denominator are the same rounded sum."""
    rng = np.random.default_rng(2)
    w0 = rng.uniform(0, 196, 100_000).astype(F)
    w = rng.uniform(1e-4, 5, 100_000).astype(F)
    num = ((F(1.0) * w0).astype(F) + (F(1.0) * w).astype(F)).astype(F)
    assert np.all((num / (w0 + w).astype(F)).astype(F) == 1.0)

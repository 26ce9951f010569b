"""The ICP stability check's eigenvalue method, emulated on the CPU.

The device computes the eigenvalues of the 6x6 FP64 normal-equation system
(kf_icp.hip: sym_eig6_wave) in two steps:
* one lane reduces it to tridiagonal form by four Householder reflections
  (H = I - beta v v^T applied as A - v w^T - w v^T with w = p - K v,
  p = beta A v, K = beta v.p / 2);
* 60 lanes then find the six eigenvalues by multisection: the 10 lanes of
  eigenvalue k probe 10 interior points of its interval (initially the
  Gershgorin interval of the tridiagonal matrix), each counting the
  eigenvalues below its point from the sign changes of the Sturm sequence
  p_0 = 1, p_1 = d_0 - x, p_i = (d_{i-1} - x) p_{i-1} - e_{i-2}^2 p_{i-2},
  and keep the sub-interval where the count passes k; 16 rounds shrink the
  interval by 11^16.
The reference (icp_optimizer.cpp:319-347) takes the spectrum from Eigen's
SelfAdjointEigenSolver; this pins the method against numpy.linalg.eigvalsh on
systems as ill-conditioned as the stability check sees (condition numbers up
to 1e8), and on diagonal, rank-deficient and non-finite systems.
"""
import numpy as np

ROUNDS, PROBES = 16, 10


def tridiagonal(a):
    """Householder reduction of a symmetric 6x6 matrix: diagonal d[6] and
    sub-diagonal e[5]."""
    a = np.array(a, np.float64)
    n = 6
    d, e = np.zeros(n), np.zeros(n - 1)
    for k in range(n - 2):
        x = a[k + 1:, k].copy()
        alpha = np.sqrt(np.sum(x * x))
        if x[0] > 0:
            alpha = -alpha
        e[k] = alpha
        v = x.copy()
        v[0] -= alpha
        vn = np.sum(v * v)
        if vn > 0.0:
            beta = 2.0 / vn
            sub = a[k + 1:, k + 1:]
            p = beta * (sub @ v)
            kk = 0.5 * beta * np.dot(v, p)
            w = p - kk * v
            a[k + 1:, k + 1:] = sub - np.outer(v, w) - np.outer(w, v)
        d[k] = a[k, k]
    d[4], d[5], e[4] = a[4, 4], a[5, 5], a[5, 4]
    return d, e


def count_below(d, e, x):
    p_prev, p = 1.0, d[0] - x
    changes = int(p < 0.0 or (p == 0.0 and p_prev > 0.0))
    sign = -1.0 if p < 0.0 or (p == 0.0 and p_prev > 0.0) else 1.0
    for i in range(1, 6):
        p_next = (d[i] - x) * p - e[i - 1] * e[i - 1] * p_prev
        s = -sign if p_next == 0.0 else np.sign(p_next)
        changes += int(s != sign)
        sign, p_prev, p = s, p, p_next
    return changes


def eigenvalues(a):
    d, e = tridiagonal(a)
    if not (np.all(np.isfinite(d)) and np.all(np.isfinite(e))):
        return np.full(6, np.nan)
    ee = np.concatenate([[0.0], np.abs(e), [0.0]])
    lo = np.min(d - ee[:-1] - ee[1:])
    hi = np.max(d + ee[:-1] + ee[1:])
    pad = (hi - lo) * 1e-12 + 1e-300
    lo, hi = lo - pad, hi + pad
    out = np.zeros(6)
    for k in range(6):
        l, h = lo, hi
        for _ in range(ROUNDS):
            xs = [l + (h - l) * (j / (PROBES + 1)) for j in range(1, PROBES + 1)]
            above = [count_below(d, e, x) > k for x in xs]
            j = above.index(True) if True in above else PROBES
            nl = xs[j - 1] if j > 0 else l
            nh = xs[j] if j < PROBES else h
            l, h = nl, nh
        out[k] = 0.5 * (l + h)
    return out


def test_eigenvalues_match_eigvalsh_on_ill_conditioned_systems():
    rng = np.random.default_rng(7)
    worst_norm, worst_min = 0.0, 0.0
    for _ in range(300):
        q, _ = np.linalg.qr(rng.standard_normal((6, 6)))
        ev = np.exp(rng.uniform(0.0, np.log(1e8), 6)) * rng.uniform(1e-2, 1e6)
        a = q @ np.diag(ev) @ q.T
        a = (a + a.T) / 2
        got = eigenvalues(a)
        ref = np.linalg.eigvalsh(a)
        worst_norm = max(worst_norm, np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
        worst_min = max(worst_min, abs(got[0] - ref[0]) / ref[0])
    print(f"worst error / norm {worst_norm:.2e}; worst relative error of the smallest {worst_min:.2e}")
    assert worst_norm < 1e-13
    assert worst_min < 1e-6  # condition numbers up to 1e8


def test_diagonal_rank_deficient_and_non_finite_systems():
    d = np.diag([5.0, 4.0, 3.0, 2.0, 1.0, 0.5])
    assert np.max(np.abs(eigenvalues(d) - np.sort(np.diag(d)))) < 1e-14 * 5.0
    rng = np.random.default_rng(3)
    b = rng.standard_normal((6, 4))
    a = b @ b.T  # rank 4: two zero eigenvalues
    ref = np.linalg.eigvalsh(a)
    assert np.max(np.abs(eigenvalues(a) - ref)) < 1e-13 * np.max(ref)
    a[2, 3] = a[3, 2] = np.nan
    assert np.all(np.isnan(eigenvalues(a)))

"""The pose network's FP64 oracle pinned by the mathematics: central finite
differences of its loss for the analytic gradients, a brute-force
convolution for im2col, the adjoint identity for col2im, and Adam's first
step, which moves every parameter by the learning rate against the sign of
its gradient."""
import numpy as np
import pytest

from oracle import posenet_oracle as P


def small_problem(seed=0, batch=2):
    rng = np.random.default_rng(seed)
    params = np.zeros(P.parameter_count())
    off = 0
    for kind, cin, cout, hin, win, hout, wout, fan in P.layer_shapes():
        params[off: off + cout * fan] = rng.uniform(-1, 1, cout * fan) * np.sqrt(6.0 / fan)
        off += cout * fan
        params[off: off + cout] = rng.uniform(-0.05, 0.05, cout)
        off += cout
    x0 = rng.uniform(0.4, 3.0, (batch, P.IN_H, P.IN_W, 2))
    x0[rng.uniform(size=x0.shape) < 0.05] = 0.0
    targets = rng.normal(0, 0.03, (batch, 6))
    return params, x0, targets


def test_parameter_count():
    # 5 convolutions + 2560 -> 256 -> 6, about 0.79 M parameters
    assert P.parameter_count() == (2 * 9 * 16 + 16 + 16 * 9 * 32 + 32 + 32 * 9 * 64 + 64 + 64 * 9 * 64 + 64
                                   + 64 * 9 * 128 + 128 + 2560 * 256 + 256 + 256 * 6 + 6)
    shapes = P.layer_shapes()
    assert [(s[5], s[6]) for s in shapes[:5]] == [(60, 80), (30, 40), (15, 20), (8, 10), (4, 5)]


def test_im2col_matches_direct_convolution():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 7, 9, 3))
    W = rng.normal(size=(4, 3, 3, 3))  # [out][ky][kx][in]
    col = P.im2col(x)
    y = (col @ W.reshape(4, -1).T).reshape(2, 4, 5, 4)
    direct = np.zeros((2, 4, 5, 4))
    for b in range(2):
        for yo in range(4):
            for xo in range(5):
                for ky in range(3):
                    for kx in range(3):
                        yi, xi = 2 * yo - 1 + ky, 2 * xo - 1 + kx
                        if 0 <= yi < 7 and 0 <= xi < 9:
                            direct[b, yo, xo] += W[:, ky, kx, :] @ x[b, yi, xi]
    assert np.allclose(y, direct, atol=1e-12)


def test_col2im_is_the_adjoint_of_im2col():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(2, 9, 11, 3))
    d = rng.normal(size=P.im2col(x).shape)
    assert np.isclose(np.sum(P.im2col(x) * d), np.sum(x * P.col2im(d, 2, 9, 11, 3)), rtol=1e-12)


def test_gradient_matches_central_differences():
    params, x0, targets = small_problem()
    loss, grad = P.loss_and_grad(params, x0, targets)
    rng = np.random.default_rng(3)
    # a few parameters of every layer, weights and biases
    picks, off = [], 0
    for kind, cin, cout, hin, win, hout, wout, fan in P.layer_shapes():
        picks += list(rng.integers(off, off + cout * fan, 4)) + [off + cout * fan + int(rng.integers(0, cout))]
        off += cout * fan + cout
    eps = 1e-6
    for i in picks:
        p1, p2 = params.copy(), params.copy()
        p1[i] += eps
        p2[i] -= eps
        fd = (P.loss_and_grad(p1, x0, targets)[0] - P.loss_and_grad(p2, x0, targets)[0]) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, grad[i])


def test_bf16_rounding_known_answers():
    assert P.bf16(np.float32(1.0)) == 1.0
    # ties go to even: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7
    assert P.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0
    assert P.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6


def test_adam_first_step_moves_by_the_learning_rate():
    g = np.array([0.3, -2.0, 1e-3])
    p, m, v = P.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, lr=0.01, eps=0.0)
    assert np.allclose(p, -0.01 * np.sign(g), rtol=1e-14)
    g = np.array([0.3, -2.0, 1e-3, 0.0])
    p, m, v = P.adam(np.zeros(4), g, np.zeros(4), np.zeros(4), 1, lr=0.01)
    assert abs(p[3]) == 0.0 and np.allclose(p[:3], -0.01 * np.sign(g[:3]), rtol=1e-4)

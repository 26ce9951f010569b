"""FP64 numpy restatement of the pose-proposal network, its gradients and
Adam -- test infrastructure only (the checker of kinectfusion_amd's
BF16 MFMA kernels in libkf_hip.so).

The network has no counterpart in the reference; its definition is the one
of include/kf_hip.h (kf_posenet_*):
  input  (B, 120, 160, 2) metres minus 1.5, 0 where invalid (previous,
         current coarsest-level depth)
  conv   3x3, stride 2, pad 1, ReLU: 16, 32, 64, 64, 128 channels
  fc     2560 -> 256 (ReLU) -> 6; twist = output * OUT_SCALE
  loss   mean over (B, 6) of (output - target / OUT_SCALE)^2
Parameters are packed per layer as W[out][in] then b[out]; a convolution's
input index is (ky * 3 + kx) * Cin + ci, the flattened map is (y, x, c).

`emulate_bf16` rounds what the device rounds (GEMM operands: weights, layer
outputs, gradients of layer outputs) to bf16 with round-to-nearest-even, so
the comparison isolates accumulation order.  The gradients themselves are
pinned by central finite differences in tests/test_posenet_oracle.py.
"""
from __future__ import annotations

import numpy as np

IN_H, IN_W = 120, 160
CHANNELS = (2, 16, 32, 64, 64, 128)
HIDDEN, OUT = 256, 6
OUT_SCALE = np.array([0.1, 0.1, 0.1, 0.03, 0.03, 0.03])
DEPTH_CENTRE = 1.5


def layer_shapes():
    """[(kind, cin, cout, hin, win, hout, wout, fan_in)] in parameter order."""
    out = []
    h, w = IN_H, IN_W
    for l in range(5):
        ho, wo = (h + 1) // 2, (w + 1) // 2
        out.append(("conv", CHANNELS[l], CHANNELS[l + 1], h, w, ho, wo, 9 * CHANNELS[l]))
        h, w = ho, wo
    flat = h * w * CHANNELS[5]
    out.append(("fc", flat, HIDDEN, 1, 1, 1, 1, flat))
    out.append(("fc", HIDDEN, OUT, 1, 1, 1, 1, HIDDEN))
    return out


def parameter_count() -> int:
    return sum(s[2] * s[7] + s[2] for s in layer_shapes())


def unpack(params: np.ndarray):
    layers, off = [], 0
    for s in layer_shapes():
        cout, fan = s[2], s[7]
        W = params[off: off + cout * fan].reshape(cout, fan)
        off += cout * fan
        b = params[off: off + cout]
        off += cout
        layers.append((W, b))
    return layers


def pack(layers) -> np.ndarray:
    return np.concatenate([np.concatenate([W.reshape(-1), b.reshape(-1)]) for W, b in layers])


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bf16 (nearest even), returned as float64."""
    f = np.asarray(x, np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def inputs_from_pairs(pairs: np.ndarray, depth_scale: float = 1000.0) -> np.ndarray:
    """(B, 2, 120, 160) raw u16 -> (B, 120, 160, 2) centred metres, as the
    device forms them (float32 raw * float32(1 / scale) - 1.5, then bf16)."""
    inv = np.float32(1.0) / np.float32(depth_scale)
    m = np.where(pairs == 0, np.float32(0.0), pairs.astype(np.float32) * inv - np.float32(DEPTH_CENTRE))
    return bf16(np.transpose(m.astype(np.float32), (0, 2, 3, 1)))


def im2col(x: np.ndarray) -> np.ndarray:
    """(B, H, W, C) -> (B * Ho * Wo, 9 * C), k = (ky * 3 + kx) * C + c."""
    B, H, W, Cc = x.shape
    Ho, Wo = (H + 1) // 2, (W + 1) // 2
    xp = np.zeros((B, H + 2, W + 2, Cc), x.dtype)
    xp[:, 1:H + 1, 1:W + 1] = x
    cols = []
    for ky in range(3):
        for kx in range(3):
            cols.append(xp[:, ky: ky + 2 * Ho: 2, kx: kx + 2 * Wo: 2, :])
    return np.concatenate(cols, axis=3).reshape(B * Ho * Wo, 9 * Cc)


def col2im(dcol: np.ndarray, B: int, H: int, W: int, Cc: int) -> np.ndarray:
    Ho, Wo = (H + 1) // 2, (W + 1) // 2
    d = dcol.reshape(B, Ho, Wo, 9, Cc)
    dxp = np.zeros((B, H + 2, W + 2, Cc))
    for ky in range(3):
        for kx in range(3):
            dxp[:, ky: ky + 2 * Ho: 2, kx: kx + 2 * Wo: 2, :] += d[:, :, :, ky * 3 + kx, :]
    return dxp[:, 1:H + 1, 1:W + 1]


def forward(params: np.ndarray, x0: np.ndarray, emulate_bf16: bool = False):
    """Returns (output (B, 6) normalised twist, cache)."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unpack(np.asarray(params, np.float64))
    shapes = layer_shapes()
    B = x0.shape[0]
    acts, cols = [r(x0)], []
    x = acts[0]
    for l in range(5):
        W, b = layers[l]
        _, cin, cout, h, w, ho, wo, fan = shapes[l]
        col = im2col(x)
        cols.append(col)
        y = np.maximum(col @ r(W).T + b, 0.0)
        x = r(y).reshape(B, ho, wo, cout)
        acts.append(x)
    W, b = layers[5]
    h = r(np.maximum(x.reshape(B, -1) @ r(W).T + b, 0.0))
    acts.append(h)
    W, b = layers[6]
    out = h @ r(W).T + b
    return out, {"acts": acts, "cols": cols, "layers": layers, "emulate": emulate_bf16}


def loss_and_grad(params: np.ndarray, x0: np.ndarray, targets: np.ndarray, emulate_bf16: bool = False):
    """(loss, packed gradient) of the mean squared error."""
    out, cache = forward(params, x0, emulate_bf16)
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    B = x0.shape[0]
    diff = out - np.asarray(targets, np.float64) / OUT_SCALE
    loss = float(np.mean(diff * diff))
    dout = r(2.0 * diff / diff.size)
    acts, cols, layers = cache["acts"], cache["cols"], cache["layers"]
    shapes = layer_shapes()
    grads = [None] * 7
    # fc2
    W, _ = layers[6]
    h = acts[6]
    grads[6] = (dout.T @ h, dout.sum(0))
    dh = r((dout @ r(W)) * (h > 0))
    # fc1
    W, _ = layers[5]
    x5 = acts[5].reshape(B, -1)
    grads[5] = (dh.T @ x5, dh.sum(0))
    dz = r((dh @ r(W)) * (x5 > 0)).reshape(acts[5].shape)
    for l in range(4, -1, -1):
        W, _ = layers[l]
        _, cin, cout, hh, ww, ho, wo, fan = shapes[l]
        dz2 = dz.reshape(-1, cout)
        grads[l] = (dz2.T @ cols[l], dz2.sum(0))
        if l == 0:
            break
        dcol = dz2 @ r(W)
        dz = r(col2im(dcol, B, hh, ww, cin) * (acts[l] > 0))
    return loss, pack(grads)


def adam(params, grads, m, v, step: int, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, algorithm 1) in FP64; returns
    (params, m, v)."""
    m = beta1 * m + (1.0 - beta1) * grads
    v = beta2 * v + (1.0 - beta2) * grads * grads
    c1 = 1.0 / (1.0 - beta1 ** step)
    c2 = 1.0 / (1.0 - beta2 ** step)
    return params - lr * (m * c1) / (np.sqrt(v * c2) + eps), m, v

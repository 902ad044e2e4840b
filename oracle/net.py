"""Float64 numpy restatement of the learned denoiser (features, forward,
backward, loss, Adam).

TEST INFRASTRUCTURE. The learned denoiser has no counterpart in the
reference (its only denoiser is the a-trous filter, crates/core/src/denoise.rs);
its inputs are the reference's Accum channels (gpu_layout.rs:345). This module
is the checker the HIP kernels are compared with; it is pinned by brute-force
convolution loops, central finite differences and Adam's closed form in the
CPU test suite. Only tests import it.

With `bf16=True` every tensor the device stores in BF16 (features, weights,
activations, activation gradients) is rounded to BF16 (round to nearest even)
at the same points, so the remaining difference from the device is the order
of FP32 accumulation.
"""
from __future__ import annotations

import numpy as np

# (cin, cout, relu) per layer; encoder levels 0..3, decoder 2..0, output.
LAYERS = [
    (16, 32, 1), (32, 32, 1), (32, 48, 1), (48, 48, 1), (48, 64, 1), (64, 64, 1), (64, 96, 1), (96, 96, 1),
    (160, 64, 1), (64, 64, 1), (112, 48, 1), (48, 48, 1), (80, 32, 1), (32, 32, 1), (32, 16, 0),
]
LEVEL = [0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0]
INPUTS_USED = 15


def param_count() -> int:
    return sum(9 * ci * co + co for ci, co, _ in LAYERS)


def unpack(params: np.ndarray):
    """Flat parameters -> [(W[9, cin, cout], b[cout])]."""
    out, off = [], 0
    for ci, co, _ in LAYERS:
        W = params[off:off + 9 * ci * co].reshape(9, ci, co)
        off += 9 * ci * co
        out.append((W, params[off:off + co]))
        off += co
    return out


def pack(layers) -> np.ndarray:
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def to_bf16(x: np.ndarray) -> np.ndarray:
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


# ---------------------------------------------------------------------------
# Features
# ---------------------------------------------------------------------------

def _means(accum: np.ndarray):
    """accum: float array [..., 16] in GpuAccum order."""
    a = np.asarray(accum, np.float64)
    ns = np.maximum(a[..., 3], 1.0)
    inv = 1.0 / ns
    al = np.maximum(a[..., 4:7] * inv[..., None], 1e-3)
    rd = a[..., 0:3] * inv[..., None]
    return a, ns, inv, al, rd


def features(accum: np.ndarray, base: np.ndarray | None = None) -> np.ndarray:
    """16 input channels per pixel from Accum sums (shape [..., 16]); `base`
    is the a-trous output of the same film (Accum-shaped), channels 12..14."""
    a, ns, inv, al, rd = _means(accum)
    f = np.zeros(a.shape[:-1] + (16,))
    f[..., 0:3] = np.log1p(np.maximum(rd / al, 0.0))
    f[..., 3:6] = al
    f[..., 6:9] = a[..., 8:11] * inv[..., None]
    dep = np.maximum(a[..., 7] * inv, 0.0)
    f[..., 9] = dep / (1.0 + dep)
    lw = np.array([0.2126, 0.7152, 0.0722])
    var = np.maximum(a[..., 12:15] * inv[..., None] - rd * rd, 0.0)
    var_l = (var * lw).sum(-1)
    al_l = (al * lw).sum(-1)
    f[..., 10] = np.log1p(np.sqrt(var_l * inv) / al_l)
    f[..., 11] = np.log2(ns) / 16.0
    if base is not None:
        b = np.asarray(base, np.float64)
        br = b[..., 0:3] / np.maximum(b[..., 3], 1.0)[..., None]
        f[..., 12:15] = np.log1p(np.maximum(br / al, 0.0))
    return f


def target(accum_in: np.ndarray, accum_ref: np.ndarray) -> np.ndarray:
    """log1p(reference radiance / input albedo), 3 channels."""
    _, _, _, al, _ = _means(accum_in)
    r = np.asarray(accum_ref, np.float64)
    rr = r[..., 0:3] / np.maximum(r[..., 3], 1.0)[..., None]
    return np.log1p(np.maximum(rr / al, 0.0))


def compose(out: np.ndarray, accum_in: np.ndarray, x: np.ndarray, base: np.ndarray | None = None) -> np.ndarray:
    """Network output + the a-trous estimate -> Accum-shaped film (radiance,
    samples = 1, guides). `base` is the estimate's log channels at full
    precision (the device keeps them in FP32); by default the BF16-stored
    channels 12..14 of the features `x`."""
    a, _, inv, al, _ = _means(accum_in)
    res = np.zeros(a.shape[:-1] + (16,))
    b = x[..., 12:15] if base is None else base
    res[..., 0:3] = np.expm1(np.maximum(out[..., 0:3] + b, 0.0)) * al
    res[..., 3] = 1.0
    res[..., 4:7] = al
    res[..., 7] = a[..., 7] * inv
    res[..., 8:11] = a[..., 8:11] * inv[..., None]
    return res


# ---------------------------------------------------------------------------
# Building blocks (NHWC, float64)
# ---------------------------------------------------------------------------

def conv3x3(x, W, b=None):
    """y[p, co] = b[co] + sum_{tap, ci} x[p + (dy - 1, dx - 1), ci] W[tap, ci, co], zero padding."""
    n, h, w, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((n, h, w, W.shape[2]))
    for tap in range(9):
        dy, dx = divmod(tap, 3)
        y += xp[:, dy:dy + h, dx:dx + w, :] @ W[tap]
    if b is not None:
        y += b
    return y


def conv3x3_grads(x, W, dy_out):
    """(dx, dW, db) of conv3x3 given dL/dy."""
    n, h, w, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dW = np.zeros_like(W)
    flat_dy = dy_out.reshape(-1, dy_out.shape[-1])
    for tap in range(9):
        dy, dx = divmod(tap, 3)
        sh = xp[:, dy:dy + h, dx:dx + w, :].reshape(-1, x.shape[-1])
        dW[tap] = sh.T @ flat_dy
        dxp[:, dy:dy + h, dx:dx + w, :] += dy_out @ W[tap].T
    return dxp[:, 1:-1, 1:-1, :], dW, flat_dy.sum(0)


def pool(x):
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def pool_bwd(d):
    return np.repeat(np.repeat(d, 2, axis=1), 2, axis=2) * 0.25


def upcat(low, skip):
    up = np.repeat(np.repeat(low, 2, axis=1), 2, axis=2)
    return np.concatenate([up, skip], axis=-1)


def upcat_bwd(d, c1):
    n, h, w, _ = d.shape
    dlow = d[..., :c1].reshape(n, h // 2, 2, w // 2, 2, c1).sum(axis=(2, 4))
    return dlow, d[..., c1:]


# ---------------------------------------------------------------------------
# Network
# ---------------------------------------------------------------------------

def forward(params, x, bf16=False):
    """Returns the activations: 'x', 'l0'..'l14', 'p0'..'p2', 'c0'..'c2'."""
    R = to_bf16 if bf16 else (lambda v: v)
    layers = [(R(W), b) for W, b in unpack(np.asarray(params, np.float64))]
    A = {"x": R(x)}

    def conv(i, inp):
        W, b = layers[i]
        y = conv3x3(inp, W, b)
        if LAYERS[i][2]:
            y = np.maximum(y, 0.0)
        A[f"l{i}"] = R(y)
        return A[f"l{i}"]

    conv(1, conv(0, A["x"]))
    A["p0"] = R(pool(A["l1"]))
    conv(3, conv(2, A["p0"]))
    A["p1"] = R(pool(A["l3"]))
    conv(5, conv(4, A["p1"]))
    A["p2"] = R(pool(A["l5"]))
    conv(7, conv(6, A["p2"]))
    A["c2"] = upcat(A["l7"], A["l5"])
    conv(9, conv(8, A["c2"]))
    A["c1"] = upcat(A["l9"], A["l3"])
    conv(11, conv(10, A["c1"]))
    A["c0"] = upcat(A["l11"], A["l1"])
    conv(14, conv(13, conv(12, A["c0"])))
    return A


def activations_from_layers(x, layer_outputs, bf16=False):
    """The full activation set from given layer outputs l0..l14 (e.g. the
    device's), with the pools and concatenations recomputed from them."""
    R = to_bf16 if bf16 else (lambda v: v)
    A = {"x": np.asarray(x, np.float64)}
    for i, v in enumerate(layer_outputs):
        A[f"l{i}"] = np.asarray(v, np.float64)
    A["p0"], A["p1"], A["p2"] = R(pool(A["l1"])), R(pool(A["l3"])), R(pool(A["l5"]))
    A["c2"] = upcat(A["l7"], A["l5"])
    A["c1"] = upcat(A["l9"], A["l3"])
    A["c0"] = upcat(A["l11"], A["l1"])
    return A


def loss_and_grad(params, x, tgt, bf16=False, acts=None, base=None):
    """Mean squared error of (output + the a-trous estimate) against the
    target over the 3 used output channels, and its gradient
    with respect to the flat parameters. Also returns the activation
    gradients ('d*' keys), computed as the device stores them. `acts`
    (from activations_from_layers) replaces the forward pass, so the backward
    pass can be checked on its own. `base` is the estimate's log channels at
    full precision (the device keeps them in FP32), by default channels
    12..14 of x."""
    R = to_bf16 if bf16 else (lambda v: v)
    A = forward(params, x, bf16) if acts is None else acts
    layers = [(R(W), b) for W, b in unpack(np.asarray(params, np.float64))]
    out = A["l14"]
    e = (out[..., :3] + (A["x"][..., 12:15] if base is None else base)) - tgt
    loss = float((e * e).mean())
    d = np.zeros_like(out)
    d[..., :3] = R(2.0 / e.size * e)
    D = {"l14": d}
    grads = [None] * len(LAYERS)

    def back(i, inp):
        """d(layer i output) in D -> grads[i]; returns d(input)."""
        dy = D[f"l{i}"]
        if LAYERS[i][2]:
            dy = dy * (A[f"l{i}"] > 0)
        dx, dW, db = conv3x3_grads(inp, layers[i][0], dy)
        grads[i] = (dW, db)
        return R(dx)

    D["l13"] = back(14, A["l13"])
    D["l12"] = back(13, A["l12"])
    dc0 = back(12, A["c0"])
    dl11, dskip0 = upcat_bwd(dc0, 48)
    D["l11"] = R(dl11)
    D["l10"] = back(11, A["l10"])
    dc1 = back(10, A["c1"])
    dl9, dskip1 = upcat_bwd(dc1, 64)
    D["l9"] = R(dl9)
    D["l8"] = back(9, A["l8"])
    dc2 = back(8, A["c2"])
    dl7, dskip2 = upcat_bwd(dc2, 96)
    D["l7"] = R(dl7)
    D["l6"] = back(7, A["l6"])
    dp2 = back(6, A["p2"])
    D["l5"] = R(pool_bwd(dp2) + dskip2)
    D["l4"] = back(5, A["l4"])
    dp1 = back(4, A["p1"])
    D["l3"] = R(pool_bwd(dp1) + dskip1)
    D["l2"] = back(3, A["l2"])
    dp0 = back(2, A["p0"])
    D["l1"] = R(pool_bwd(dp0) + dskip0)
    D["l0"] = back(1, A["l0"])
    back(0, A["x"])
    return loss, pack(grads), A, D


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba, Algorithm 1) at step t >= 1; returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights on the used channels, zero biases (numpy stream; the
    device initialiser uses its own stream)."""
    rng = np.random.default_rng(seed)
    layers = []
    for i, (ci, co, _) in enumerate(LAYERS):
        ci_used = INPUTS_USED if i == 0 else ci
        co_used = 3 if i == len(LAYERS) - 1 else co
        W = np.zeros((9, ci, co))
        std = np.sqrt(2.0 / (9 * ci_used)) * (0.1 if i == len(LAYERS) - 1 else 1.0)
        W[:, :ci_used, :co_used] = rng.normal(size=(9, ci_used, co_used)) * std
        layers.append((W, np.zeros(co)))
    return pack(layers)

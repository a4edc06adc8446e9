"""CPU oracle of the learned predictor's network (numpy). TEST INFRASTRUCTURE ONLY: the checker for the
GPU forward/backward/Adam kernels; the product never imports it.

MLP 14 -> 128 -> 128 -> 128 -> 3 with SiLU, parameters laid out W1[128x14], b1, W2[128x128], b2,
W3[128x128], b3, W4[3x128], b4 (W maps inputs to outputs, row-major), loss = mean((y - t)^2).
`bf16=True` rounds every matrix-product operand to BF16 (round to nearest even) and accumulates in FP32,
which is what the MFMA kernels do; `bf16=False` is plain FP64 (used to pin the backward pass by finite
differences).
"""
import numpy as np

N_IN, N_HID, N_OUT = 14, 128, 3
SHAPES = [("W1", (N_HID, N_IN)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
          ("W3", (N_HID, N_HID)), ("b3", (N_HID,)), ("W4", (N_OUT, N_HID)), ("b4", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = p[o:o + n].reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([d[name].ravel() for name, _ in SHAPES])


def to_bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def _mm(a, b, bf16):
    if bf16:
        return to_bf16(a).astype(np.float32) @ to_bf16(b).astype(np.float32)
    return a @ b


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, bf16=False):
    dt = np.float32 if bf16 else np.float64
    P = {k: v.astype(dt) for k, v in unpack(np.asarray(p)).items()}
    a = np.asarray(x, dtype=dt)
    cache = [a]
    zs = []
    for l in (1, 2, 3):
        z = _mm(a, P[f"W{l}"].T, bf16) + P[f"b{l}"]
        zs.append(z)
        a = silu(z).astype(dt)
        cache.append(a)
    y = _mm(a, P["W4"].T, bf16) + P["b4"]
    return y, zs, cache


def loss_grad(p, x, t, bf16=False):
    dt = np.float32 if bf16 else np.float64
    P = {k: v.astype(dt) for k, v in unpack(np.asarray(p)).items()}
    y, zs, acts = forward(p, x, bf16)
    n = y.shape[0]
    d = y - np.asarray(t, dtype=dt)
    loss = float(np.mean(d.astype(np.float64) ** 2))
    dy = (2.0 * d / (n * N_OUT)).astype(dt)
    g = {}
    g["W4"] = _mm(dy.T, acts[3], bf16)
    g["b4"] = dy.sum(0)
    da = _mm(dy, P["W4"], bf16)
    for l in (3, 2, 1):
        dz = (da * dsilu(zs[l - 1])).astype(dt)
        g[f"W{l}"] = _mm(dz.T, acts[l - 1], bf16)
        g[f"b{l}"] = dz.sum(0)
        if l > 1:
            da = _mm(dz, P[f"W{l}"], bf16)
    return loss, pack({k: np.asarray(v, dtype=np.float64) for k, v in g.items()})


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1 = 1.0 / (1.0 - b1 ** step)
    c2 = 1.0 / (1.0 - b2 ** step)
    return p - lr * (m * c1) / (np.sqrt(v * c2) + eps), m, v

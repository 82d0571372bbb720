"""CPU oracle of the learned heat surrogate (numpy, float64).  TEST INFRASTRUCTURE ONLY: imported
by tests/ as the checker of the gfx950 convolution and U-Net kernels; nothing in the product
imports it.

There is no counterpart in the reference (pantometry has no learned component), so this module
is pinned by mathematics alone (tests/test_surrogate_oracle.py):
  - conv3d against scipy.ndimage.correlate, channel pair by channel pair;
  - conv3d's data and weight gradients against central finite differences, and the adjoint
    identity <conv(x), dy> = <x, conv^T(dy)>;
  - pool / upsample against their adjoints and hand-computed 2x2x2 cases;
  - gelu' against finite differences of gelu;
  - the whole network's loss gradient against central finite differences on sampled parameters.

Layout (the device's): activations [B][D][H][W][C] channels-last; conv weights [Cout][27][Cin]
with tap t = 9*dz + 3*dy + dx for offsets (dz-1, dy-1, dx-1); zero padding.  With `bf16=True` the
forward and backward round to bf16 wherever the device stores bf16, so the two agree up to fp32
accumulation order.
"""
from __future__ import annotations

import math
from typing import Dict, List, Tuple

import numpy as np

T_SHIFT, T_SCALE, K_SCALE, OUT_SCALE = 300.0, 100.0, 0.5, 10.0

# (cin, cout, level) of the fourteen 3x3x3 convolutions, encoder then decoder.
LAYERS: List[Tuple[int, int, int]] = [
    (16, 16, 0), (16, 16, 0),
    (16, 32, 1), (32, 32, 1),
    (32, 64, 2), (64, 64, 2),
    (64, 128, 3), (128, 128, 3),
    (192, 64, 2), (64, 64, 2),
    (96, 32, 1), (32, 32, 1),
    (48, 16, 0), (16, 16, 0),
]


def round_bf16(x: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even to bf16 through fp32, returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _id(x):
    return x


# ---- layers ---------------------------------------------------------------------------------

def _taps():
    for t in range(27):
        yield t, t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1


def _shifted(xp: np.ndarray, dz: int, dy: int, dx: int, D: int, H: int, W: int) -> np.ndarray:
    """xp padded by one on D/H/W: the view x[p + (dz, dy, dx)] over all positions p."""
    return xp[:, 1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W, :]


def conv3d(x: np.ndarray, w: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
    """y[p][co] = sum_t sum_ci w[co][t][ci] x[p + off(t)][ci] + b[co]."""
    B, D, H, W, Cin = x.shape
    Cout = w.shape[0]
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((B, D, H, W, Cout))
    for t, dz, dy, dx in _taps():
        y += _shifted(xp, dz, dy, dx, D, H, W) @ w[:, t, :].T
    if b is not None:
        y += b
    return y


def conv3d_backward(x: np.ndarray, w: np.ndarray, dy: np.ndarray):
    """(dx, dw, db) of sum(y * dy) for y = conv3d(x, w, b)."""
    B, D, H, W, Cin = x.shape
    Cout = w.shape[0]
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    dyf = dy.reshape(-1, Cout)
    for t, dz, dyy, dx in _taps():
        dw[:, t, :] = dyf.T @ _shifted(xp, dz, dyy, dx, D, H, W).reshape(-1, Cin)
        dxp[:, 1 + dz:1 + dz + D, 1 + dyy:1 + dyy + H, 1 + dx:1 + dx + W, :] += dy @ w[:, t, :]
    return dxp[:, 1:-1, 1:-1, 1:-1, :], dw, dyf.sum(axis=0)


_GK = 0.7978845608028654


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(_GK * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(_GK * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * _GK * (1.0 + 3.0 * 0.044715 * z * z)


def pool(x: np.ndarray) -> np.ndarray:
    B, D, H, W, C = x.shape
    return x.reshape(B, D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4, 6))


def pool_backward(dy: np.ndarray) -> np.ndarray:
    return np.repeat(np.repeat(np.repeat(dy, 2, 1), 2, 2), 2, 3) * 0.125


def upsample(x: np.ndarray) -> np.ndarray:
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def upsample_backward(dy: np.ndarray) -> np.ndarray:
    B, D, H, W, C = dy.shape
    return dy.reshape(B, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6))


# ---- the network ----------------------------------------------------------------------------

def parameter_layout() -> Tuple[List[Tuple[int, int]], int, int]:
    """[(weight offset, bias offset)] per layer, the head's offset, and the total count."""
    off, out = 0, []
    for cin, cout, _ in LAYERS:
        w = off
        off += cout * 27 * cin
        out.append((w, off))
        off += cout
    return out, off, off + 17


def unpack(params: np.ndarray):
    lay, head, _ = parameter_layout()
    ws, bs = [], []
    for (cin, cout, _), (wo, bo) in zip(LAYERS, lay):
        ws.append(params[wo:wo + cout * 27 * cin].reshape(cout, 27, cin))
        bs.append(params[bo:bo + cout])
    return ws, bs, params[head:head + 16], params[head + 16]


K_STEPS = 16


def pack_input(t: np.ndarray, k: np.ndarray, coupling, bf16: bool = False) -> np.ndarray:
    """[B][D][H][W][16] input channels of cubic samples t, k [B][z][y][x] with the per-sample
    coupling a = dt * dx / C:
      0 (T - 300)/100, 1 log k / 0.5, 2..7 10 a k_f for faces x-, x+, y-, y+, z-, z+ (harmonic
      mean, 0 on the surface), 8..13 a k_f (T_nb - T) * 16 / 10, 14 their sum, 15 one."""
    t = np.asarray(t, dtype=np.float64)
    k = np.asarray(k, dtype=np.float64)
    B = t.shape[0]
    a = np.broadcast_to(np.asarray(coupling, dtype=np.float64), (B,)).astype(np.float32)
    x = np.zeros(t.shape + (16,))
    x[..., 0] = ((t - T_SHIFT) / T_SCALE).astype(np.float32)
    x[..., 1] = (np.log(k) / K_SCALE).astype(np.float32)
    flux_sum = np.zeros(t.shape, dtype=np.float32)
    q = 0
    for axis in (3, 2, 1):                 # x, y, z
        for side in (-1, 1):
            kn = np.roll(k, -side, axis=axis)
            tn = np.roll(t, -side, axis=axis)
            kf = 2.0 * k * kn / (k + kn)
            edge = [slice(None)] * 4
            edge[axis] = 0 if side < 0 else -1
            on = np.ones(t.shape, dtype=bool)
            on[tuple(edge)] = False
            af = np.where(on, (a[:, None, None, None] * kf).astype(np.float32), np.float32(0.0))
            fl = np.where(on, (a[:, None, None, None].astype(np.float64) * kf * (tn - t)
                                * (K_STEPS / OUT_SCALE)).astype(np.float32), np.float32(0.0))
            x[..., 2 + q] = np.float32(10.0) * af
            x[..., 8 + q] = fl
            flux_sum = (flux_sum + fl).astype(np.float32)
            q += 1
    x[..., 14] = flux_sum
    x[..., 15] = 1.0
    return round_bf16(x) if bf16 else x


def forward(params: np.ndarray, x: np.ndarray, bf16: bool = False):
    """The network output y [B][D][H][W] and the cache for `backward`."""
    r = round_bf16 if bf16 else _id
    ws, bs, hw, hb = unpack(params)
    if bf16:
        ws = [round_bf16(w) for w in ws]
    cache: Dict[str, object] = {"x": x, "ws": ws, "hw": hw}
    z, a = [None] * 14, [None] * 14

    def conv(i, inp):
        zz = conv3d(inp, ws[i], bs[i])
        z[i] = r(zz)
        a[i] = r(gelu(zz))
        return a[i]

    h = conv(1, conv(0, x))
    p0 = r(pool(a[1]))
    conv(3, conv(2, p0))
    p1 = r(pool(a[3]))
    conv(5, conv(4, p1))
    p2 = r(pool(a[5]))
    conv(7, conv(6, p2))
    c2 = np.concatenate([upsample(a[7]), a[5]], axis=-1)
    conv(9, conv(8, c2))
    c1 = np.concatenate([upsample(a[9]), a[3]], axis=-1)
    conv(11, conv(10, c1))
    c0 = np.concatenate([upsample(a[11]), a[1]], axis=-1)
    conv(13, conv(12, c0))
    y = a[13] @ hw + hb
    cache.update(z=z, a=a, p=[p0, p1, p2], c=[c0, c1, c2])
    del h
    return y, cache


def loss(y: np.ndarray, target: np.ndarray) -> float:
    return 0.5 * float(np.mean((y - target) ** 2))


def backward(params: np.ndarray, cache, y: np.ndarray, target: np.ndarray,
             bf16: bool = False) -> np.ndarray:
    """The gradient of loss(y, target) with respect to the flat parameter array."""
    r = round_bf16 if bf16 else _id
    ws, hw = cache["ws"], cache["hw"]
    z, a = cache["z"], cache["a"]
    p0, p1, p2 = cache["p"]
    c0, c1, c2 = cache["c"]
    lay, head, total = parameter_layout()
    g = np.zeros(total)
    dy = (y - target) / y.size
    g[head:head + 16] = np.tensordot(dy, a[13], axes=(list(range(dy.ndim)), list(range(dy.ndim))))
    g[head + 16] = dy.sum()
    da = [None] * 14
    da[13] = r(dy[..., None] * hw)

    def back(i, inp, want_dx=True):
        cin, cout, _ = LAYERS[i]
        dz = r(da[i] * gelu_grad(z[i]))
        dx, dw, db = conv3d_backward(inp, ws[i], dz)
        wo, bo = lay[i]
        g[wo:wo + cout * 27 * cin] = dw.reshape(-1)
        g[bo:bo + cout] = db
        return r(dx) if want_dx else None

    da[12] = back(13, a[12])
    dc0 = back(12, c0)
    da[11] = r(upsample_backward(dc0[..., :32]))
    da[1] = dc0[..., 32:]
    da[10] = back(11, a[10])
    dc1 = back(10, c1)
    da[9] = r(upsample_backward(dc1[..., :64]))
    da[3] = dc1[..., 64:]
    da[8] = back(9, a[8])
    dc2 = back(8, c2)
    da[7] = r(upsample_backward(dc2[..., :128]))
    da[5] = dc2[..., 128:]
    da[6] = back(7, a[6])
    dp2 = back(6, p2)
    da[5] = r(da[5] + pool_backward(dp2))
    da[4] = back(5, a[4])
    dp1 = back(4, p1)
    da[3] = r(da[3] + pool_backward(dp1))
    da[2] = back(3, a[2])
    dp0 = back(2, p0)
    da[1] = r(da[1] + pool_backward(dp0))
    da[0] = back(1, a[0])
    back(0, cache["x"], want_dx=False)
    return g


def init_parameters(seed: int) -> np.ndarray:
    """He-uniform weights (bound sqrt(6 / (27 cin))) from xorshift64* on the seed, zero biases,
    zero head: the untrained network predicts no change.  The device's initialisation."""
    lay, head, total = parameter_layout()
    p = np.zeros(total, dtype=np.float32)
    st = seed if seed != 0 else 0x9E3779B97F4A7C15
    mask = (1 << 64) - 1
    for (cin, cout, _), (wo, _) in zip(LAYERS, lay):
        bound = math.sqrt(6.0 / (27 * cin))
        n = cout * 27 * cin
        vals = np.empty(n)
        for j in range(n):
            st ^= st >> 12
            st ^= (st << 25) & mask
            st ^= st >> 27
            u = (((st * 0x2545F4914F6CDD1D) & mask) >> 11) / float(1 << 53)
            vals[j] = u * 2.0 - 1.0
        p[wo:wo + n] = (vals * bound).astype(np.float32)
    return p


def adam_step(p: np.ndarray, g: np.ndarray, m: np.ndarray, v: np.ndarray, t: int, lr: float,
              b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8, dtype=np.float64):
    """One bias-corrected Adam step (Kingma and Ba, 2015, Algorithm 1), step number t counted
    from 1: m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2, p -= lr (m / (1-b1^t)) /
    (sqrt(v / (1-b2^t)) + eps).  Returns (p, m, v) in `dtype`.  With dtype=np.float32 every
    constant and operation is rounded to float32 in this order, as the device's optimiser runs
    it (fp32 master parameters and moments; the corrections 1-b^t formed in fp32 on the host)."""
    f = dtype
    p, g, m, v = (np.asarray(a, dtype=f) for a in (p, g, m, v))
    b1, b2, lr, eps = f(b1), f(b2), f(lr), f(eps)
    one = f(1.0)
    m = b1 * m + (one - b1) * g
    v = b2 * v + (one - b2) * g * g
    c1 = one - f(b1 ** f(t))
    c2 = one - f(b2 ** f(t))
    return (p - lr * (m / c1) / (np.sqrt(v / c2) + eps)).astype(f), m, v

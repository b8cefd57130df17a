"""CPU oracle of the learned-safety network: forward pass, loss and gradient
in numpy float64.  Test infrastructure only.

The network is the library's (vecgeom_amd/csrc/mlp.h): 51 positional-encoding
features of the normalised point u in [-1, 1]^3 (u, then for k = 0..7 the
sines and cosines of 2^k pi u per axis), four 128-unit SiLU layers and one
linear output.  Public parameter order: W1[128][51], b1[128], W2[128][128],
b2, W3, b3, W4, b4, w5[128], b5.

With emulate=True the BF16 rounding points of the GPU kernels are reproduced
(features, weights and activations rounded to BF16 before each product; the
output-layer gradient dZ rounded before the products that use it); sums are
float64 where the GPU accumulates in FP32.  With emulate=False everything is
exact float64 arithmetic, which the finite-difference tests pin.

It has no counterpart in the reference (VecGeom has no learned safety); the
quantity it approximates is VSafetyEstimator's safety
(VecGeom/navigation/VSafetyEstimator.h:45-80), computed by the shape oracles.
"""
import numpy as np

H = 128
FREQS = 8
NIN = 3 + 6 * FREQS
NPARAMS = H * NIN + H + 3 * (H * H + H) + H + 1
assert NPARAMS == 56321


def bf16(x):
    """Round to BF16 (nearest, ties to even) through float32; float64 result."""
    f = np.asarray(x, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def features(u):
    """(n, 3) normalised coordinates -> (n, 51) features."""
    u = np.asarray(u, dtype=np.float64).reshape(-1, 3)
    cols = [u[:, 0], u[:, 1], u[:, 2]]
    for k in range(FREQS):
        a = (2.0 ** k) * np.pi * u
        cols += [np.sin(a[:, 0]), np.sin(a[:, 1]), np.sin(a[:, 2])]
        cols += [np.cos(a[:, 0]), np.cos(a[:, 1]), np.cos(a[:, 2])]
    return np.stack(cols, axis=1)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0

    def take(*shape):
        nonlocal o
        n = int(np.prod(shape))
        r = p[o:o + n].reshape(shape)
        o += n
        return r

    W1, b1 = take(H, NIN), take(H)
    W2, b2 = take(H, H), take(H)
    W3, b3 = take(H, H), take(H)
    W4, b4 = take(H, H), take(H)
    w5, b5 = take(H), take(1)
    assert o == NPARAMS
    return [W1, W2, W3, W4], [b1, b2, b3, b4], w5, b5[0]


def pack(Ws, bs, w5, b5):
    parts = []
    for W, b in zip(Ws, bs):
        parts += [np.ravel(W), np.ravel(b)]
    parts += [np.ravel(w5), np.atleast_1d(b5)]
    return np.concatenate(parts)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, u, emulate=True):
    """Returns y (n,) and the cache for backward."""
    r = bf16 if emulate else (lambda x: np.asarray(x, dtype=np.float64))
    Ws, bs, w5, b5 = unpack(params)
    h = r(features(u))
    hs, zs = [h], []
    for W, b in zip(Ws, bs):
        z = h @ r(W).T + b
        h = r(silu(z))
        zs.append(z)
        hs.append(h)
    y = h @ w5 + b5
    return y, (hs, zs)


def loss_and_grad(params, u, target, over_weight=4.0, emulate=True):
    """Mean asymmetric squared error (overestimates weighted by over_weight)
    and its gradient in public parameter order."""
    r = bf16 if emulate else (lambda x: np.asarray(x, dtype=np.float64))
    Ws, bs, w5, b5 = unpack(params)
    y, (hs, zs) = forward(params, u, emulate)
    t = np.asarray(target, dtype=np.float64)
    n = len(t)
    e = y - t
    w = np.where(e > 0, over_weight, 1.0)
    loss = float(np.sum(w * e * e) / n)
    dy = 2.0 * w * e / n
    dw5 = hs[4].T @ dy
    db5 = np.sum(dy)
    g = np.outer(dy, w5)
    dWs, dbs = [None] * 4, [None] * 4
    for l in range(3, -1, -1):
        dz = g * dsilu(zs[l])
        dbs[l] = dz.sum(axis=0)
        dzr = r(dz)
        dWs[l] = dzr.T @ hs[l]
        if l:
            g = dzr @ r(Ws[l])
    return loss, pack(dWs, dbs, dw5, db5)


def random_params(seed, scale=1.0):
    """Parameters drawn like torch's Linear initialisation (times `scale`)."""
    rng = np.random.default_rng(seed)
    Ws, bs = [], []
    fan = NIN
    for _ in range(4):
        a = scale / np.sqrt(fan)
        Ws.append(rng.uniform(-a, a, size=(H, fan)))
        bs.append(rng.uniform(-a, a, size=H))
        fan = H
    a = scale / np.sqrt(H)
    return pack(Ws, bs, rng.uniform(-a, a, size=H), rng.uniform(-a, a))

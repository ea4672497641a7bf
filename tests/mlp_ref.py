"""float64 reference of the radiance MLP (test helper).

Same network as srt_amd/csrc/mlp.hip: 9 raw inputs -> [u, sin(2^k pi u),
cos(2^k pi u), k = 0..2] (63, +1 zero pad) -> 4 x (64, ReLU) -> 3; relative
L2 loss mean((y - t)^2 / (y^2 + 0.01)) with y detached in the denominator.
`emulate=True` rounds to BF16 exactly where the kernels do (encoded input,
weights, post-ReLU activations, back-propagated gradients); `emulate=False` is
the plain float64 network used to check the backward pass against finite
differences.
"""
import numpy as np

SHAPES = [(64, 63), (64, 64), (64, 64), (64, 64), (3, 64)]


def n_params():
    return sum(r * c + r for r, c in SHAPES)


def unpack(p):
    Ws, bs, o = [], [], 0
    for r, c in SHAPES:
        Ws.append(p[o:o + r * c].reshape(r, c)); o += r * c
        bs.append(p[o:o + r]); o += r
    return Ws, bs


def pack(Ws, bs):
    return np.concatenate([np.concatenate([W.reshape(-1), b]) for W, b in zip(Ws, bs)])


def bf16(x):
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def encode(u):
    u32 = np.asarray(u, np.float32)
    feats = [u32.astype(np.float64)]
    for k in range(3):
        x = (u32 * np.float32(2 ** k)) * np.float32(np.pi)
        feats += [np.sin(x.astype(np.float64)), np.cos(x.astype(np.float64))]
    return np.concatenate(feats, axis=1)            # (n, 63)


def forward(p, u, emulate=True):
    Ws, bs = unpack(p)
    r = bf16 if emulate else (lambda z: z)
    h = r(encode(u))
    hs = [h]
    for l in range(4):
        z = h @ r(Ws[l]).T + bs[l]
        h = r(np.maximum(z, 0.0))
        hs.append(h)
    y = h @ r(Ws[4]).T + bs[4]
    return y, hs


def loss(p, u, t, emulate=False):
    y, _ = forward(p, u, emulate)
    return np.mean((y - t) ** 2 / (y ** 2 + 0.01))


def backward(p, u, t, emulate=True):
    """Gradient of the loss (denominator treated as a constant)."""
    Ws, bs = unpack(p)
    r = bf16 if emulate else (lambda z: z)
    y, hs = forward(p, u, emulate)
    n = y.shape[0]
    g = r(2.0 * (y - t) / (y ** 2 + 0.01) / (3 * n))
    dWs, dbs = [None] * 5, [None] * 5
    for l in range(4, -1, -1):
        dWs[l] = g.T @ hs[l]
        dbs[l] = g.sum(0)
        if l == 0:
            break
        dh = g @ r(Ws[l])
        g = r(dh * (hs[l] > 0))
    return pack(dWs, dbs)


def loss_detached_grad_fd(p, u, t, eps=1e-6, idx=None):
    """Central differences of sum((y-t)^2 / (y0^2 + 0.01)) / 3n with y0 frozen."""
    y0, _ = forward(p, u, False)
    den = y0 ** 2 + 0.01

    def f(q):
        y, _ = forward(q, u, False)
        return np.mean((y - t) ** 2 / den)

    out = []
    for i in idx:
        a = p.copy(); a[i] += eps
        b = p.copy(); b[i] -= eps
        out.append((f(a) - f(b)) / (2 * eps))
    return np.array(out)


def init_params(seed):
    rng = np.random.default_rng(seed)
    Ws, bs = [], []
    for r, c in SHAPES:
        a = np.sqrt(6.0 / c)
        Ws.append(rng.uniform(-a, a, size=(r, c)))
        bs.append(rng.uniform(-0.1, 0.1, size=r))
    return pack(Ws, bs).astype(np.float32).astype(np.float64)

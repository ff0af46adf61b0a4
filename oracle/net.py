"""FP64 numpy restatement of the warm-start network (forward, analytic
backward, Adam) used as the checker of the GPU kernels.

TEST INFRASTRUCTURE ONLY (imported by tests/ and bench.py's checks; the
product never imports it).

The network has no counterpart in the reference: it produces the seed of the
FlatStart hook (src/RL-Biggs-Andrews.cpp:305-316), which replaces Y_0 of the
RL-Biggs-Andrews iteration.  Its definition (same as the library's):

    u  = raw / mean(raw)                               per sample
    e1 = relu(conv3(u; W1) + b1)                       1 -> 16, full size
    e2 = relu(conv3(pool(e1); W2) + b2)                16 -> 32, 1/2
    e3 = relu(conv3(pool(e2); W3) + b3)                32 -> 64, 1/4
    d3 = relu(conv3(e3; W4) + b4)                      64 -> 32, 1/4
    d2 = relu(conv3(up(d3) + e2; W5) + b5)             32 -> 16, 1/2
    y  = u + bh + (up(d2) + e1) . Wh                   1x1x1 head
    loss = mean((y - target / mean(raw))^2),   seed = max(y, 0) * mean(raw)

conv3 is the zero-padded 3x3x3 correlation out[v] = sum_t x[v + t] W[t]
(t = (dz+1)*9 + (dy+1)*3 + (dx+1)), pool the 2x2x2 mean, up the nearest 2x
upsampling.  Arrays are channels-last (nb, nz, ny, nx, c).

``bf16=True`` rounds (round-to-nearest-even) at the points where the GPU
stores BF16: the weights, bf16(u) as the first layer's input, and every
activation tensor after its layer (the head reads u in FP32).  Gradients are
exact FP64 gradients of the forward that was run.

Pinned by tests/test_net_oracle.py: conv3 against a brute-force loop, the
adjoint identities of conv3 / pool / up, and the analytic gradient against
central finite differences of the FP64 forward.
"""
from __future__ import annotations

import numpy as np

CIN = (1, 16, 32, 64, 32)
COUT = (16, 32, 64, 32, 16)
TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def layout():
    """(w_off, b_off, head_off, n) of the flat parameter vector."""
    o = 0
    w_off, b_off = [], []
    for ci, co in zip(CIN, COUT):
        w_off.append(o)
        o += 27 * ci * co
        b_off.append(o)
        o += co
    return w_off, b_off, o, o + 17


def unpack(p):
    w_off, b_off, h, _ = layout()
    W = [np.asarray(p[w_off[l]:w_off[l] + 27 * CIN[l] * COUT[l]], np.float64).reshape(27, CIN[l], COUT[l])
         for l in range(5)]
    b = [np.asarray(p[b_off[l]:b_off[l] + COUT[l]], np.float64) for l in range(5)]
    return W, b, np.asarray(p[h:h + 16], np.float64), float(p[h + 16])


def bf16(x):
    """Round to bfloat16 (nearest even) and return as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def shift(x, t):
    """S_t(x)[v] = x[v + t], zero outside the volume (per sample)."""
    dz, dy, dx = t
    _, nz, ny, nx, _ = x.shape
    out = np.zeros_like(x)
    zs, ys, xs = slice(max(0, -dz), nz - max(0, dz)), slice(max(0, -dy), ny - max(0, dy)), \
        slice(max(0, -dx), nx - max(0, dx))
    zd, yd, xd = slice(max(0, dz), nz + min(0, dz)), slice(max(0, dy), ny + min(0, dy)), \
        slice(max(0, dx), nx + min(0, dx))
    out[:, zs, ys, xs] = x[:, zd, yd, xd]
    return out


def conv3(x, W, b=None):
    out = sum(shift(x, t) @ W[i] for i, t in enumerate(TAPS))
    return out if b is None else out + b


def conv3_adjoint(g, W):
    """Adjoint of x -> conv3(x, W): sum_t S_{-t}(g W[t]^T)."""
    return sum(shift(g @ W[i].T, (-t[0], -t[1], -t[2])) for i, t in enumerate(TAPS))


def conv3_wgrad(x, g):
    """dW[t] = sum_v S_t(x)[v]^T g[v]."""
    gf = g.reshape(-1, g.shape[-1])
    return np.stack([shift(x, t).reshape(-1, x.shape[-1]).T @ gf for t in TAPS])


def pool(x):
    nb, nz, ny, nx, c = x.shape
    return x.reshape(nb, nz // 2, 2, ny // 2, 2, nx // 2, 2, c).mean(axis=(2, 4, 6))


def up(x):
    return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def down_sum(x):
    nb, nz, ny, nx, c = x.shape
    return x.reshape(nb, nz // 2, 2, ny // 2, 2, nx // 2, 2, c).sum(axis=(2, 4, 6))


def relu(x):
    return np.maximum(x, 0.0)


def forward(p, raw, bf=False):
    """Forward pass; raw (nb, nz, ny, nx).  Returns a dict of every tensor."""
    W, b, wh, bh = unpack(p)
    R = (lambda a: bf16(a)) if bf else (lambda a: a)
    if bf:
        W = [bf16(w) for w in W]
    raw = np.asarray(raw, np.float64)
    nb = raw.shape[0]
    mean = raw.reshape(nb, -1).mean(axis=1)
    inv_s = np.where(mean > 0, 1.0 / np.where(mean > 0, mean, 1.0), 1.0)
    if bf:
        inv_s = inv_s.astype(np.float32).astype(np.float64)
        u32 = (raw.astype(np.float32) * inv_s.astype(np.float32)[:, None, None, None]).astype(np.float64)
    else:
        u32 = raw * inv_s[:, None, None, None]
    u = u32[..., None]
    uin = R(u)
    T = {"inv_s": inv_s, "u": u, "uin": uin}
    T["e1"] = R(relu(conv3(uin, W[0], b[0])))
    T["p1"] = R(pool(T["e1"]))
    T["e2"] = R(relu(conv3(T["p1"], W[1], b[1])))
    T["p2"] = R(pool(T["e2"]))
    T["e3"] = R(relu(conv3(T["p2"], W[2], b[2])))
    T["d3"] = R(relu(conv3(T["e3"], W[3], b[3])))
    T["a2"] = R(up(T["d3"]) + T["e2"])
    T["d2"] = R(relu(conv3(T["a2"], W[4], b[4])))
    T["a1"] = up(T["d2"]) + T["e1"]
    T["y"] = u[..., 0] + bh + T["a1"] @ wh
    T["W"], T["b"], T["wh"], T["bh"] = W, b, wh, bh
    return T


def seed(p, raw, bf=False):
    T = forward(p, raw, bf)
    return np.maximum(T["y"], 0.0) / T["inv_s"][:, None, None, None]


def loss_grad(p, raw, target, bf=False):
    """(loss, flat gradient, tensors) of mean((y - target * inv_s)^2)."""
    T = forward(p, raw, bf)
    W, wh = T["W"], T["wh"]
    t = np.asarray(target, np.float64) * T["inv_s"][:, None, None, None]
    res = T["y"] - t
    n = res.size
    loss = float((res ** 2).sum() / n)
    dy = 2.0 * res / n
    G = {}
    G["dy"] = dy
    G["wh"] = T["a1"].reshape(-1, 16).T @ dy.ravel()
    G["bh"] = dy.sum()
    dA1 = dy[..., None] * wh
    G["dPre5"] = down_sum(dA1) * (T["d2"] > 0)
    G["W5"], G["b5"] = conv3_wgrad(T["a2"], G["dPre5"]), G["dPre5"].sum(axis=(0, 1, 2, 3))
    G["dA2"] = conv3_adjoint(G["dPre5"], W[4])
    G["dPre4"] = down_sum(G["dA2"]) * (T["d3"] > 0)
    G["W4"], G["b4"] = conv3_wgrad(T["e3"], G["dPre4"]), G["dPre4"].sum(axis=(0, 1, 2, 3))
    G["dE3"] = conv3_adjoint(G["dPre4"], W[3])
    G["dPre3"] = G["dE3"] * (T["e3"] > 0)
    G["W3"], G["b3"] = conv3_wgrad(T["p2"], G["dPre3"]), G["dPre3"].sum(axis=(0, 1, 2, 3))
    G["dP2"] = conv3_adjoint(G["dPre3"], W[2])
    G["dPre2"] = (G["dA2"] + up(G["dP2"]) / 8.0) * (T["e2"] > 0)
    G["W2"], G["b2"] = conv3_wgrad(T["p1"], G["dPre2"]), G["dPre2"].sum(axis=(0, 1, 2, 3))
    G["dP1"] = conv3_adjoint(G["dPre2"], W[1])
    G["dPre1"] = (dA1 + up(G["dP1"]) / 8.0) * (T["e1"] > 0)
    G["W1"], G["b1"] = conv3_wgrad(T["uin"], G["dPre1"]), G["dPre1"].sum(axis=(0, 1, 2, 3))
    w_off, b_off, h, npar = layout()
    flat = np.zeros(npar)
    for l in range(5):
        flat[w_off[l]:w_off[l] + 27 * CIN[l] * COUT[l]] = G["W%d" % (l + 1)].ravel()
        flat[b_off[l]:b_off[l] + COUT[l]] = G["b%d" % (l + 1)]
    flat[h:h + 16] = G["wh"]
    flat[h + 16] = G["bh"]
    return loss, flat, T, G


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t = step number from 1); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mhat = m / (1 - b1 ** t)
    vhat = v / (1 - b2 ** t)
    return p - lr * mhat / (np.sqrt(vhat) + eps), m, v


def he_init(seed_value):
    """A He-normal parameter vector (numpy's generator; zero biases and head)."""
    rng = np.random.default_rng(seed_value)
    w_off, b_off, h, npar = layout()
    p = np.zeros(npar)
    for l in range(5):
        p[w_off[l]:w_off[l] + 27 * CIN[l] * COUT[l]] = rng.standard_normal(27 * CIN[l] * COUT[l]) * np.sqrt(
            2.0 / (27 * CIN[l]))
    return p

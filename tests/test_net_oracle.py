"""Pins the FP64 oracle of the warm-start network (oracle/net.py) by checks
independent of its own formulas:

* conv3 against a brute-force 7-deep loop;
* adjoint identities <A x, g> = <x, A^T g> for conv3 (input and weight
  gradient), pool and up;
* the analytic gradient of the whole network against central finite
  differences of its FP64 forward, per layer, along random directions and on
  single coordinates (a dropped term, a transposed operand, a wrong tap flip
  or a missing 1/8 of the pooling adjoint each fail here);
* bf16 rounding on known bit patterns.
"""
import numpy as np
import pytest

from oracle import net as N


def test_conv3_matches_brute_force():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 3, 4, 5, 2))
    W = rng.standard_normal((27, 2, 3))
    b = rng.standard_normal(3)
    got = N.conv3(x, W, b)
    nb, nz, ny, nx, ci = x.shape
    ref = np.zeros((nb, nz, ny, nx, 3))
    for s in range(nb):
        for z in range(nz):
            for y in range(ny):
                for xx in range(nx):
                    acc = b.copy()
                    for t, (dz, dy, dx) in enumerate(N.TAPS):
                        zz, yy, xs = z + dz, y + dy, xx + dx
                        if 0 <= zz < nz and 0 <= yy < ny and 0 <= xs < nx:
                            for c in range(ci):
                                acc += x[s, zz, yy, xs, c] * W[t, c]
                    ref[s, z, y, xx] = acc
    assert np.abs(got - ref).max() < 1e-12


def test_adjoint_identities():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 4, 6, 8, 3))
    W = rng.standard_normal((27, 3, 5))
    g = rng.standard_normal((2, 4, 6, 8, 5))
    lhs = np.vdot(N.conv3(x, W), g)
    assert abs(lhs - np.vdot(x, N.conv3_adjoint(g, W))) < 1e-9 * abs(lhs)
    assert abs(lhs - np.vdot(W, N.conv3_wgrad(x, g))) < 1e-9 * abs(lhs)
    y = rng.standard_normal((2, 2, 3, 4, 3))
    assert abs(np.vdot(N.pool(x), y) - np.vdot(x, N.up(y) / 8.0)) < 1e-12 * x.size
    assert abs(np.vdot(N.up(y), x) - np.vdot(y, N.down_sum(x))) < 1e-12 * x.size


def _params(seed):
    rng = np.random.default_rng(seed)
    p = N.he_init(seed)
    w_off, b_off, h, n = N.layout()
    for l in range(5):
        p[b_off[l]:b_off[l] + N.COUT[l]] = 0.1 * rng.standard_normal(N.COUT[l])
    p[h:h + 17] = 0.3 * rng.standard_normal(17)
    return p


def _case(seed=2):
    rng = np.random.default_rng(seed)
    raw = rng.gamma(2.0, 50.0, (2, 4, 8, 8))
    tgt = raw * rng.uniform(0.5, 1.5, raw.shape)
    return raw, tgt


@pytest.mark.parametrize("layer", range(6))
def test_gradient_matches_finite_differences(layer):
    p = _params(3)
    raw, tgt = _case()
    loss, g, _, _ = N.loss_grad(p, raw, tgt)
    w_off, b_off, h, n = N.layout()
    if layer < 5:
        lo, hi = w_off[layer], b_off[layer] + N.COUT[layer]
    else:
        lo, hi = h, n
    rng = np.random.default_rng(10 + layer)

    def fd_err(d, an):
        # two step sizes: a ReLU kink inside one step spoils only that one
        errs = []
        for eps in (1e-6, 1e-7):
            fd = (N.loss_grad(p + eps * d, raw, tgt)[0] - N.loss_grad(p - eps * d, raw, tgt)[0]) / (2 * eps)
            errs.append(abs(fd - an))
        return min(errs)

    # random directions inside the layer
    for _ in range(3):
        d = np.zeros(n)
        d[lo:hi] = rng.standard_normal(hi - lo)
        an = float(g @ d)
        assert fd_err(d, an) <= 1e-5 * max(abs(an), 1e-8) + 1e-9, an
    # single coordinates (weights and biases)
    for i in list(rng.integers(lo, hi, 4)) + [hi - 1]:
        e = np.zeros(n)
        e[i] = 1.0
        assert fd_err(e, g[i]) <= 1e-5 * max(abs(g[i]), 1e-7) + 1e-9, (i, g[i])


def test_zero_head_is_identity():
    p = N.he_init(4)
    raw, _ = _case(5)
    s = N.seed(p, raw)
    assert np.abs(s - raw).max() < 1e-9 * raw.max()


def test_bf16_rounding_known_values():
    one = np.float32(1.0)
    ulp = 2.0 ** -7
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, -2.5, 3.0e-30], np.float64)
    got = N.bf16(vals)
    assert got[0] == 1.0
    assert got[1] == 1.0                    # tie -> even
    assert got[2] == 1.0 + 2 * ulp          # tie -> even (up)
    assert got[3] == 1.0 + ulp              # above the tie -> up
    assert got[4] == -2.5
    assert abs(got[5] - 3.0e-30) < 3.0e-30 * 2 ** -8
    assert one == 1.0

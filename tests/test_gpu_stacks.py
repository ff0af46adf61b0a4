"""Several stacks of one shape in one engine (cudaDecon_rl_create_stacks):
every pass runs once per iteration over all stacks, each stack keeping its
own acceleration factor.  Each stack's result and factors must equal,
bitwise, those of a one-stack engine on the same data (same kernels, same
per-row arithmetic, same per-stack dot-product partials summed in the same
order)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPE = (32, 64, 128)  # nz, ny, nx: 2048 rows per stack
NST = 5
ITERS = 8


def _setup(cdn):
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import gaussian_psf
    tab = make_otf(gaussian_psf(1.5, 3.0))
    ks = (np.float32(0.3), np.float32(0.3), np.float32(0.6))
    rng = np.random.default_rng(11)
    stacks = rng.gamma(2.0, 40.0, (NST,) + SHAPE).astype(np.float32)
    stacks[2] *= 3.0  # different scales give different factors per stack
    return tab, ks, stacks


def _single(cdn, tab, ks, raw):
    e = cdn.RLEngine(SHAPE, tab, ks)
    e.set_raw(raw)
    lam, _ = e.iterate(ITERS)
    X = e.get()
    e.close()
    return X, lam


def test_stacks_equal_single_engines_bitwise(cdn):
    tab, ks, stacks = _setup(cdn)
    print("multi-stack engine", NST, SHAPE, flush=True)
    m = cdn.RLEngine(SHAPE, tab, ks, stacks=NST)
    m.set_raw(stacks)
    lam, _ = m.iterate(ITERS)
    X = m.get()
    assert X.shape == (NST,) + SHAPE and lam.shape == (ITERS, NST)
    for s in range(NST):
        Xs, ls = _single(cdn, tab, ks, stacks[s])
        assert np.array_equal(X[s], Xs), s
        assert np.array_equal(lam[:, s], ls), (s, lam[:, s], ls)
        assert np.array_equal(m.get_stack(s), Xs)
    assert len(set(np.round(lam[-1], 12))) == NST  # the stacks really have their own factors
    # restage one stack: it restarts from its new data, the others from theirs
    new = np.random.default_rng(12).gamma(2.0, 60.0, SHAPE).astype(np.float32)
    m.set_raw_stack(3, new)
    lam2, _ = m.iterate(ITERS)
    Xn, ln = _single(cdn, tab, ks, new)
    assert np.array_equal(m.get_stack(3), Xn)
    assert np.array_equal(lam2[:, 3], ln)
    assert np.array_equal(m.get_stack(1), X[1])
    # continuation in two calls equals one call
    m.set_raw(stacks)
    m.iterate(3)
    m.iterate(ITERS - 3, k0=3)
    assert np.array_equal(m.get(), X)
    m.close()


def test_z256_stacks(cdn, orc):
    """z length 256 (the 6-wave z filter) with the kx scale of a wide stack:
    three stacks in one engine equal one-stack engines bitwise, and one of
    them the oracle."""
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import gaussian_psf
    tab = make_otf(gaussian_psf(1.5, 3.0))
    shape = (256, 16, 128)
    ks = (np.float32(0.03), np.float32(0.3), np.float32(0.13))
    rng = np.random.default_rng(21)
    stacks = rng.gamma(2.0, 40.0, (3,) + shape).astype(np.float32)
    stacks[1] *= 2.5
    print("z256 stacks", shape, flush=True)
    m = cdn.RLEngine(shape, tab, ks, stacks=3)
    m.set_raw(stacks)
    lam, _ = m.iterate(5)
    X = m.get()
    m.close()
    for s in range(3):
        e = cdn.RLEngine(shape, tab, ks)
        e.set_raw(stacks[s])
        ls, _ = e.iterate(5)
        assert np.array_equal(X[s], e.get()), s
        assert np.array_equal(lam[:, s], ls), s
        e.close()
    nz, ny, nx = shape
    H = orc.otf_interp(tab, (tab.shape[0], 1, tab.shape[1]), (nx, ny, nz), ks)
    ref, _ = orc.rl_ba(stacks[1].astype(np.float64), H, 5)
    assert np.abs(X[1] - ref).max() <= 1e-4 * np.abs(ref).max()


def test_stack_shape_and_single_stack_entry_points_are_checked(cdn):
    tab, ks, _ = _setup(cdn)
    with pytest.raises(RuntimeError, match="multiple"):
        cdn.RLEngine((7, 9, 128), tab, ks, stacks=2)  # 63 rows per stack: workgroups would straddle stacks
    m = cdn.RLEngine(SHAPE, tab, ks, stacks=2)
    with pytest.raises(RuntimeError, match="one-stack"):
        m.snapshot()
    with pytest.raises(RuntimeError, match="one-stack"):
        m.set_raw_block(np.ones((4, 4, 4), np.float32), (0, 0, 0), 1.0)
    with pytest.raises(RuntimeError):
        m.set_raw_stack(2, np.ones(SHAPE, np.float32))  # index out of range
    m.close()


def test_warm_start_seeds_every_stack(cdn):
    """The network runs once over all stacks (one sample each, normalised by
    its own mean); warm start + iterations per stack agree with one-stack
    engines to FP32 rounding (the per-sample means are summed in blocks
    whose count depends on the number of samples)."""
    from cudadecon_amd.warmstart import load_params
    tab, ks, stacks = _setup(cdn)
    p, _ = load_params()
    net = cdn.WarmNet(params=p)
    m = cdn.RLEngine(SHAPE, tab, ks, stacks=NST)
    m.set_raw(stacks)
    m.warm_start(net)
    m.iterate(4, seeded=True)
    X = m.get()
    m.close()
    for s in (0, 2, 4):
        e = cdn.RLEngine(SHAPE, tab, ks)
        e.set_raw(stacks[s])
        e.warm_start(net)
        e.iterate(4, seeded=True)
        Xs = e.get()
        e.close()
        err = np.abs(X[s] - Xs).max() / np.abs(Xs).max()
        assert err <= 1e-5, (s, err)
    net.close()

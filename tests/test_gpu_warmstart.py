"""What the trained warm-start network gains over plain RL-Biggs-Andrews.

Held-out stack (64 x 512 x 512, made as the 128 x 512 x 512 light-sheet
workload with seed 5; the weights were trained on seeds 100000+ and
validated on 200000).  Reference: the 50-iteration plain RL result.  The
network's seed followed by k iterations must be closer to it than plain RL
after the same k iterations (k = 0, 2, 5, 10), and reaching the residual of
the seed itself must take plain RL several times the network's time.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def curves(cdn):
    from cudadecon_amd import warmstart as ws
    from cudadecon_amd.otf import make_otf
    p, meta = ws.load_params()
    net = cdn.WarmNet(params=p)
    R, psf = ws.light_sheet_stack((64, 512, 512), 5)
    tab = make_otf(psf)
    res = ws.evaluate(net, R, tab, ws.kscale_for(R.shape, tab), iters=50, warm_iters=12)
    print({k: res[k] for k in ("iteration_ms", "net_ms", "plain", "warm")}, flush=True)
    return res


def test_warm_start_is_closer_at_equal_iterations(curves):
    for k in ("0", "2", "5", "10"):
        assert curves["warm"][k] <= 0.85 * curves["plain"][k], (k, curves["warm"][k], curves["plain"][k])


def test_warm_start_reaches_its_residual_faster(curves):
    r0 = curves["curves"]["warm"][0]
    plain = curves["curves"]["plain"]
    k_plain = next(k for k, v in enumerate(plain) if v <= r0)
    t_plain = k_plain * curves["iteration_ms"]
    assert k_plain >= 10
    assert curves["net_ms"] <= 0.5 * t_plain, (curves["net_ms"], t_plain)


def test_plain_curve_reaches_reference(curves):
    plain = curves["curves"]["plain"]
    assert plain[-1] == 0.0 and plain[0] > plain[10] > plain[40]


def test_warm_start_full_size_256x1024x1024(cdn):
    """At the config-3 inference size: a 256 x 1024 x 1024 light-sheet stack
    (seed 6, neither a training nor a validation seed).  The seed is finite
    and non-negative; warm start + k iterations is closer to the plain
    50-iteration result than plain RL after k, for k = 0, 2, 5, 10; and the
    network costs less than two RL-BA iterations."""
    from cudadecon_amd import warmstart as ws
    from cudadecon_amd.otf import make_otf
    p, _ = ws.load_params()
    net = cdn.WarmNet(params=p)
    R, psf = ws.light_sheet_stack(ws.EVAL_STACK, 6)
    tab = make_otf(psf)
    ks = ws.kscale_for(R.shape, tab)
    eng = cdn.RLEngine(R.shape, tab, ks)
    eng.set_raw(R)
    eng.warm_start(net)
    s = eng.buffer_stats(2)
    print("seed sum, min, max", s, flush=True)
    assert np.isfinite(s[0]) and s[1] >= 0.0 and np.isfinite(s[2])
    eng.close()
    res = ws.evaluate(net, R, tab, ks, iters=50, warm_iters=11)
    net.close()
    print({k: res[k] for k in ("iteration_ms", "net_ms", "plain", "warm")}, flush=True)
    for k in ("0", "2", "5", "10"):
        assert res["warm"][k] < res["plain"][k], (k, res["warm"][k], res["plain"][k])
    assert res["net_ms"] < 2 * res["iteration_ms"], (res["net_ms"], res["iteration_ms"])

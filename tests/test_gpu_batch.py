"""Batched RL-BA over stacks of uneven shape: every execution mode (one at a
time, round-robin streams, one captured hipGraph replayed) reproduces the
single-engine result bit for bit, including repeated graph replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(64, 96, 128), (96, 128, 64), (64, 64, 256), (70, 72, 150), (128, 64, 96)]


def test_batch_modes_match_single_engines(cdn, orc):
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import gaussian_psf, make_stack
    tab = make_otf(gaussian_psf(1.5, 3.0))
    ks, raws, refs = [], [], []
    for s in SHAPES:
        g = orc.init_geometry(s[2], s[1], s[0], 0.104, 0.2, 0.104, 0.2, 0.0, 0.0, 0, False,
                              (2 * tab.shape[1], tab.shape[0], 1))
        ks.append(g["kscale"])
        raw, _, _ = make_stack("beads_64x128x128", shape=s)
        R = np.maximum(raw.astype(np.float32) - 100.0, 0.0)
        raws.append(R)
        eng = cdn.RLEngine(s, tab, g["kscale"])
        eng.set_raw(R)
        eng.iterate(6)
        refs.append(eng.get())
        eng.close()
    b = cdn.Batch(SHAPES, tab, ks, nstreams=3)
    for i, R in enumerate(raws):
        b.set_raw(i, R)
    for mode in ("sequential", "streams", "graph", "graph"):
        print("batch mode", mode, flush=True)
        b.run(6, mode)
        for i in range(len(SHAPES)):
            assert np.array_equal(b.get(i), refs[i]), (mode, i)
    b.close()


def _singles(cdn, tab, shapes, ks, raws, iters):
    out = []
    for s, k, R in zip(shapes, ks, raws):
        eng = cdn.RLEngine(s, tab, k)
        eng.set_raw(R)
        eng.iterate(iters)
        out.append(eng.get())
        eng.close()
    return out


def test_graph_replay_after_restage_and_longer_runs(cdn, orc):
    """A replayed graph leaves every stack's result where the capture left
    it, even after set_raw moved the engines' X pointers (odd iteration
    count), and is captured again when a longer run regrew a buffer it had
    baked in."""
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import gaussian_psf, make_stack
    tab = make_otf(gaussian_psf(1.5, 3.0))
    shapes = SHAPES[:3]
    ks, raws = [], []
    for s in shapes:
        g = orc.init_geometry(s[2], s[1], s[0], 0.104, 0.2, 0.104, 0.2, 0.0, 0.0, 0, False,
                              (2 * tab.shape[1], tab.shape[0], 1))
        ks.append(g["kscale"])
        raw, _, _ = make_stack("beads_64x128x128", shape=s)
        raws.append(np.maximum(raw.astype(np.float32) - 100.0, 0.0))
    ref5 = _singles(cdn, tab, shapes, ks, raws, 5)
    ref6 = _singles(cdn, tab, shapes, ks, raws, 6)
    ref12 = _singles(cdn, tab, shapes, ks, raws, 12)
    b = cdn.Batch(shapes, tab, ks, nstreams=2)
    for i, R in enumerate(raws):
        b.set_raw(i, R)
    print("graph 5, restage, graph 5", flush=True)
    b.run(5, "graph")
    for i in range(len(shapes)):
        assert np.array_equal(b.get(i), ref5[i]), ("graph 5", i)
    for i, R in enumerate(raws):
        b.set_raw(i, R)
    b.run(5, "graph")
    for i in range(len(shapes)):
        assert np.array_equal(b.get(i), ref5[i]), ("graph 5 after set_raw", i)
    print("graph 6, sequential 12, graph 6", flush=True)
    b.run(6, "graph")
    b.run(12, "sequential")
    for i in range(len(shapes)):
        assert np.array_equal(b.get(i), ref12[i]), ("sequential 12", i)
    b.run(6, "graph")
    for i in range(len(shapes)):
        assert np.array_equal(b.get(i), ref6[i]), ("graph 6 after a longer run", i)
    b.close()

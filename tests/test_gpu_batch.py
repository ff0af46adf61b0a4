"""Batched RL-BA over stacks of uneven shape: every execution mode (one at a
time, round-robin streams, one captured hipGraph replayed, grouped launches
spanning stacks of one transform length, and those captured in a hipGraph)
reproduces the single-engine result bit for bit, including repeated graph
replays; the 48-stack workload (three channel PSFs) matches the FP64
oracle's digest."""
import hashlib
import os

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


# x lengths 128 / 64 / 256 / 150 / 96 (150 run-time compiled), a stack whose
# ny nz (72 x 70 = 5040) is not a multiple of two update workgroups' rows,
# repeated lengths so that groups hold several stacks
MIXED = [(64, 96, 128), (96, 128, 64), (64, 64, 256), (70, 72, 150), (128, 64, 96), (96, 96, 128),
         (64, 128, 64), (70, 72, 150), (64, 96, 256)]


def _tables():
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import BATCH_CHANNELS, gaussian_psf
    return [make_otf(gaussian_psf(sxy, sz)) for sxy, sz in BATCH_CHANNELS]


def test_grouped_modes_match_single_engines(cdn, orc):
    """Grouped launches (one per pass and transform length, spanning stacks
    of different shapes and OTF tables) give every stack bitwise the
    result and the factors of its own one-stack engine, in every mode."""
    from cudadecon_amd.synth import make_stack
    tabs = _tables()
    table_of = [i % 3 for i in range(len(MIXED))]
    ks, raws, refs, lams = [], [], [], []
    for i, s in enumerate(MIXED):
        t = tabs[table_of[i]]
        g = orc.init_geometry(s[2], s[1], s[0], 0.104, 0.2, 0.104, 0.2, 0.0, 0.0, 0, False,
                              (2 * t.shape[1], t.shape[0], 1))
        ks.append(g["kscale"])
        raw, _, _ = make_stack("beads_64x128x128", shape=s, seed=40 + i)
        R = np.maximum(raw.astype(np.float32) - 100.0, 0.0)
        raws.append(R)
        eng = cdn.RLEngine(s, t, g["kscale"])
        eng.set_raw(R)
        lam, _ = eng.iterate(7)
        refs.append(eng.get())
        lams.append(lam)
        eng.close()
    b = cdn.Batch(MIXED, tabs, ks, nstreams=3, table_of=table_of)
    # distinct nx, ny, nz: 5, 4, 4 (an x group splits after a stack whose rows
    # do not fill whole update workgroups)
    assert 3 * 5 + 4 * 4 + 2 * 4 + 2 <= b.launches() <= 3 * 7 + 4 * 4 + 2 * 4 + 2 < 9 * len(MIXED)
    for i, R in enumerate(raws):
        b.set_raw(i, R)
    for mode in ("grouped", "grouped_graph", "sequential", "grouped_graph", "graph", "grouped"):
        print("batch mode", mode, flush=True)
        b.run(7, mode)
        for i in range(len(MIXED)):
            assert np.array_equal(b.get(i), refs[i]), (mode, i)
            assert np.array_equal(b.lambdas(i, 7), lams[i]), (mode, i)
    print("persistent column passes", flush=True)
    b.set_grid_cap(64)
    b.run(7, "grouped")
    for i in range(len(MIXED)):
        assert np.array_equal(b.get(i), refs[i]), ("persistent", i)
    b.set_grid_cap(0)
    one = b.launches()
    for parts in (2, 3):
        b.set_partitions(parts)
        # the schedule is rebuilt at once: each partition launches its own
        # passes (and its own two factor launches)
        assert b.launches() >= one + parts - 1, (parts, b.launches(), one)
        for mode in ("grouped", "grouped_graph", "grouped_graph"):
            print("partitions", parts, mode, flush=True)
            b.run(7, mode)
            for i in range(len(MIXED)):
                assert np.array_equal(b.get(i), refs[i]), (parts, mode, i)
                assert np.array_equal(b.lambdas(i, 7), lams[i]), (parts, mode, i)
    with pytest.raises(RuntimeError):
        b.set_partitions(4)  # more than the batch's streams
    b.close()


def _sha(a):
    return hashlib.sha1(np.ascontiguousarray(a).view(np.uint8)).hexdigest()


def test_batch_48_uneven_stacks(cdn, orc):
    """The 48-stack workload: shapes drawn with seed 2 from z in {64..256},
    y, x in {256..1024}; per-stack seeded bead fields imaged through three
    channel PSFs; 10 iterations as one captured grouped schedule.  Three
    stacks (one per channel, the largest shape drawn among them) against
    the FP64 oracle's digest (tests/golden/make_batch_digest.py), every
    stack bitwise against a one-stack engine in the grouped-graph (one and
    four partitions) and per-stack graph modes."""
    from cudadecon_amd import synth
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "batch_48_digest.npz"))
    shapes = synth.batch_shapes()
    tabs = _tables()
    table_of = [synth.batch_channel(i)[0] for i in range(len(shapes))]
    ks = []
    for i, s in enumerate(shapes):
        t = tabs[table_of[i]]
        g = orc.init_geometry(s[2], s[1], s[0], synth.BATCH_DR, synth.BATCH_DZ, synth.BATCH_DR, synth.BATCH_DZ,
                              0.0, 0.0, 0, False, (2 * t.shape[1], t.shape[0], 1))
        ks.append(g["kscale"])
    b = cdn.Batch(shapes, tabs, ks, nstreams=8, table_of=table_of)
    single, digest_stacks = {}, [int(v) for v in d["stacks"]]
    for i, s in enumerate(shapes):
        raw = synth.batch_stack(i, s)
        if i in digest_stacks:
            assert tuple(d[f"s{i}_shape"]) == s and raw.dtype == np.uint16
            assert float(raw.astype(np.float64).sum()) == float(d[f"s{i}_raw_sum"]), i
        R = np.maximum(raw.astype(np.float32) - np.float32(100.0), np.float32(0.0))
        del raw
        b.set_raw(i, R)
        eng = cdn.RLEngine(s, tabs[table_of[i]], ks[i])
        eng.set_raw(R)
        eng.iterate(synth.BATCH_ITERS)
        single[i] = _sha(eng.get())
        eng.close()
        if i % 8 == 7:
            print(f"staged {i + 1} stacks", flush=True)
    for mode, parts in (("grouped_graph", 1), ("graph", 1), ("grouped_graph", 4)):
        b.set_partitions(parts)
        ms = b.run(synth.BATCH_ITERS, mode)
        print(f"mode {mode}, {parts} partition(s): {ms:.1f} ms for 10 iterations of all 48 stacks", flush=True)
        for i, s in enumerate(shapes):
            X = b.get(i)
            assert _sha(X) == single[i], (mode, parts, i)
            if i in digest_stacks and mode == "grouped_graph" and parts == 1:
                idx = d[f"s{i}_idx"]
                err = float(np.abs(X.reshape(-1)[idx] - d[f"s{i}_val"]).max() / d[f"s{i}_xmax"])
                lam_err = float(np.abs(b.lambdas(i, synth.BATCH_ITERS)[2:] - d[f"s{i}_lam"][2:]).max())
                print(f"stack {i} {s}: max |gpu - fp64| / max = {err:.2e}, factors {lam_err:.1e}", flush=True)
                assert err <= 1e-4 and lam_err <= 1e-4, (i, err, lam_err)
    b.close()

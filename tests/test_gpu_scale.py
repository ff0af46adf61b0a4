"""512 x 2048 x 2048 checks (, 2^31 voxels, ~52 GB resident in HBM).

The FP64 oracle cannot run at this size, so size-independent properties of
RL-BA are checked instead:

* circular-shift equivariance: a bead block on a constant background placed
  near the start of the volume and the same block placed near its far end
  (byte offsets beyond 2^32 in every array) must deconvolve to the same
  values (the reference's 32-bit index products overflow here);
* flux conservation: sum X = sum R for an OTF with unit DC gain;
* positivity and finiteness of X.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPE = (512, 2048, 2048)


def test_512x2048x2048_shift_equivariance_and_flux(cdn, orc):
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import gaussian_psf, make_stack
    nz, ny, nx = SHAPE
    tab = make_otf(gaussian_psf(1.5, 3.0))
    g = orc.init_geometry(nx, ny, nz, 0.104, 0.4, 0.104, 0.4, 0.0, 0.0, 0, False,
                          (2 * tab.shape[1], tab.shape[0], 1))
    raw, _, _ = make_stack("beads_64x128x128")  # 64 x 128 x 128 bead block
    block = np.maximum(raw.astype(np.float32) - 100.0, 0.0)
    bg = 1.0
    offA = (24, 24, 40)
    offB = (424, 1896, 1896)  # window ends at the last voxel (index 2^31 - 1); byte offsets > 2^32
    margin = 24
    win = (block.shape[0] + 2 * margin, block.shape[1] + 2 * margin, block.shape[2] + 2 * margin)
    eng = cdn.RLEngine(SHAPE, tab, g["kscale"])
    out = {}
    for name, off in (("A", offA), ("B", offB)):
        print("scale run", name, off, flush=True)
        eng.set_raw_block(block, off, background=bg)
        eng.timing(1)
        lam, ms = eng.iterate(10)
        pms, calls = eng.timing(0)
        print("  per-pass ms:", {k: round(pms[k] / max(calls[k], 1), 3) for k in pms}, flush=True)
        print(f"  10 iterations at {SHAPE}: {ms:.1f} ms ({ms / 10:.2f} ms/iter, "
              f"{np.prod(SHAPE) / (ms / 10 * 1e-3) / 1e9:.1f} Gvoxel/s)", flush=True)
        st = eng.stats()
        rs, xs = st["raw"][0], st["x"][0]
        assert abs(xs / rs - 1) < 1e-4, (xs, rs)
        assert st["x"][1] >= 0 and np.isfinite(st["x"][2])
        wo = tuple(o - margin for o in off)
        out[name] = (eng.get_block(win, wo), lam)
    eng.close()
    a, la = out["A"]
    b, lb = out["B"]
    assert np.abs(a - b).max() <= 1e-4 * np.abs(a).max()
    assert np.abs(la - lb).max() < 1e-4


def test_512x2048x2048_full_pipeline(cdn):
    """The 512 x 2048 x 2048 pipeline resident in HBM: camera offset
    subtraction of a uint16 stack, learned warm start, 10 RL-BA iterations,
    deskew (31.5 deg, dz 0.4, dx 0.104 -> 3675 wide) and an affine with a
    small rotation and scale (seed 4).  Checks stage results (finite,
    non-negative, bounded by the stack) and prints the stage times."""
    from cudadecon_amd import pipeline, warmstart
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import make_stack
    nz, ny, nx = SHAPE
    print("pipeline", SHAPE, flush=True)
    blk, psf, _ = make_stack("lightsheet_128x512x512", shape=(64, 128, 128), seed=4)
    raw = np.full(SHAPE, 200, np.uint16)
    for off in ((100, 300, 500), (300, 1500, 900), (420, 100, 1800)):
        raw[off[0]:off[0] + 64, off[1]:off[1] + 128, off[2]:off[2] + 128] = blk
    tab = make_otf(psf)
    p, _ = warmstart.load_params()
    net = cdn.WarmNet(params=p)
    eng = cdn.RLEngine(SHAPE, tab, warmstart.kscale_for(SHAPE, tab, dz=0.4))
    t = pipeline.run(eng, raw, net)
    print(t, flush=True)
    assert t["deskewed_nx"] == 3675 and t["output_zyx"] == [512, 2048, 3675]
    st = eng.stats()
    assert st["x"][1] >= 0 and np.isfinite(st["x"][2]) and st["x"][2] <= 100 * st["raw"][2]
    s = eng.post_stats()
    assert np.isfinite(s[0]) and s[1] >= 0 and s[2] <= st["x"][2] * (1 + 1e-6)
    eng.close()
    net.close()

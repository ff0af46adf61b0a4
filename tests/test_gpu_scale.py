"""cfg4-size checks (512 x 2048 x 2048, 2^31 voxels, ~52 GB resident in HBM).

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


def test_cfg4_size_shift_equivariance_and_flux(cdn, orc):
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import gaussian_psf, make_stack
    nz, ny, nx = SHAPE
    tab = make_otf(gaussian_psf(1.5, 3.0))
    g = orc.init_geometry(nx, ny, nz, 0.104, 0.4, 0.104, 0.4, 0.0, 0.0, 0, False,
                          (2 * tab.shape[1], tab.shape[0], 1))
    raw, _, _ = make_stack("cfg0")  # 64 x 128 x 128 bead block
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

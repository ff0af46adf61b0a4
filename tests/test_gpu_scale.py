"""512 x 2048 x 2048 checks (, 2^31 voxels, ~52 GB resident in HBM).

The FP64 oracle cannot run at this size, so size-independent properties of
RL-BA are checked instead:

* circular-shift equivariance: a bead block on a constant background placed
  near the start of the volume and the same block placed near its far end
  (byte offsets beyond 2^32 in every array) must deconvolve to the same
  values (the reference's 32-bit index products overflow here);
* flux conservation: sum X = sum R for an OTF with unit DC gain;
* positivity and finiteness of X;
* the full pipeline's deskew and affine stages against the oracle's
  formulas on blocks read back from the device.
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


def _affine_samples(eng, orc, M, shape, xmax, n, seed):
    """Affine result at n random output voxels against the oracle's texture
    sampling (orc.affine) of each voxel's 2 x 2 x 2 source cell of the
    deskewed volume, read back by blocks.  The source coordinates are the
    reference's float expression; the GPU contracts it into FMAs, which can
    move a coordinate by a few ulps (2.4e-4 voxel at 3675), hence a
    tolerance of 1e-3 of the cell's value range plus 1e-5 of max |X|.
    Returns the worst error / tolerance and the number of samples inside."""
    nz, ny, nx = shape
    rng = np.random.default_rng(seed)
    m = np.asarray(M, np.float32).reshape(-1)
    f = np.float32
    worst, inside = 0.0, 0
    for _ in range(n):
        x, y, z = int(rng.integers(0, nx)), int(rng.integers(0, ny)), int(rng.integers(0, nz))
        u, v, w = f(x), f(y), f(z)
        t = [m[4 * r] * u + m[4 * r + 1] * v + m[4 * r + 2] * w + m[4 * r + 3] + f(0.5) for r in range(3)]
        b = [f(tc - f(0.5)) for tc in t]
        fl = [int(np.floor(bc)) for bc in b]
        got = eng.post_get_block(1, (1, 1, 1), (z, y, x))[0, 0, 0]
        if not (0 <= fl[0] < nx - 1 and 0 <= fl[1] < ny - 1 and 0 <= fl[2] < nz - 1):
            if fl[0] < -1 or fl[0] > nx or fl[1] < -1 or fl[1] > ny or fl[2] < -1 or fl[2] > nz:
                assert got == 0.0, (x, y, z)  # every texel outside: border 0
            continue
        cell = eng.post_get_block(0, (2, 2, 2), (fl[2], fl[1], fl[0]))
        mc = np.zeros(16, np.float32)
        mc[3], mc[7], mc[11] = b[0] - f(fl[0]), b[1] - f(fl[1]), b[2] - f(fl[2])  # exact (Sterbenz)
        want = orc.affine(cell, mc)[0, 0, 0]
        tol = 1e-3 * float(cell.max() - cell.min()) + 1e-5 * xmax
        worst = max(worst, abs(float(got) - float(want)) / tol)
        inside += 1
    return worst, inside


def test_512x2048x2048_full_pipeline(cdn, orc):
    """The 512 x 2048 x 2048 pipeline resident in HBM: camera offset
    subtraction of a uint16 stack, learned warm start, 10 RL-BA iterations,
    deskew (31.5 deg, dz 0.4, dx 0.104 -> 3675 wide) and an affine with a
    small rotation and scale (seed 4).  Checks stage results (finite,
    non-negative, bounded by the stack), the deskew of the deconvolved
    volume against the oracle's deskew (whole x-z slabs at 8 random y, read
    back by blocks) and the affine against the oracle's texture sampling at
    3,000 random output voxels, and prints the stage times."""
    import math

    from cudadecon_amd.pipeline import small_affine
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
    xmax = st["x"][2]
    # deskew (src/geometryTransform.cu:4-48) of X, slab by slab: a deskew
    # only moves voxels along x within their own (z, y) row
    factor = math.cos(31.5 * math.pi / 180.0) * float(np.float32(0.4)) / float(np.float32(0.104))
    rng = np.random.default_rng(44)
    worst = 0.0
    for y in sorted(int(v) for v in rng.choice(ny, 8, replace=False)):
        X = eng.get_block((nz, 1, nx), (0, y, 0))
        D = eng.post_get_block(0, (nz, 1, 3675), (0, y, 0))
        want = orc.deskew(X, factor, 3675)
        worst = max(worst, float(np.abs(D - want).max()) / xmax)
    print(f"deskew: max |gpu - oracle| / max X = {worst:.2e}", flush=True)
    assert worst <= 1e-5
    M, _ = small_affine((nz, ny, 3675), seed=4)
    ratio, inside = _affine_samples(eng, orc, M, (nz, ny, 3675), xmax, 3000, 45)
    print(f"affine: {inside} voxels inside, worst error / tolerance = {ratio:.3f}", flush=True)
    assert inside > 2500 and ratio <= 1.0
    eng.close()
    net.close()

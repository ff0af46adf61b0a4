"""Autotuner of the FFT launch shapes (cudaDecon_autotune).

Variants only regroup lines into workgroups: column-pass variants (columns
per tile, threads per line, non-temporal loads, workgroup-to-XCD
renumbering) leave every result bitwise unchanged; an
x-pass variant also regroups the per-block partial sums of the acceleration
factor's dot products (FP64 sums in another order), so RL results agree to
FP32 rounding.  Variants are forced through the tuner's decision cache in a
private cache directory, then a real tuning run is checked the same way."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPE = (32, 96, 192)   # nz, ny, nx: lengths 32 (run-time compiled), 96, 192


def _rl(cdn, R, tab, n=6):
    eng = cdn.RLEngine(R.shape, tab, (np.float32(0.3), np.float32(0.3), np.float32(0.6)))
    eng.set_raw(R)
    lam, _ = eng.iterate(n)
    X = eng.get()
    eng.close()
    return X, lam


@pytest.fixture(scope="module")
def case(cdn):
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import gaussian_psf
    R = np.random.default_rng(3).uniform(0, 200, SHAPE).astype(np.float32)
    tab = make_otf(gaussian_psf(1.5, 3.0))
    cdn.autotune_clear()
    X0, lam0 = _rl(cdn, R, tab)
    return R, tab, X0, lam0


def _force(cdn, tmp_path, lines, shape=SHAPE):
    d = tmp_path / "cache"
    (d / "cudadecon_amd").mkdir(parents=True)
    (d / "cudadecon_amd" / "tune_gfx950_v2.txt").write_text("".join(l + "\n" for l in lines))
    old = os.environ.get("XDG_CACHE_HOME")
    os.environ["XDG_CACHE_HOME"] = str(d)
    try:
        return cdn.autotune([shape], reps=1, use_cache=True)
    finally:
        if old is None:
            del os.environ["XDG_CACHE_HOME"]
        else:
            os.environ["XDG_CACHE_HOME"] = old


def test_forced_column_variants_are_bitwise(cdn, case, tmp_path):
    R, tab, X0, lam0 = case
    g = "x".join(map(str, SHAPE))
    rep = _force(cdn, tmp_path, [f"192 x {g} 0 0 -1 -1 0 0 0", f"96 y {g} 8 0 1 0 0 0 0", f"32 z {g} 128 0 1 2 1 2 4"])
    print(rep, flush=True)
    assert cdn.fft_variant(96) == dict(cols=8, xthreads=0, col_nt_load=1, xcd_remap=0, col_tpt=0, zf_waves=0, zf_rows=0)
    # z length 32: one thread per line, 128 columns per tile (all 97 in one),
    # z filter built for 2 waves per SIMD, 4 OTF rows staged in LDS
    assert cdn.fft_variant(32) == dict(cols=128, xthreads=0, col_nt_load=1, xcd_remap=2, col_tpt=1, zf_waves=2, zf_rows=4)
    assert cdn.fft_variant(192) is None
    X1, lam1 = _rl(cdn, R, tab)
    assert np.array_equal(X1, X0)
    assert np.array_equal(lam1, lam0)
    cdn.autotune_clear()


# the wide column tiles the tuner picks at 512 x 2048 x 2048, on small grids:
# (nz, ny, nx), the variant's cache line, the length it applies to
WIDE = [((512, 32, 64), "512 z {g} 32 0 -1 -1 16 0 0", 512),
        ((32, 2048, 64), "2048 y {g} 16 0 -1 2 64 0 0", 2048)]


@pytest.mark.parametrize("shape,line,n", WIDE)
def test_forced_wide_tiles_are_bitwise(cdn, tmp_path, shape, line, n):
    """Twice the columns per tile at half the threads per line (32 elements
    per thread): the 512-point z filter (rotationally averaged OTF kind) and
    the 2048-point y transforms give the built-in shapes' results bit for
    bit."""
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import gaussian_psf
    R = np.random.default_rng(4).uniform(0, 200, shape).astype(np.float32)
    tab = make_otf(gaussian_psf(1.5, 3.0))
    cdn.autotune_clear()
    X0, lam0 = _rl(cdn, R, tab, n=4)
    g = "x".join(map(str, shape))
    roles = {shape[2]: "x", shape[1]: "y", shape[0]: "z"}
    lines = [line.format(g=g)] + [f"{m} {r} {g} 0 0 -1 -1 0 0 0" for m, r in roles.items() if m != n]
    _force(cdn, tmp_path, lines, shape)
    v = cdn.fft_variant(n)
    assert v is not None and v["col_tpt"] > 0 and v["cols"] > 0, v
    X1, lam1 = _rl(cdn, R, tab, n=4)
    assert np.array_equal(X1, X0)
    assert np.array_equal(lam1, lam0)
    cdn.autotune_clear()


def test_forced_x_variant_agrees_to_rounding(cdn, case, tmp_path):
    R, tab, X0, lam0 = case
    g = "x".join(map(str, SHAPE))
    _force(cdn, tmp_path, [f"192 x {g} 0 256 -1 -1 0 0 0", f"96 y {g} 0 0 -1 -1 0 0 0", f"32 z {g} 0 0 -1 -1 0 0 0"])
    assert cdn.fft_variant(192) == dict(cols=0, xthreads=256, col_nt_load=-1, xcd_remap=-1, col_tpt=0, zf_waves=0, zf_rows=0)
    X1, lam1 = _rl(cdn, R, tab)
    assert np.abs(X1 - X0).max() <= 1e-5 * np.abs(X0).max()
    assert np.abs(lam1 - lam0).max() <= 1e-9
    cdn.autotune_clear()


def test_tuning_run_keeps_results(cdn, case, tmp_path):
    R, tab, X0, lam0 = case
    old = os.environ.get("XDG_CACHE_HOME")
    os.environ["XDG_CACHE_HOME"] = str(tmp_path / "fresh")
    try:
        rep = cdn.autotune([SHAPE, (64, 128, 128)], reps=3, use_cache=True)
    finally:
        if old is None:
            del os.environ["XDG_CACHE_HOME"]
        else:
            os.environ["XDG_CACHE_HOME"] = old
    print(rep, flush=True)
    ns = sorted(e["n"] for e in rep["lengths"])
    assert ns == [32, 64, 96, 128, 192]
    for e in rep["lengths"]:
        assert "chosen" in e
    X1, lam1 = _rl(cdn, R, tab)
    assert np.abs(X1 - X0).max() <= 1e-5 * np.abs(X0).max()
    cdn.autotune_clear()

"""GPU parity of the RL-Biggs-Andrews path against the FP64 oracle.

Tolerance (the stated per-config bar): max |X_gpu - X_oracle| <= 1e-4 *
max |X_oracle| after the stated iteration count; acceleration factors agree
to 1e-4 absolute.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _preprocessed(raw_u16, bg=100.0):
    return np.maximum(raw_u16.astype(np.float64) - bg, 0.0)


def _otf_and_geom(orc, shape, sxy=1.5, sz=3.0, dr=0.104, dz=0.2):
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import gaussian_psf
    tab = make_otf(gaussian_psf(sxy, sz))
    g = orc.init_geometry(shape[2], shape[1], shape[0], dr, dz, dr, dz, 0.0, 0.0, 0, False,
                          (2 * tab.shape[1], tab.shape[0], 1))
    return tab, g


def _relmax(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("shape,iters", [((64, 128, 128), 10), ((64, 96, 810), 6), ((96, 64, 64), 12),
                                         ((70, 72, 150), 8), ((256, 16, 1024), 6)],
                         ids=["beads_64x128x128", "x810", "z96", "runtime-compiled", "z256"])
def test_engine_matches_oracle(cdn, orc, shape, iters):
    print("engine", shape, iters, flush=True)
    from cudadecon_amd.synth import make_stack
    raw, _, _ = make_stack("beads_64x128x128", shape=shape)
    R = _preprocessed(raw)
    tab, g = _otf_and_geom(orc, shape)
    H = orc.otf_interp(tab, g["otf_dims"], g["fft_dims"], g["kscale"])
    ref, lam_ref = orc.rl_ba(R, H, iters)
    eng = cdn.RLEngine(shape, tab, g["kscale"])
    eng.set_raw(R.astype(np.float32))
    lam, _ = eng.iterate(iters)
    got = eng.get()
    eng.close()
    assert _relmax(got, ref) <= TOL
    assert np.abs(lam - lam_ref).max() < 1e-4


def test_engine_continuation_equals_single_run(cdn, orc):
    shape = (64, 128, 128)
    from cudadecon_amd.synth import make_stack
    raw, _, _ = make_stack("beads_64x128x128", shape=shape)
    R = _preprocessed(raw).astype(np.float32)
    tab, g = _otf_and_geom(orc, shape)
    eng = cdn.RLEngine(shape, tab, g["kscale"])
    eng.set_raw(R)
    eng.iterate(7)
    a = eng.get()
    eng.iterate(3, k0=0)
    eng.iterate(4, k0=3)
    b = eng.get()
    eng.close()
    assert np.array_equal(a, b)


def test_seeded_start_matches_oracle(cdn, orc):
    shape = (64, 128, 128)
    from cudadecon_amd.synth import make_stack
    raw, _, _ = make_stack("beads_64x128x128", shape=shape)
    R = _preprocessed(raw)
    seed = np.full(shape, R.mean())
    tab, g = _otf_and_geom(orc, shape)
    H = orc.otf_interp(tab, g["otf_dims"], g["fft_dims"], g["kscale"])
    ref, _ = orc.rl_ba(R, H, 5, seed=seed)
    eng = cdn.RLEngine(shape, tab, g["kscale"])
    eng.set_raw(R.astype(np.float32))
    eng.set_seed(seed.astype(np.float32))
    eng.iterate(5, seeded=True)
    got = eng.get()
    eng.close()
    assert _relmax(got, ref) <= TOL


def _write_otf(tmp_path, tab):
    from cudadecon_amd.otf import write_otf_tiff
    p = str(tmp_path / "otf.tif")
    write_otf_tiff(p, tab)
    return p


def _run_both(cdn, orc, tmp_path, raw, tab, dr, dz, angle, rot, iters, **kw):
    nz, ny, nx = raw.shape
    p = _write_otf(tmp_path, tab)
    skewed = kw.get("skewed_decon", False)
    cdn.rl_cleanup()
    cdn.rl_init(raw.shape, p, dzdata=dz, dxdata=dr, dzpsf=dz, dxpsf=dr, deskew=angle, rotate=rot,
                skewed_decon=skewed)
    g = orc.init_geometry(nx, ny, nz, dr, dz, dr, dz, angle, rot, 0, skewed, (2 * tab.shape[1], tab.shape[0], 1))
    assert cdn.output_shape() == (g["out_dims"][2], g["out_dims"][1], g["out_dims"][0])
    sess = orc.Session(g, tab)
    ref, _, _ = orc.rl_interface(sess, raw, 100.0, False, False, iters, 0, **kw)
    got = cdn.rl_decon(raw, background=100.0, n_iters=iters, **kw)
    cdn.rl_cleanup()
    return got, ref


def test_rl_interface_beads_64x128x128(cdn, orc, tmp_path):
    from cudadecon_amd.synth import make_stack
    raw, _, c = make_stack("beads_64x128x128")
    tab, _ = _otf_and_geom(orc, raw.shape)
    got, ref = _run_both(cdn, orc, tmp_path, raw, tab, 0.104, 0.2, 0.0, 0.0, c["iters"])
    assert got.shape == ref.shape
    assert _relmax(got, ref) <= TOL


def test_rl_interface_deskew_rotate_small(cdn, orc, tmp_path):
    # the light-sheet geometry on a reduced stack: 64 x 128 x 100 raw, deskew 31.5 deg,
    # dz 0.3 um -> deskewed width findOptimalDimension(100 + 157) = 256
    from cudadecon_amd.synth import make_stack
    raw, _, _ = make_stack("lightsheet_128x512x512", shape=(64, 128, 100))
    tab, _ = _otf_and_geom(orc, raw.shape, 1.3, 4.0, 0.104, 0.3)
    nz, ny, nx = raw.shape
    g = orc.init_geometry(nx, ny, nz, 0.104, 0.3, 0.104, 0.3, 31.5, 31.5, 0, False,
                          (2 * tab.shape[1], tab.shape[0], 1))
    if not cdn.lib().cudaDecon_fft_supported(g["deskewed_nx"]):
        pytest.skip(f"deskewed width {g['deskewed_nx']} not compiled in")
    got, ref = _run_both(cdn, orc, tmp_path, raw, tab, 0.104, 0.3, 31.5, 31.5, 4)
    assert _relmax(got, ref) <= TOL


def test_rl_interface_deskew_to_runtime_compiled_width(cdn, orc, tmp_path):
    # 64 x 128 x 100 raw at -32.8 deg: deskewed width findOptimalDimension(255) = 252
    from cudadecon_amd.synth import make_stack
    raw, _, _ = make_stack("lightsheet_128x512x512", shape=(64, 128, 100))
    tab, _ = _otf_and_geom(orc, raw.shape, 1.3, 4.0, 0.104, 0.3)
    got, ref = _run_both(cdn, orc, tmp_path, raw, tab, 0.104, 0.3, -32.8, 0.0, 5)
    assert got.shape == ref.shape == (64, 128, 252)
    assert _relmax(got, ref) <= TOL


def test_rl_interface_preprocessing_only_with_rotation(cdn, orc, tmp_path):
    from cudadecon_amd.synth import make_stack
    raw, _, _ = make_stack("beads_64x128x128", shape=(64, 64, 64))
    tab, _ = _otf_and_geom(orc, raw.shape)
    got, ref = _run_both(cdn, orc, tmp_path, raw, tab, 0.104, 0.2, 0.0, 20.0, 0, napodize=5)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-4 * np.abs(ref).max()


# RL_interface options against the oracle's restatement of RichardsonLucy_GPU
# (src/RL-Biggs-Andrews.cpp:56-450): skewed decon (iterate on the raw grid,
# deskew after), z blend, the doubled reversed stack, the deskewed raw output,
# and the bleach rescale, whose first call fixes the reference intensity.
# 64 x 64 x 100 raw at 31.5 deg, dz 0.3: deskewed width 256 (built in).
OPTION_CASES = {
    "skewed_decon": dict(skewed=True, kw=dict()),
    "z_blend_4": dict(skewed=False, kw=dict(nzblend=4)),
    "dup_rev_stack": dict(skewed=False, kw=dict(dup=True)),
    "save_deskewed": dict(skewed=False, kw=dict(save=True)),
    "rescale_second_call": dict(skewed=False, kw=dict(rescale=True)),
    "shift_and_pad": dict(skewed=False, kw=dict(shift=7, pad=50.0)),
    "output_width_300": dict(skewed=False, kw=dict(width=300)),
}


@pytest.mark.parametrize("case", list(OPTION_CASES), ids=list(OPTION_CASES))
def test_rl_interface_options(cdn, orc, tmp_path, case):
    print("RL_interface option", case, flush=True)
    from cudadecon_amd.synth import make_stack
    raw, _, _ = make_stack("lightsheet_128x512x512", shape=(64, 64, 100))
    tab, _ = _otf_and_geom(orc, raw.shape, 1.3, 4.0, 0.104, 0.3)
    nz, ny, nx = raw.shape
    c = OPTION_CASES[case]
    skewed, kw = c["skewed"], c["kw"]
    p = _write_otf(tmp_path, tab)
    cdn.rl_cleanup()
    width = kw.get("width", 0)
    cdn.rl_init(raw.shape, p, dzdata=0.3, dxdata=0.104, dzpsf=0.3, dxpsf=0.104, deskew=31.5, rotate=0.0,
                width=width, skewed_decon=skewed)
    g = orc.init_geometry(nx, ny, nz, 0.104, 0.3, 0.104, 0.3, 31.5, 0.0, width, skewed,
                          (2 * tab.shape[1], tab.shape[0], 1))
    assert cdn.output_shape() == (g["out_dims"][2], g["out_dims"][1], g["out_dims"][0])
    sess = orc.Session(g, tab)
    iters = 4
    save = kw.get("save", False)
    resc = kw.get("rescale", False)
    pad = kw.get("pad", 0.0)
    shift = kw.get("shift", 0)
    args = dict(napodize=0, nzblend=kw.get("nzblend", 0), pad_val=pad, dup_rev_stack=kw.get("dup", False),
                skewed_decon=skewed)
    gargs = dict(napodize=0, nz_blend=kw.get("nzblend", 0), pad_val=pad, dup_reversed_z=kw.get("dup", False),
                 skewed_decon=skewed, shift=shift)
    calls = [raw, np.clip(raw.astype(np.int32) * 3 // 4 + 25, 0, 65535).astype(np.uint16)] if resc else [raw]
    for im in calls:
        ref, ref_desk, _ = orc.rl_interface(sess, im, 100.0, resc, save, iters, shift, **args)
        got = cdn.rl_decon(im, background=100.0, n_iters=iters, save_deskewed=save, rescale=resc, **gargs)
        if save:
            got, got_desk = got
            assert got_desk.shape == ref_desk.shape
            assert np.abs(got_desk - ref_desk).max() <= 1e-5 * np.abs(ref_desk).max()
        assert got.shape == ref.shape
        assert _relmax(got, ref) <= TOL, (case, _relmax(got, ref))
    cdn.rl_cleanup()


def test_deskew_interface(cdn, orc):
    rng = np.random.default_rng(11)
    im = rng.uniform(0, 1000, (32, 40, 100)).astype(np.float32)
    for angle, width, shift, pad in [(31.5, 160, 0, 0.0), (-32.8, 180, 3, 7.0), (0.0, 120, -2, 1.0)]:
        got = cdn.deskew(im, angle, 0.3, 0.104, width=width, shift=shift, pad_val=pad)
        ref = orc.deskew_interface(im, 0.3, 0.104, angle, width, shift, pad)
        assert got.shape == ref.shape
        assert np.abs(got - ref).max() <= 1e-4 * np.abs(ref).max()


def test_deskew_large_buffers_round_trip_exactly(cdn):
    """Host buffers of 16 MB and more cross through two 32 MB pinned staging
    buffers (src/RL-Biggs-Andrews.cpp:724-773 copies them with cudaMemcpy).
    A zero-angle deskew is a copy with the last column padded (xin = x +
    extraShift exactly, interpolation weight 0), so the result is bitwise
    the input: 72 MB each way is three chunks, the last partial, which takes
    both pipelines through buffer reuse.  A small call in between goes
    straight through the runtime; the large call after it reuses the
    staging buffers."""
    rng = np.random.default_rng(21)
    big = rng.uniform(-1000, 1000, (67, 263, 1021)).astype(np.float32)
    small = rng.uniform(-1000, 1000, (5, 31, 77)).astype(np.float32)
    assert big.nbytes > 2 * (32 << 20) and big.nbytes % (32 << 20) != 0
    for im, shift in [(big, 0), (small, 0), (big, 2)]:
        nx = im.shape[2]
        got = cdn.deskew(im, 0.0, 0.3, 0.104, shift=shift, pad_val=-7.0)
        assert got.shape == im.shape
        assert np.array_equal(got[..., : nx - 1 - shift], im[..., shift : nx - 1])
        assert np.all(got[..., nx - 1 - shift :] == np.float32(-7.0))


def test_affine_interfaces(cdn, orc):
    rng = np.random.default_rng(12)
    im = rng.uniform(0, 1000, (24, 48, 64)).astype(np.float32)
    th = np.deg2rad(4.0)
    m = np.eye(4, dtype=np.float32)
    m[0, 0], m[0, 1], m[1, 0], m[1, 1] = np.cos(th) * 1.03, -np.sin(th), np.sin(th), np.cos(th) * 0.97
    m[0, 3], m[1, 3], m[2, 3] = 2.5, -1.25, 0.5
    got = cdn.affine(im, m)
    ref = orc.affine(im, m)
    assert np.abs(got - ref).max() <= 1e-4 * np.abs(ref).max()
    got = cdn.affine_ra(im, m, 0.104, 0.104, 0.3)
    ref = orc.affine_ra(im, 0.104, 0.104, 0.3, m)
    assert np.abs(got - ref).max() <= 1e-4 * np.abs(ref).max()


def test_rl_interface_lightsheet_full_size_against_fp64_digest(cdn, tmp_path):
    """The light-sheet workload at full size: 128 x 512 x 512 raw, deskew 31.5 deg -> 128 x 512 x 810,
    15 RL-BA iterations, rotation -> 192 x 512 x 791, against the FP64 oracle's
    result sampled at 40,000 seeded voxels (tests/golden/make_lightsheet_digest.py)."""
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import make_stack
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "lightsheet_128x512x512_digest.npz"))
    raw, psf, c = make_stack("lightsheet_128x512x512")
    # the regenerated input is the one the digest was made from (a changed
    # random stream would otherwise read as a parity failure)
    assert raw.shape == (128, 512, 512) and raw.dtype == np.uint16
    assert float(raw.astype(np.float64).sum()) == float(d["raw_sum"])
    tab = make_otf(psf)
    p = _write_otf(tmp_path, tab)
    cdn.rl_cleanup()
    cdn.rl_init(raw.shape, p, dzdata=c["dz"], dxdata=c["dr"], dzpsf=c["dz"], dxpsf=c["dr"], deskew=c["angle"],
                rotate=c["rot"])
    got = cdn.rl_decon(raw, background=100.0, n_iters=c["iters"])
    cdn.rl_cleanup()
    assert got.shape == tuple(d["shape"]) == (192, 512, 791)
    err = float(np.abs(got.reshape(-1)[d["idx"]] - d["val"]).max() / d["xmax"])
    print(f"light-sheet max |gpu - fp64| / max = {err:.2e}", flush=True)
    assert err <= TOL


def test_device_pipeline_matches_host_entry_points(cdn, orc):
    """set_raw_u16 / deskew / affine on the engine's buffers give bitwise the
    results of the float path and of Deskew_interface / Affine_interface."""
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.pipeline import small_affine
    from cudadecon_amd.synth import make_stack
    raw, psf, c = make_stack("beads_64x128x128")
    tab = make_otf(psf)
    g = orc.init_geometry(128, 128, 64, 0.104, 0.2, 0.104, 0.2, 0.0, 0.0, 0, False, (2 * tab.shape[1], tab.shape[0], 1))
    eng = cdn.RLEngine(raw.shape, tab, g["kscale"])
    eng.set_raw(np.maximum(raw.astype(np.float32) - 100.0, 0.0))
    eng.iterate(4)
    X = eng.get()
    eng.set_raw_u16(raw, 100.0)
    eng.iterate(4)
    assert np.array_equal(eng.get(), X)
    w = eng.deskew(0.4, 0.104, 31.5, width=300)
    assert w == 300
    d = eng.post_get()
    ref = cdn.deskew(X, 31.5, 0.4, 0.104, width=300)
    assert d.shape == ref.shape and np.array_equal(d, ref)
    M, _ = small_affine(d.shape, seed=4)
    eng.affine(M)
    a = eng.post_get()
    assert np.array_equal(a, cdn.affine(d, M))
    s = eng.post_stats()
    assert abs(s[0] - float(a.astype(np.float64).sum())) <= 1e-9 * abs(s[0]) and s[1] == a.min() and s[2] == a.max()
    eng.close()


def test_engine_random_shapes_match_oracle(cdn, orc):
    """Seeded random 2,3,5,7-smooth shapes (72 and 100 compiled at run time),
    3 iterations each, against the FP64 oracle."""
    rng = np.random.default_rng(2024)
    lens = [64, 96, 128, 192, 72, 100]
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import gaussian_psf
    tab = make_otf(gaussian_psf(1.5, 3.0))
    for _ in range(5):
        shape = tuple(int(v) for v in rng.choice(lens, 3))
        print("random shape", shape, flush=True)
        nz, ny, nx = shape
        g = orc.init_geometry(nx, ny, nz, 0.104, 0.2, 0.104, 0.2, 0.0, 0.0, 0, False,
                              (2 * tab.shape[1], tab.shape[0], 1))
        R = rng.gamma(2.0, 40.0, shape)
        H = orc.otf_interp(tab, g["otf_dims"], g["fft_dims"], g["kscale"])
        ref, lam_ref = orc.rl_ba(R, H, 3)
        eng = cdn.RLEngine(shape, tab, g["kscale"])
        eng.set_raw(R.astype(np.float32))
        lam, _ = eng.iterate(3)
        got = eng.get()
        eng.close()
        assert _relmax(got, ref) <= TOL, (shape, _relmax(got, ref))
        assert np.abs(lam - lam_ref).max() < 1e-4


def test_engine_larger_than_memory_fails_cleanly(cdn):
    """40 stacks of 512 x 2048 x 2048 need 344 GB per volume, more than the
    GPU holds: creation reports the failed allocation (no exit, no crash) and
    keeps nothing; an ordinary engine works afterwards."""
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import gaussian_psf
    tab = make_otf(gaussian_psf(1.5, 3.0))
    ks = (np.float32(0.25), np.float32(0.25), np.float32(0.5))
    with pytest.raises(RuntimeError, match="(?i)memory"):
        cdn.RLEngine((512, 2048, 2048), tab, ks, stacks=40)
    e = cdn.RLEngine((64, 128, 128), tab, ks)
    e.set_raw(np.random.default_rng(1).uniform(0, 100, (64, 128, 128)).astype(np.float32))
    e.iterate(2)
    assert np.isfinite(e.get()).all()
    e.close()


def test_nan_inputs_are_rejected(cdn):
    """A NaN or Inf in a staged stack or seed is reported (it would spread
    through the FFT to every voxel), and the engine stays usable."""
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import gaussian_psf
    tab = make_otf(gaussian_psf(1.5, 3.0))
    shape = (64, 64, 64)
    eng = cdn.RLEngine(shape, tab, (0.5, 0.5, 0.5))
    R = np.full(shape, 10.0, np.float32)
    R[3, 4, 5] = np.nan
    with pytest.raises(RuntimeError, match="NaN or Inf"):
        eng.set_raw(R)
    R[3, 4, 5] = np.inf
    with pytest.raises(RuntimeError, match="1 NaN or Inf"):
        eng.set_raw(R)
    R[3, 4, 5] = 10.0
    eng.set_raw(R)
    S = R.copy()
    S[0, 0, 0] = np.nan
    with pytest.raises(RuntimeError, match="NaN or Inf"):
        eng.set_seed(S)
    with pytest.raises(RuntimeError, match="negative iteration count"):
        eng.iterate(-1)
    lam, ms = eng.iterate(0)
    assert ms == 0.0
    eng.iterate(2)
    assert np.isfinite(eng.get()).all()
    eng.close()


def test_copy_probe_keeps_a_loaded_seed(cdn):
    """The copy-rate probe writes a buffer of its own: a seed loaded before
    it still starts the seeded iteration."""
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import gaussian_psf
    tab = make_otf(gaussian_psf(1.5, 3.0))
    shape = (64, 64, 64)
    rng = np.random.default_rng(5)
    R = rng.uniform(1, 100, shape).astype(np.float32)
    seed = rng.uniform(1, 100, shape).astype(np.float32)
    outs = []
    for probe in (False, True):
        eng = cdn.RLEngine(shape, tab, (0.5, 0.5, 0.5))
        eng.set_raw(R)
        eng.set_seed(seed)
        if probe:
            eng.copy_probe(2)
        eng.iterate(3, seeded=True)
        outs.append(eng.get())
        eng.close()
    assert np.array_equal(outs[0], outs[1])

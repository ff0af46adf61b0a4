"""Pins the CPU oracle (oracle/oracle.cpp + oracle/__init__.py).

The reference ships no golden vectors or known-answer tests for this path
(its only tests exercise buffer copies), so each oracle function is pinned
here by a library routine, brute force, a closed form or an invariant:

| oracle function            | pinned by                                                        |
|----------------------------|------------------------------------------------------------------|
| fft1d / rfftn / irfftn     | numpy pocketfft FP64 fixtures (tests/golden), brute-force DFT,   |
|                            | Parseval, delta, shift theorem, round trip                       |
| otf_interp (RA and 3-D)    | constant / linear tables, kr cut-off, kz wrap at nzotf-1         |
| filt                       | H = 1 identity, delta -> point-spread, adjoint identity of conj  |
| rl_ba                      | H = 1 fixed point, flux conservation, brute-force DFT RL on a    |
|                            | tiny volume, lambda range and k < 2 rule                         |
| deskew / rotate / affine   | zero factor copy with padded last column, integer shifts,        |
|                            | identity rotation = z flip, identity / translation affine        |
| apodize / zblend / bgsub   | closed forms on linear ramps                                      |
| mean_above                 | brute force                                                      |
| find_optimal_dimension     | exhaustive against a brute factoriser                            |
| init_geometry              | the light-sheet sizes (128x512x512 raw -> 810 wide, 192 x 791 rotated)  |
"""
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz")


# ----------------------------------------------------------------------------
# FFT
# ----------------------------------------------------------------------------
@pytest.fixture(scope="module")
def golden():
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_golden
    return np.load(GOLDEN), make_golden


def test_fft1d_matches_pocketfft_fixtures(orc, golden):
    g, mg = golden
    for n in mg.LENGTHS_1D:
        x = mg.inputs_1d(n)
        f = orc.fft1d(x)
        b = orc.fft1d(x, inverse=True)
        assert np.abs(f - g[f"fft_{n}"]).max() <= 1e-12 * np.abs(g[f"fft_{n}"]).max(), n
        assert np.abs(b - g[f"ifft_{n}"]).max() <= 1e-12 * np.abs(g[f"ifft_{n}"]).max(), n


def test_rfftn_irfftn_match_pocketfft_fixtures(orc, golden):
    g, mg = golden
    for s in mg.SHAPES_3D:
        key = "x".join(map(str, s))
        v = mg.inputs_3d(s)
        S = orc.rfftn(v)
        assert np.abs(S - g[f"rfftn_{key}"]).max() <= 1e-12 * np.abs(S).max(), s
        w = orc.irfftn(g[f"irfftn_in_{key}"], s[2])
        ref = g[f"irfftn_{key}"]
        assert np.abs(w - ref).max() <= 1e-12 * np.abs(ref).max(), s


@pytest.mark.parametrize("shape", [(6, 10, 14), (7, 9, 15), (3, 4, 5)])
def test_rfftn_brute_force_dft(orc, shape):
    rng = np.random.default_rng(3)
    v = rng.standard_normal(shape)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    S = orc.rfftn(v)
    for kz in range(nz):
        for ky in range(ny):
            for kx in range(nx // 2 + 1):
                ph = np.exp(-2j * np.pi * (kz * z / nz + ky * y / ny + kx * x / nx))
                assert abs((v * ph).sum() - S[kz, ky, kx]) < 1e-10


def test_fft_identities(orc):
    rng = np.random.default_rng(4)
    v = rng.standard_normal((8, 12, 30))
    S = orc.rfftn(v)
    # round trip (unnormalised both ways)
    assert np.abs(orc.irfftn(S, 30) / v.size - v).max() < 1e-12
    # Parseval with half-spectrum weights
    w = np.full(S.shape[2], 2.0)
    w[0] = 1.0
    w[-1] = 1.0  # nx even: kx = nx/2 counted once
    assert abs((np.abs(S) ** 2 * w).sum() / v.size - (v ** 2).sum()) < 1e-8 * (v ** 2).sum()
    # delta -> all-ones spectrum
    d = np.zeros((4, 6, 10))
    d[0, 0, 0] = 1.0
    assert np.abs(orc.rfftn(d) - 1.0).max() < 1e-14
    # shift theorem along x
    s = np.roll(v, 3, axis=2)
    kx = np.arange(16)
    assert np.abs(orc.rfftn(s) - S * np.exp(-2j * np.pi * 3 * kx / 30)).max() < 1e-10


# ----------------------------------------------------------------------------
# OTF interpolation (src/RLgpuImpl.cu:228-366)
# ----------------------------------------------------------------------------
def test_otf_ra_constant_table_and_cutoff(orc):
    nr, nzo = 9, 12
    tab = np.full((nr, nzo), 0.5 + 0.25j, np.complex64)
    grid = (32, 16, 8)
    ks = (np.float32(0.3), np.float32(0.4), np.float32(0.9))
    H = orc.otf_interp(tab, (nr, 1, nzo), grid, ks)
    nx, ny, nz = grid
    kx = np.arange(nx // 2 + 1) * float(ks[0])
    kyi = np.arange(ny)
    kyi = np.where(kyi > ny // 2, kyi - ny, kyi) * float(ks[1])
    kr = np.sqrt(kx[None, :] ** 2 + kyi[:, None] ** 2)
    inside = kr < nr - 1
    for z in range(nz):
        assert np.allclose(H[z][inside], 0.5 + 0.25j)
        assert np.all(H[z][~inside] == 0)


def test_otf_ra_linear_in_r_is_exact(orc):
    nr, nzo = 10, 6
    r = np.arange(nr)
    tab = np.repeat((2.0 * r + 1.0)[:, None], nzo, axis=1).astype(np.complex64)
    grid = (24, 20, 6)
    ks = (np.float32(0.37), np.float32(0.41), np.float32(1.0))
    H = orc.otf_interp(tab, (nr, 1, nzo), grid, ks)
    nx, ny, nz = grid
    kx = np.arange(nx // 2 + 1) * float(ks[0])
    ky = np.arange(ny)
    ky = np.where(ky > ny // 2, ky - ny, ky) * float(ks[1])
    kr = np.sqrt(kx[None, :] ** 2 + ky[:, None] ** 2)
    inside = kr < nr - 1
    assert np.allclose(H[0].real[inside], 2 * kr[inside] + 1, atol=1e-6)


def test_otf_ra_kz_wrap_at_last_row(orc):
    # kz index in [nzotf-1, nzotf) interpolates between the last and first columns
    nr, nzo = 4, 5
    tab = np.zeros((nr, nzo), np.complex64)
    tab[:, nzo - 1] = 1.0
    tab[:, 0] = 3.0
    # grid nz = 10, kzscale 0.45: kz = -1 -> -0.45 + 5 = 4.55 -> 0.45 * 1 + 0.55 * ... from index 4 to 0
    H = orc.otf_interp(tab, (nr, 1, nzo), (4, 1, 10), (np.float32(0.0), np.float32(0.0), np.float32(0.45)))
    az = 4.55 - 4
    assert abs(H[9, 0, 0] - ((1 - az) * 1.0 + az * 3.0)) < 1e-6


def test_otf_3d_trilinear_is_exact_for_linear_table(orc):
    nxo, nyo, nzo = 9, 16, 12
    z, y, x = np.meshgrid(np.arange(nzo), np.arange(nyo), np.arange(nxo), indexing="ij")
    tab = (1.0 + 0.5 * x + 0.0 * y + 0.0 * z).astype(np.complex64)
    H = orc.otf_interp(tab, (nxo, nyo, nzo), (20, 8, 6), (np.float32(0.7), np.float32(0.5), np.float32(0.5)))
    kx = np.arange(11) * 0.7
    inside = kx < nxo - 1
    assert np.allclose(H[0, 0].real[inside], 1.0 + 0.5 * kx[inside], atol=1e-5)
    assert np.all(H[0, 0][~inside] == 0)


# ----------------------------------------------------------------------------
# filter and RL-BA
# ----------------------------------------------------------------------------
def test_filter_unit_otf_is_identity_and_conj_is_adjoint(orc):
    rng = np.random.default_rng(5)
    shape = (6, 10, 14)
    v = rng.standard_normal(shape)
    H1 = np.ones((6, 10, 8), np.complex128)
    assert np.abs(orc.filt(v, H1) - v).max() < 1e-12
    H = orc.rfftn(np.exp(-0.5 * ((np.arange(14) - 7.0) / 1.3) ** 2)[None, None, :] * np.ones(shape))
    H = H / np.abs(H).max() * np.exp(0.3j)
    w = rng.standard_normal(shape)
    lhs = (orc.filt(v, H) * w).sum()
    rhs = (v * orc.filt(w, H, conj=True)).sum()
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)


def test_filter_delta_gives_point_spread(orc):
    shape = (8, 8, 12)
    psf = np.random.default_rng(6).uniform(0, 1, shape)
    psf /= psf.sum()
    H = orc.rfftn(psf)
    d = np.zeros(shape)
    d[0, 0, 0] = 1.0  # conv(delta, H) = C2R(H) / N = psf
    assert np.abs(orc.filt(d, H) - psf).max() < 1e-15


def test_rl_unit_otf_is_fixed_point(orc):
    rng = np.random.default_rng(7)
    raw = rng.uniform(1, 100, (6, 10, 14))
    X, lam = orc.rl_ba(raw, np.ones((6, 10, 8), np.complex128), 5)
    eps = np.finfo(np.float32).eps
    assert np.abs(X - raw).max() <= 20 * eps * raw.max()
    assert lam[0] == 0 and lam[1] == 0
    assert np.all((lam >= 0) & (lam <= 1))


def test_rl_conserves_flux_with_normalised_otf(orc):
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import gaussian_psf, make_stack
    raw, _, _ = make_stack("beads_64x128x128", shape=(32, 64, 64))
    R = np.maximum(raw.astype(np.float64) - 100, 0)
    tab = make_otf(gaussian_psf(1.5, 3.0))
    g = orc.init_geometry(64, 64, 32, 0.104, 0.2, 0.104, 0.2, 0.0, 0.0, 0, False, (66, 17, 1))
    H = orc.otf_interp(tab, g["otf_dims"], g["fft_dims"], g["kscale"])
    X, lam = orc.rl_ba(R, H, 8)
    assert abs(X.sum() / R.sum() - 1) < 1e-5
    assert np.all((lam >= 0) & (lam <= 1)) and lam[2] > 0


def _rl_bruteforce(raw, psf_full, n_iter):
    """Appendix-A RL-BA with circular convolution by explicit DFT matrices."""
    nz, ny, nx = raw.shape
    N = raw.size

    def dftm(n, s):
        k = np.arange(n)
        return np.exp(s * 2j * np.pi * np.outer(k, k) / n)

    Fz, Fy, Fx = dftm(nz, -1), dftm(ny, -1), dftm(nx, -1)
    Iz, Iy, Ix = dftm(nz, 1), dftm(ny, 1), dftm(nx, 1)
    fwd = lambda v: np.einsum("az,by,cx,zyx->abc", Fz, Fy, Fx, v)
    inv = lambda S: np.einsum("az,by,cx,zyx->abc", Iz, Iy, Ix, S).real
    H = fwd(psf_full)
    conv = lambda v, h: inv(h * fwd(v / N))
    eps = float(np.finfo(np.float32).eps)
    X = raw.copy()
    Xp = G1 = G2 = None
    for k in range(n_iter):
        if k > 1:
            lam = min(max((G1 * G2).sum() / ((G2 * G2).sum() + eps), 0.0), 1.0)
            Y = np.maximum(X + lam * (X - Xp), 0)
        else:
            Y = X.copy()
        Xp = X.copy()
        if k > 0:
            G2 = G1.copy()
        C = conv(Y, H)
        C = np.where(np.abs(C) > 0, C, eps)
        C = np.clip(raw / C + eps, -10, 10)
        C = conv(C, np.conj(H))
        X = np.maximum(C * Y, 0)
        G1 = X - Y
    return X


def test_rl_matches_brute_force_dft(orc):
    rng = np.random.default_rng(8)
    shape = (4, 6, 8)
    psf = rng.uniform(0, 1, shape) ** 4
    psf /= psf.sum()
    raw = rng.uniform(0, 50, shape)
    H = orc.rfftn(psf)
    X, _ = orc.rl_ba(raw, H, 6)
    ref = _rl_bruteforce(raw, psf, 6)
    assert np.abs(X - ref).max() < 1e-9 * np.abs(ref).max()


# ----------------------------------------------------------------------------
# preprocessing and geometry
# ----------------------------------------------------------------------------
def test_bgsub_and_mean_above(orc):
    rng = np.random.default_rng(9)
    v = rng.uniform(0, 200, (5, 7, 9))
    b = orc.bgsub(v, 100.0)
    assert np.array_equal(b, np.maximum(v - 100.0, 0))
    m = np.float32(b.sum() / b.size)
    sel = b.astype(np.float32) > m
    assert abs(orc.mean_above(b) - b[sel].sum() / sel.sum()) < 1e-9


def test_apodize_closed_form(orc):
    nz, ny, nx = 2, 10, 12
    img = np.zeros((nz, ny, nx))
    img[:, -1, :] = 8.0  # last row 8, first row 0 -> diff = 4 along y
    out = orc.apodize(img, 3)
    for l in range(3):
        fact = 4.0 * (1 - math.sin((l + 0.5) / 3 * math.pi / 2))
        assert abs(out[0, l, 5] - fact) < 1e-5
        assert abs(out[0, ny - 1 - l, 5] - ((8.0 if l == 0 else 0.0) - fact)) < 1e-5


def test_zblend_closed_form(orc):
    img = np.zeros((8, 3, 4))
    img[-1] = 6.0
    out = orc.zblend(img, 2)
    for zi in range(2):
        fact = 6.0 * (1 - math.sin((zi + 0.5) / 2 * math.pi / 2))
        assert abs(out[zi, 1, 1] - fact) < 1e-5
        assert abs(out[7 - zi, 1, 1] - ((6.0 if zi == 0 else 0.0) - fact)) < 1e-5


def test_deskew_zero_factor_is_copy_with_padded_last_column(orc):
    rng = np.random.default_rng(10)
    v = rng.uniform(0, 1, (3, 4, 10))
    out = orc.deskew(v, 0.0, 10, 0, pad_val=-7.0)
    assert np.array_equal(out[..., :9], v[..., :9])
    assert np.all(out[..., 9] == -7.0)
    sh = orc.deskew(v, 0.0, 10, 2, pad_val=0.0)  # xin = x + 2
    assert np.array_equal(sh[..., :7], v[..., 2:9])


def test_rotate_identity_is_z_flip(orc):
    rng = np.random.default_rng(11)
    v = rng.uniform(0, 1, (6, 3, 8))
    out = orc.rotate(v, np.array([1, 0, 0, 1], np.float32), 8, 6)
    assert np.array_equal(out[::-1][:5, :, :7], v[:5, :, :7])
    assert np.all(out[0] == 0) and np.all(out[:, :, 7] == 0)


def test_affine_identity_and_translation(orc):
    rng = np.random.default_rng(12)
    v = rng.uniform(0, 1, (5, 6, 7))
    eye = np.eye(4, dtype=np.float32)
    assert np.array_equal(orc.affine(v, eye), v)
    t = eye.copy()
    t[0, 3] = 2.0  # pull from x + 2
    out = orc.affine(v, t)
    assert np.array_equal(out[..., :5], v[..., 2:])
    assert np.all(out[..., 5:] == 0)
    # transformKernelRA omits the +0.5 texel-centre offset of transformKernel
    # (src/geometryTransform.cu:206-208 vs :241-248): an identity matrix
    # samples at integer texture coordinates, i.e. the 2x2x2 average of
    # voxels (x-1..x, y-1..y, z-1..z) with zeros outside the volume
    p = np.pad(v, ((1, 0), (1, 0), (1, 0)))
    avg = sum(p[a:a + 5, b:b + 6, c:c + 7] for a in (0, 1) for b in (0, 1) for c in (0, 1)) / 8.0
    assert np.allclose(orc.affine_ra(v, 0.1, 0.1, 0.3, eye), avg, atol=1e-6)


def test_find_optimal_dimension_exhaustive(orc):
    def smooth(n):
        if n in (2, 3, 5, 7):
            return True
        for p in (2, 3, 5, 7):
            while n % p == 0 and n > 1:
                n //= p
        return n == 1

    for n in range(2, 8193):
        o = orc.find_optimal_dimension(n)
        assert smooth(o) and o <= n and all(not smooth(m) for m in range(o + 1, n + 1)), n
    assert orc.find_optimal_dimension(826) == 810 and orc.find_optimal_dimension(3727) == 3675


def test_init_geometry_lightsheet_128x512x512(orc):
    # 128 x 512 x 512 raw, 31.5 deg, dz 0.3, dx 0.104: deskewed 512 + 314 = 826 -> 810;
    # dz' = 0.15675; rotation gives 192 x 791 (z, x)
    g = orc.init_geometry(512, 512, 128, 0.104, 0.3, 0.104, 0.15675, 31.5, 31.5, 0, False, (66, 17, 1))
    assert g["deskewed_nx"] == 810
    assert g["out_dims"] == (791, 512, 192)
    assert abs(g["deskew_factor"] - math.cos(math.radians(31.5)) * 0.3 / 0.104) < 1e-6


def test_camcor_closed_forms(orc):
    """a = 0 (or b = 0) leaves v - offset; the previous plane of z = 0 is
    plane nz-1; conversion truncates and saturates."""
    rng = np.random.default_rng(0)
    raw = rng.integers(0, 3000, (5, 4, 6)).astype(np.uint16)
    cam = np.zeros((3, 4, 6), np.float32)
    cam[2] = 100.25
    out, r = orc.camcor(raw, cam)
    assert np.array_equal(out, np.clip(np.trunc(raw.astype(np.float64) - 100.25), 0, 65535).astype(np.uint16))
    # b = 0 also removes the correction whatever a is
    cam[0] = 50.0
    out0, _ = orc.camcor(raw, cam)
    assert np.array_equal(out0, out)
    # one pixel by hand, z = 0 uses plane nz-1
    cam[1] = 0.01
    out1, r1 = orc.camcor(raw, cam)
    v, p = float(raw[0, 1, 2]), float(raw[-1, 1, 2])
    want = v - 100.25 - 0.9 * 50.0 * (1 - np.exp(-0.01 * (p - 100.25)))
    assert abs(r1[0, 1, 2] - want) < 1e-9
    v, p = float(raw[3, 1, 2]), float(raw[2, 1, 2])
    assert abs(r1[3, 1, 2] - (v - 100.25 - 45.0 * (1 - np.exp(-0.01 * (p - 100.25))))) < 1e-9
    # saturation
    big = np.full((2, 1, 2), 65535, np.uint16)
    cz = np.zeros((3, 1, 2), np.float32)
    cz[2] = -10.0
    assert np.all(orc.camcor(big, cz)[0] == 65535)
    cz[2] = 70000.0
    assert np.all(orc.camcor(big, cz)[0] == 0)


def test_wiener_closed_forms(orc):
    """H = 1 everywhere inside the cut-off: the filter is 1 / (1 + w^2) times
    the (1 - kr / rcutoff) apodisation; a cut-off above every kr with w = 0
    and H = 1 is the identity up to the apodisation of the DC term (1)."""
    rng = np.random.default_rng(1)
    v = rng.standard_normal((8, 12, 16))
    tab = np.ones((40, 8), np.complex64)   # RA table, nr = 40 > every kr used
    ks = (np.float32(1.0), np.float32(1.0), np.float32(1.0))
    # rcutoff huge -> apodisation ~ 1; w = 0 -> identity
    out = orc.wiener(v, tab, (40, 1, 8), ks, (1e-6, 1e-6), 1e3, 0.0)
    assert np.abs(out - v).max() < 1e-6
    # w = 1 halves everything
    out = orc.wiener(v, tab, (40, 1, 8), ks, (1e-6, 1e-6), 1e3, 1.0)
    assert np.abs(out - 0.5 * v).max() < 1e-6
    # background is subtracted first
    out = orc.wiener(v + 7.0, tab, (40, 1, 8), ks, (1e-6, 1e-6), 1e3, 0.0, background=7.0)
    assert np.abs(out - v).max() < 1e-6
    # a cut-off below the first lateral bin keeps only kx = ky = 0 (all kz):
    # each plane becomes its mean
    out = orc.wiener(v, tab, (40, 1, 8), ks, (1.0, 1.0), 0.5, 0.0)
    assert np.abs(out - v.mean(axis=(1, 2), keepdims=True)).max() < 1e-9


def test_otf_shape_mismatch_raises():
    """The C routines index the OTF by the volume's shape: a mismatched OTF is
    refused before the call (a skewed decon OTF on a deskewed volume, say)."""
    import oracle as orc
    v = np.zeros((4, 4, 8))
    with pytest.raises(ValueError):
        orc.rl_ba(v, np.zeros((4, 4, 4), np.complex128), 1)
    with pytest.raises(ValueError):
        orc.rl_ba_f32(v.astype(np.float32), np.zeros((4, 4, 6), np.complex64), 1)
    with pytest.raises(ValueError):
        orc.filt(v, np.zeros((4, 3, 5), np.complex128))

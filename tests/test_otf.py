"""Radially averaged OTF generation (cudadecon_amd.otf, restating
src/radialft_interface.cpp:95-518) pinned by closed forms."""
import numpy as np

from cudadecon_amd.otf import _fixorigin, center_and_background, make_otf, write_otf_tiff
from cudadecon_amd.synth import gaussian_psf


def test_center_of_centred_gaussian_and_z_quirk():
    (xc, yc, zc), bg = center_and_background(gaussian_psf(1.5, 3.0))
    assert abs(xc - 16) < 1e-6 and abs(yc - 16) < 1e-6
    # +0.6 added to the z centre (src/radialft_interface.cpp:239)
    assert abs(zc - 16.6) < 1e-5
    # no rows above yc - 20 exist: 0 / negative -> -0.0
    assert bg == 0


def test_gaussian_otf_closed_form():
    sxy, sz = 1.5, 3.0
    otf = make_otf(gaussian_psf(sxy, sz))
    assert otf.shape == (17, 33)
    assert abs(np.abs(otf).max() - 1) < 1e-6
    # lateral profile at kz = 0: the bin average of exp(-2 pi^2 sxy^2 (kr/33)^2)
    # over the (kx >= 0, ky) samples with rint(kr) = r
    kx = np.arange(17)[None, :]
    ky = np.arange(33)
    ky = np.where(ky > 16, ky - 33, ky)[:, None]
    kr = np.sqrt(kx ** 2 + ky ** 2)
    g = np.exp(-2 * np.pi ** 2 * sxy ** 2 * (kr / 33.0) ** 2)
    rb = np.floor(kr + 0.5)
    r = np.arange(1, 9)
    ref = np.array([g[rb == k].mean() for k in r])
    # ratios to r = 1 (the r = 0 row is the linear extrapolation of fixorigin)
    assert np.allclose(np.abs(otf[r, 0]) / np.abs(otf[1, 0]), ref / ref[0], rtol=2e-3)
    # axial profile at r = 0 and the phase ramp from the 0.6 z offset
    kz = np.arange(1, 6)
    refz = np.exp(-2 * np.pi ** 2 * sz ** 2 * (kz / 33.0) ** 2)
    assert np.allclose(np.abs(otf[1, kz]) / np.abs(otf[1, 0]), refz, rtol=2e-3)
    assert np.allclose(np.angle(otf[1, kz]), 2 * np.pi * 0.6 * kz / 33.0, atol=2e-3)
    # kz Hermitian symmetry
    assert np.allclose(otf[:, 1:], np.conj(otf[:, :0:-1]), atol=1e-7)


def test_fixorigin_linear_extrapolation_is_exact():
    nr, nz = 12, 4
    avg = np.zeros((nr, nz), np.complex64)
    for z in range(nz):
        avg[:, z] = (10 - 0.5 * np.arange(nr)) * (1 + 0.1j * z)
    avg[0] = 0
    assert _fixorigin(avg, nz, 10)
    assert np.allclose(avg[0], 10 * (1 + 0.1j * np.arange(nz)), atol=1e-5)


def test_tiff_layout(tmp_path):
    tab = (np.arange(17 * 33) + 1j * np.arange(17 * 33)).astype(np.complex64).reshape(17, 33)
    p = tmp_path / "o.tif"
    write_otf_tiff(str(p), tab)
    b = p.read_bytes()
    assert b[:4] == b"II*\x00"
    # pixel data follows the first IFD: width 66 floats x height 17
    data = np.frombuffer(b[8 + 2 + 12 * 10 + 4:], dtype="<f4")
    assert np.array_equal(data.view(np.complex64).reshape(17, 33), tab)

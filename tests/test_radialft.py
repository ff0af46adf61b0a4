"""libradialft.so (makeOTF, C++) against the Python restatement in
cudadecon_amd/otf.py (itself pinned by closed forms in tests/test_otf.py):
two independent implementations of src/radialft_interface.cpp:95-518 on the
same PSF stacks.  The C++ FFT runs in double, numpy's in single precision;
tables are normalised to max 1 and agree within 3e-5."""
import os
import re

import numpy as np
import pytest

from cudadecon_amd import radialft as rft
from cudadecon_amd.otf import make_otf, make_otf_3d, read_tiff_f32, write_tiff_f32

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _psf(shape, centre, sxy=1.5, sz=3.0, amp=1000.0, bg=90.0, seed=0):
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    cz, cy, cx = centre
    g = np.exp(-0.5 * (((x - cx) / sxy) ** 2 + ((y - cy) / sxy) ** 2 + ((z - cz) / sz) ** 2))
    rng = np.random.default_rng(seed)
    return (amp * g + bg + rng.uniform(0, 1, shape)).astype(np.float32)


def _run(tmp_path, psf, **kw):
    pin, pout = tmp_path / "psf.tif", tmp_path / "otf.tif"
    write_tiff_f32(str(pin), psf)
    rft.make_otf_file(str(pin), str(pout), **kw)
    return read_tiff_f32(str(pout))


def test_header_symbols_exported():
    text = open(os.path.join(ROOT, "include", "radialft.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = sorted({n for n in re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text) if not n.isupper()})
    assert names == sorted(rft.SIGNATURES)
    L = rft.lib()
    for n in names:
        assert hasattr(L, n)


@pytest.mark.parametrize("shape,centre", [((33, 33, 33), (16.3, 16.0, 15.7)), ((31, 40, 45), (14.0, 22.4, 20.2))],
                         ids=["33cube", "31x40x45"])
def test_rotational_average_matches_python(tmp_path, shape, centre):
    psf = _psf(shape, centre)
    img = _run(tmp_path, psf, interpkr=10)
    got = img[0].view(np.complex64)          # height nx/2+1, width 2*nz
    ref = make_otf(psf, interpkr=10)
    assert got.shape == ref.shape == (shape[2] // 2 + 1, shape[0])
    assert np.abs(got - ref).max() < 3e-5


def test_user_background_and_cleanup_match_python(tmp_path):
    psf = _psf((33, 64, 64), (16.0, 31.2, 32.6), sxy=1.2)
    kw = dict(lambdanm=520, NA=1.1, NIMM=1.33, dr=0.104, dz=0.2, krmax=0)
    img = _run(tmp_path, psf, interpkr=8, background=90.0, cleanup=True, **kw)
    got = img[0].view(np.complex64)
    ref = make_otf(psf, interpkr=8, background=90.0, cleanup=kw)
    assert np.array_equal(got == 0, ref == 0)
    assert np.abs(got - ref).max() < 3e-5
    # krmax below the NA limit cuts more
    kw2 = dict(kw, krmax=10)
    got2 = _run(tmp_path, psf, interpkr=8, background=90.0, cleanup=True, **kw2)[0].view(np.complex64)
    ref2 = make_otf(psf, interpkr=8, background=90.0, cleanup=kw2)
    assert np.array_equal(got2 == 0, ref2 == 0)
    assert np.count_nonzero(got2 == 0) > np.count_nonzero(got == 0)


def test_three_d_output_matches_python(tmp_path):
    psf = _psf((17, 24, 30), (8.2, 11.6, 15.1))
    img = _run(tmp_path, psf, interpkr=6, three_d=True)   # pages nz, height ny, width 2*(nx/2+1)
    got = img.view(np.complex64)
    ref = make_otf_3d(psf, interpkr=6)
    assert got.shape == ref.shape == (17, 24, 16)
    assert np.abs(got - ref).max() < 3e-5


def test_errors_return_one(tmp_path):
    with pytest.raises(RuntimeError, match="cannot open"):
        rft.make_otf_file(str(tmp_path / "missing.tif"), str(tmp_path / "o.tif"))
    psf = _psf((9, 12, 12), (4.0, 6.0, 6.0))
    with pytest.raises(RuntimeError, match="interpkr"):
        _run(tmp_path, psf, interpkr=10)

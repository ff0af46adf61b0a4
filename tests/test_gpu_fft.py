"""GPU parity of the hand-written 3-D R2C/C2R FFT and of one filterGPU call
against the FP64 oracle (FP32 on the GPU: relative max error <= 2e-6 of the
largest magnitude for the transforms, 1e-5 for the filter)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LENGTHS = [64, 96, 128, 192, 256, 384, 512, 768, 810, 1024, 2048]


def _shapes():
    out = [(64, 128, 128), (96, 192, 256)]
    for n in LENGTHS:
        out += [(64, 64, n), (64, n, 64), (n, 64, 64)]
    return out


def _relerr(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("shape", _shapes(), ids=lambda s: "x".join(map(str, s)))
def test_r2c_matches_oracle(cdn, orc, shape):
    print("r2c", shape, flush=True)
    rng = np.random.default_rng(sum(shape))
    v = rng.standard_normal(shape).astype(np.float32)
    got = cdn.fft_r2c(v)
    ref = orc.rfftn(v.astype(np.float64))
    assert _relerr(got, ref) < 2e-6


@pytest.mark.parametrize("shape", _shapes(), ids=lambda s: "x".join(map(str, s)))
def test_c2r_matches_oracle_including_non_hermitian_bins(cdn, orc, shape):
    print("c2r", shape, flush=True)
    nz, ny, nx = shape
    rng = np.random.default_rng(1 + sum(shape))
    s = (rng.standard_normal((nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((nz, ny, nx // 2 + 1))).astype(
        np.complex64)
    got = cdn.fft_c2r(s, nx)
    ref = orc.irfftn(s.astype(np.complex128), nx)
    assert _relerr(got, ref) < 2e-6


def test_roundtrip_identity(cdn):
    rng = np.random.default_rng(5)
    v = rng.standard_normal((64, 128, 810)).astype(np.float32)
    back = cdn.fft_c2r(cdn.fft_r2c(v), 810) / v.size
    assert np.abs(back - v).max() < 1e-5


def _ra_table(nr=17, nz=33, seed=0):
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import gaussian_psf
    return make_otf(gaussian_psf(1.5, 3.0))


@pytest.mark.parametrize("conj", [False, True])
# 256 x 16 x 1024: z length 256 (6-wave z filter, 32-column tiles) with the
# kx scale of a 1024-wide stack (1/32)
@pytest.mark.parametrize("shape", [(64, 128, 128), (128, 64, 810), (64, 96, 192), (256, 16, 1024)],
                         ids=lambda s: str(s))
def test_filter_ra_matches_oracle(cdn, orc, shape, conj):
    print("filter", shape, conj, flush=True)
    tab = _ra_table()
    g = orc.init_geometry(shape[2], shape[1], shape[0], 0.104, 0.2, 0.104, 0.2, 0.0, 0.0, 0, False,
                          (2 * tab.shape[1], tab.shape[0], 1))
    rng = np.random.default_rng(3)
    v = rng.uniform(0, 100, shape).astype(np.float32)
    H = orc.otf_interp(tab, g["otf_dims"], g["fft_dims"], g["kscale"])
    ref = orc.filt(v.astype(np.float64), H, conj=conj)
    got = cdn.filter_volume(v, tab, g["kscale"], conj=conj)
    assert _relerr(got, ref) < 1e-5


def test_filter_3d_otf_matches_oracle(cdn, orc):
    rng = np.random.default_rng(9)
    nzo, nyo, nxo = 24, 40, 21
    tab = (rng.standard_normal((nzo, nyo, nxo)) + 1j * rng.standard_normal((nzo, nyo, nxo))).astype(np.complex64)
    shape = (64, 96, 128)
    ks = (np.float32(0.3), np.float32(0.7), np.float32(0.45))
    v = rng.uniform(0, 10, shape).astype(np.float32)
    H = orc.otf_interp(tab, (nxo, nyo, nzo), (shape[2], shape[1], shape[0]), ks)
    ref = orc.filt(v.astype(np.float64), H)
    got = cdn.filter_volume(v, tab, ks)
    assert _relerr(got, ref) < 1e-5


# lengths outside the precompiled set: kernels compiled at run time (hipRTC)
JIT_SHAPES = [(70, 72, 150), (16, 20, 3675), (45, 250, 64), (105, 64, 96), (8, 12, 1000)]


@pytest.mark.parametrize("shape", JIT_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_runtime_compiled_lengths_match_oracle(cdn, orc, shape):
    print("jit", shape, flush=True)
    nz, ny, nx = shape
    rng = np.random.default_rng(sum(shape) + 17)
    v = rng.standard_normal(shape).astype(np.float32)
    assert _relerr(cdn.fft_r2c(v), orc.rfftn(v.astype(np.float64))) < 2e-6
    s = (rng.standard_normal((nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((nz, ny, nx // 2 + 1))).astype(
        np.complex64)
    assert _relerr(cdn.fft_c2r(s, nx), orc.irfftn(s.astype(np.complex128), nx)) < 2e-6

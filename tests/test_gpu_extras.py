"""GPU parity of the two callers beside RL: camera correction
(camcor_interface, src/camcor.cu:80-108) and the one-step Wiener filter
(cudaDecon_wiener, src/linearDecon.cpp:116-168).

Camera correction converts a float to uint16 by truncation, so a 1-count
difference can appear where the FP64 value lies within FP32 rounding of an
integer; the test requires exact equality everywhere else and at most one
count, on at most 1e-4 of the voxels, at such boundaries.  Wiener: FP32 GPU
against the FP64 oracle within 1e-5 of the largest magnitude."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(7, 33, 45), (32, 256, 256)], ids=lambda s: "x".join(map(str, s)))
def test_camcor_matches_oracle(cdn, orc, shape):
    print("camcor", shape, flush=True)
    nz, ny, nx = shape
    rng = np.random.default_rng(nz)
    raw = rng.integers(80, 4000, shape).astype(np.uint16)
    raw.flat[::97] = 65535
    cam = np.stack([rng.uniform(0, 40, (ny, nx)), rng.uniform(0, 0.02, (ny, nx)),
                    rng.uniform(90, 110, (ny, nx))]).astype(np.float32)
    got = cdn.camcor(raw, cam)
    ref, r = orc.camcor(raw, cam)
    diff = got.astype(np.int64) - ref.astype(np.int64)
    near = np.abs(r - np.round(r)) < 1e-3 * np.maximum(1.0, np.abs(r))
    assert np.all(diff[~near] == 0)
    assert np.abs(diff).max() <= 1
    assert np.count_nonzero(diff) <= max(1, 1e-4 * diff.size)


def test_camcor_rejects_other_frame_size(cdn):
    cam = np.zeros((3, 8, 8), np.float32)
    raw = np.zeros((2, 8, 8), np.uint16)
    cdn.camcor(raw, cam)
    L = cdn.lib()
    out = np.zeros((2, 8, 9), np.uint16)
    assert L.camcor_interface(np.zeros((2, 8, 9), np.uint16).ctypes.data, 9, 8, 2, out.ctypes.data) == 0
    assert "differ" in cdn.last_error()


@pytest.mark.parametrize("shape", [(64, 96, 128), (32, 128, 810)], ids=lambda s: "x".join(map(str, s)))
def test_wiener_matches_oracle(cdn, orc, shape):
    print("wiener", shape, flush=True)
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import gaussian_psf
    nz, ny, nx = shape
    tab = make_otf(gaussian_psf(1.5, 3.0))
    dr, dz = 0.104, 0.2
    g = orc.init_geometry(nx, ny, nz, dr, dz, dr, dz, 0.0, 0.0, 0, False, (2 * tab.shape[1], tab.shape[0], 1))
    rng = np.random.default_rng(4)
    v = rng.uniform(90, 400, shape).astype(np.float32)
    dkxy = (1.0 / (dr * nx), 1.0 / (dr * ny))
    rcut = 2 * 1.2 / 0.525
    got = cdn.wiener_filter(v, tab, g["kscale"], dkxy, rcut, 0.05, background=100.0)
    ref = orc.wiener(v.astype(np.float64), tab, g["otf_dims"], g["kscale"], dkxy, rcut, 0.05, background=100.0)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("wiener rel err", err, flush=True)
    assert err <= 1e-5

"""cudaDecon_rl_run_noalloc, the RL-BA run that allocates nothing: called
through the C ABI on device buffers, twice in a row (the second stack
smaller, so the static scratch array still holds the first call's larger
contents), each result bitwise that of cudaDecon_rl_create / set_raw /
iterate / get on the same stack; stacks too large for the array and FFT
lengths that are not built in are refused before any GPU work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [((64, 96, 128), 6, 0), ((64, 64, 96), 5, 1)]  # (shape, iterations, PSF) - the second smaller


def _stack(shape, seed):
    from cudadecon_amd.synth import make_stack
    raw, _, _ = make_stack("beads_64x128x128", shape=shape, seed=seed)
    return np.maximum(raw.astype(np.float32) - np.float32(100.0), np.float32(0.0))


def test_noalloc_matches_engine_twice_second_smaller(cdn, orc):
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import BATCH_CHANNELS, gaussian_psf
    for call, (shape, iters, ch) in enumerate(CASES):
        print("no-alloc call", call, shape, iters, flush=True)
        tab = make_otf(gaussian_psf(*BATCH_CHANNELS[ch]))
        nz, ny, nx = shape
        g = orc.init_geometry(nx, ny, nz, 0.104, 0.2, 0.104, 0.2, 0.0, 0.0, 0, False,
                              (2 * tab.shape[1], tab.shape[0], 1))
        R = _stack(shape, 60 + call)
        eng = cdn.RLEngine(shape, tab, g["kscale"])
        eng.set_raw(R)
        eng.iterate(iters)
        ref = eng.get()
        eng.close()
        d_raw = cdn.DeviceBuffer(R.nbytes)
        d_out = cdn.DeviceBuffer(R.nbytes)
        d_raw.upload(R)
        cdn.rl_run_noalloc(d_raw, d_out, shape, tab, g["kscale"], iters)
        got = d_out.download(shape)
        assert np.array_equal(d_raw.download(shape), R)  # the stack is read only
        d_raw.close()
        d_out.close()
        assert np.array_equal(got, ref), call


def test_noalloc_refuses_what_it_cannot_run(cdn):
    """Checked on the host before any GPU work (the pointers are never used)."""
    tab = np.ones((17, 33), np.complex64)
    fake = 0x1000
    with pytest.raises(RuntimeError, match="static scratch holds"):
        cdn.rl_run_noalloc(fake, fake, (256, 1024, 1024), tab, (1, 1, 1), 3)
    with pytest.raises(RuntimeError, match="not built into the library"):
        cdn.rl_run_noalloc(fake, fake, (64, 64, 150), tab, (1, 1, 1), 3)
    with pytest.raises(RuntimeError, match="null argument"):
        cdn.rl_run_noalloc(None, fake, (64, 64, 64), tab, (1, 1, 1), 3)

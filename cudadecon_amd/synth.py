"""Deterministic synthetic light-sheet stacks (beads, smooth background,
Gaussian PSF blur, Poisson noise, camera offset, uint16).

Every stack is a function of its seed (numpy PCG64), so the same input can
be regenerated on any host.  Sizes and parameters of the named workloads:

* ``beads_64x128x128``  64 x 128 x 128, 300 beads, PSF sigma_xy 1.5 / sigma_z 3.0 px,
  10 iterations, no deskew;
* ``lightsheet_128x512x512``  128 x 512 x 512, 5000 beads, PSF 1.3 / 4.0 px, deskew 31.5 deg,
  dz 0.3 um, dx 0.104 um, then rotation, 15 iterations;
* ``beads_256x1024x1024``  256 x 1024 x 1024 plain RL-BA (the throughput workload);
* the 48-stack batch (``batch_shapes``, ``batch_stack``): shapes drawn with
  seed 2 from z in {64, 96, 128, 192, 256} and y, x in {256, 384, 512, 768,
  1024}; stack i has its own bead field (seed 1000 + i, one bead per 2e4
  voxels) imaged through channel PSF i mod 3 (sigma_xy 1.2 / 1.5 / 1.8 px,
  sigma_z 3 / 4 / 5 px), 10 iterations each.
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

PSF_SIZE = 33
SLAB = 8  # z-slices per noise chunk (fixed, so results do not depend on threads)


def _pool():
    return ThreadPoolExecutor(max(1, min(16, os.cpu_count() or 1)))


def gaussian_psf(sigma_xy: float, sigma_z: float, size: int = PSF_SIZE) -> np.ndarray:
    """Sampled anisotropic Gaussian on a size^3 grid centred at size//2,
    normalised to sum 1 (float32, (z, y, x))."""
    c = size // 2
    r = np.arange(size) - c
    gxy = np.exp(-0.5 * (r / sigma_xy) ** 2)
    gz = np.exp(-0.5 * (r / sigma_z) ** 2)
    psf = gz[:, None, None] * gxy[None, :, None] * gxy[None, None, :]
    return (psf / psf.sum()).astype(np.float32)


def _kernels(sigma_xy, sigma_z, size=PSF_SIZE):
    c = size // 2
    r = np.arange(size) - c
    gxy = np.exp(-0.5 * (r / sigma_xy) ** 2)
    gz = np.exp(-0.5 * (r / sigma_z) ** 2)
    return (gxy / gxy.sum()).astype(np.float32), (gz / gz.sum()).astype(np.float32)


def blur(vol: np.ndarray, sigma_xy: float, sigma_z: float) -> np.ndarray:
    """Linear 'same' convolution with the separable 33^3 Gaussian PSF
    (zero outside the volume)."""
    from scipy.ndimage import convolve1d
    kxy, kz = _kernels(sigma_xy, sigma_z)
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    nz, ny, _ = vol.shape
    tmp = np.empty_like(vol)
    out = np.empty_like(vol)

    def xy(z0):
        s = slice(z0, min(z0 + SLAB, nz))
        tmp[s] = convolve1d(convolve1d(vol[s], kxy, axis=2, mode="constant", cval=0.0), kxy, axis=1,
                            mode="constant", cval=0.0)

    def zz(y0):
        s = slice(y0, min(y0 + 32, ny))
        out[:, s] = convolve1d(tmp[:, s], kz, axis=0, mode="constant", cval=0.0)

    with _pool() as p:
        list(p.map(xy, range(0, nz, SLAB)))
        list(p.map(zz, range(0, ny, 32)))
    return out


def ground_truth(shape, n_beads, rng, n_gauss=3, gauss_sigma=(10.0, 30.0), gauss_amp=50.0,
                 amp=(200.0, 2000.0)) -> np.ndarray:
    """Beads at uniform integer voxel positions with U[amp] amplitudes plus
    n_gauss isotropic Gaussians of amplitude gauss_amp (float32)."""
    nz, ny, nx = shape
    vol = np.zeros(shape, np.float32)
    z = rng.integers(0, nz, n_beads)
    y = rng.integers(0, ny, n_beads)
    x = rng.integers(0, nx, n_beads)
    a = rng.uniform(amp[0], amp[1], n_beads).astype(np.float32)
    np.add.at(vol, (z, y, x), a)
    zz = np.arange(nz, dtype=np.float32)[:, None, None]
    yy = np.arange(ny, dtype=np.float32)[None, :, None]
    xx = np.arange(nx, dtype=np.float32)[None, None, :]
    for _ in range(n_gauss):
        cz, cy, cx = rng.uniform(0, nz), rng.uniform(0, ny), rng.uniform(0, nx)
        s = np.float32(rng.uniform(gauss_sigma[0], gauss_sigma[1]))
        gz = np.exp(-0.5 * ((zz - np.float32(cz)) / s) ** 2)
        gy = np.exp(-0.5 * ((yy - np.float32(cy)) / s) ** 2)
        gx = np.exp(-0.5 * ((xx - np.float32(cx)) / s) ** 2)
        vol += np.float32(gauss_amp) * (gz * gy * gx)
    return vol


def image(blurred: np.ndarray, seed_seq: np.random.SeedSequence, background=100.0, offset=100.0) -> np.ndarray:
    """+background, Poisson, +camera offset, round, clamp to uint16.  Each
    slab of SLAB z-slices draws from its own child generator of seed_seq."""
    nz = blurred.shape[0]
    out = np.empty(blurred.shape, np.uint16)
    kids = seed_seq.spawn((nz + SLAB - 1) // SLAB)

    def slab(i):
        s = slice(i * SLAB, min((i + 1) * SLAB, nz))
        lam = np.maximum(blurred[s].astype(np.float64) + background, 0.0)
        counts = np.random.default_rng(kids[i]).poisson(lam).astype(np.float64) + offset
        out[s] = np.clip(np.rint(counts), 0, 65535).astype(np.uint16)

    with _pool() as p:
        list(p.map(slab, range(len(kids))))
    return out


CONFIGS = {
    "beads_64x128x128": dict(shape=(64, 128, 128), beads=300, sxy=1.5, sz=3.0, seed=0, iters=10, dr=0.104, dz=0.2,
                 angle=0.0, rot=0.0),
    "lightsheet_128x512x512": dict(shape=(128, 512, 512), beads=5000, sxy=1.3, sz=4.0, seed=1, iters=15, dr=0.104, dz=0.3,
                 angle=31.5, rot=31.5),
    "beads_256x1024x1024": dict(shape=(256, 1024, 1024), beads=300 * 256, sxy=1.5, sz=3.0, seed=7, iters=10, dr=0.104, dz=0.2,
                 angle=0.0, rot=0.0),
}


# ---------------------------------------------------------------------------
# 48 stacks of uneven shape, three channels
# ---------------------------------------------------------------------------
BATCH_N = 48
BATCH_ZS = (64, 96, 128, 192, 256)
BATCH_XYS = (256, 384, 512, 768, 1024)
BATCH_CHANNELS = ((1.2, 3.0), (1.5, 4.0), (1.8, 5.0))  # (sigma_xy, sigma_z) px
BATCH_DR, BATCH_DZ = 0.104, 0.2  # um
BATCH_BEAD_DENSITY = 1.0 / 2e4   # beads per voxel
BATCH_ITERS = 10


def batch_shapes(n: int = BATCH_N, seed: int = 2):
    """(nz, ny, nx) of the n stacks: z, y, x drawn independently."""
    rng = np.random.default_rng(seed)
    return [(int(rng.choice(BATCH_ZS)), int(rng.choice(BATCH_XYS)), int(rng.choice(BATCH_XYS))) for _ in range(n)]


def batch_channel(i: int):
    """Channel index and (sigma_xy, sigma_z) of stack i."""
    c = i % len(BATCH_CHANNELS)
    return c, BATCH_CHANNELS[c]


def blur_fft(vol: np.ndarray, sigma_xy: float, sigma_z: float) -> np.ndarray:
    """The linear 'same' convolution of ``blur`` (separable 33^3 Gaussian,
    zero outside the volume) computed with zero-padded float32 FFTs on all
    host cores.  The PSF's transform is the outer product of its three 1-D
    transforms (it is separable), so only the volume is transformed."""
    import scipy.fft as sf
    kxy, kz = _kernels(sigma_xy, sigma_z)
    pad = PSF_SIZE // 2
    nz, ny, nx = vol.shape
    s = [sf.next_fast_len(n + 2 * pad, real=True) for n in vol.shape]
    workers = max(1, min(16, os.cpu_count() or 1))
    F = sf.rfftn(np.ascontiguousarray(vol, dtype=np.float32), s=s, workers=workers)

    def dft(k, n, half=False):
        h = np.zeros(n, np.float64)
        h[:k.size] = k
        return (np.fft.rfft(h) if half else np.fft.fft(h)).astype(np.complex64)

    hz, hy, hx = dft(kz, s[0]), dft(kxy, s[1]), dft(kxy, s[2], half=True)
    hyx = hy[:, None] * hx[None, :]

    def plane(z):
        F[z] *= hz[z] * hyx

    with _pool() as p:
        list(p.map(plane, range(s[0])))
    full = sf.irfftn(F, s=s, workers=workers)
    del F
    return np.ascontiguousarray(full[pad:pad + nz, pad:pad + ny, pad:pad + nx])


def batch_stack(i: int, shape):
    """uint16 raw stack i of the batch (its own seeded bead field, channel
    PSF i mod 3, +100 background, Poisson, +100 camera offset)."""
    _, (sxy, sz) = batch_channel(i)
    ss_truth, ss_noise = np.random.SeedSequence(1000 + i).spawn(2)
    beads = max(1, int(round(np.prod(shape) * BATCH_BEAD_DENSITY)))
    gt = ground_truth(tuple(shape), beads, np.random.default_rng(ss_truth), n_gauss=0)
    return image(blur_fft(gt, sxy, sz), ss_noise)


def make_stack(name: str, shape=None, beads=None, seed=None):
    """uint16 raw stack and the PSF of a named workload (optionally resized;
    bead count then scales with the volume; optionally another seed)."""
    c = dict(CONFIGS[name])
    if seed is not None:
        c["seed"] = int(seed)
    if shape is not None:
        vol0 = np.prod(c["shape"])
        c["beads"] = max(1, int(round(c["beads"] * np.prod(shape) / vol0))) if beads is None else beads
        c["shape"] = tuple(shape)
    ss_truth, ss_noise = np.random.SeedSequence(c["seed"]).spawn(2)
    gt = ground_truth(c["shape"], c["beads"], np.random.default_rng(ss_truth))
    raw = image(blur(gt, c["sxy"], c["sz"]), ss_noise)
    psf = gaussian_psf(c["sxy"], c["sz"])
    return raw, psf, c

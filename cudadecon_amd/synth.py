"""Deterministic synthetic light-sheet stacks (beads, smooth background,
Gaussian PSF blur, Poisson noise, camera offset, uint16).

Every stack is a function of its seed (numpy PCG64), so the same input can
be regenerated on any host.  Sizes and parameters of the named workloads:

* ``beads_64x128x128``  64 x 128 x 128, 300 beads, PSF sigma_xy 1.5 / sigma_z 3.0 px,
  10 iterations, no deskew;
* ``lightsheet_128x512x512``  128 x 512 x 512, 5000 beads, PSF 1.3 / 4.0 px, deskew 31.5 deg,
  dz 0.3 um, dx 0.104 um, then rotation, 15 iterations;
* ``beads_256x1024x1024``  256 x 1024 x 1024 plain RL-BA (the throughput workload).
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

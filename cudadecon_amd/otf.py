"""Radially averaged OTF generation and OTF TIFF files (host side).

``make_otf`` restates the reference's ``makeOTF`` for the rotationally
averaged output (src/radialft_interface.cpp:95-196 with its helpers at
:199-518): bead centre by parabolic peak fit with the reference's +0.6 on z
(:237-239), background from the in-focus rows above yc-20 (:249-254), R2C of
the PSF, phase shift of the centre to the origin (:286-319), averaging over
rint(kr) bins with kz Hermitian symmetrisation (:321-381), kr = 0 by linear
extrapolation over kr = 1..interpkr with the retry rule (:175-185, :415-443),
and normalisation to max |OTF| = 1 (:488-504).  The FFT is numpy's pocketfft
in float32 (the reference uses FFTW FP32).

The result is a complex64 table (nr, nzotf) = (nx//2+1, nz), kz fastest,
i.e. the layout of the OTF TIFF (width 2*nzotf floats x height nr).
"""
from __future__ import annotations

import struct

import numpy as np


def _fitparabola(a1, a2, a3):
    # src/radialft_interface.cpp:262-280
    slope = np.float32(0.5) * (np.float32(a3) - np.float32(a1))
    curve = (np.float32(a3) + np.float32(a1)) - np.float32(2) * np.float32(a2)
    if curve == 0:
        return np.float32(0)
    peak = -slope / curve
    if peak > 1.5 or peak < -1.5:
        return np.float32(0)
    return np.float32(peak)


def center_and_background(psf: np.ndarray):
    """determine_center_and_background (src/radialft_interface.cpp:199-255)."""
    nz, ny, nx = psf.shape
    flat = psf.reshape(-1)
    # first strict maximum in (z, y, x) scan order, starting from 0
    maxval = np.float32(0)
    mk = mi = mj = 0
    idx = int(np.argmax(flat))
    if flat[idx] > 0:
        maxval = flat[idx]
        mk, rem = divmod(idx, ny * nx)
        mi, mj = divmod(rem, nx)
    km, kp = (mk - 1) % nz, (mk + 1) % nz
    im, ip = (mi - 1) % ny, (mi + 1) % ny
    jm, jp = (mj - 1) % nx, (mj + 1) % nx
    zc = np.float32(mk + _fitparabola(psf[km, mi, mj], maxval, psf[kp, mi, mj]))
    zc = np.float32(zc + np.float32(0.6))
    yc = np.float32(mi + _fitparabola(psf[mk, im, mj], maxval, psf[mk, ip, mj]))
    xc = np.float32(mj + _fitparabola(psf[mk, mi, jm], maxval, psf[mk, mi, jp]))
    infocus = int(np.floor(zc))
    nrows = 0
    s = 0.0
    i = 0
    while i < yc - 20:
        s += float(np.sum(psf[infocus, i, :], dtype=np.float64))
        i += 1
        nrows += 1
    denom = (float(yc) - 20.0) * nx
    bg = np.float32(s / denom) if denom != 0 else np.float32(np.nan)
    return (xc, yc, zc), bg


def _fixorigin(avg: np.ndarray, nz: int, kx2: int) -> bool:
    # src/radialft_interface.cpp:415-443 (avg is (nr, nz) complex64, modified in place)
    mean_kx = (kx2 + 1) / 2.0
    xs = np.arange(1, kx2 + 1, dtype=np.float64)
    for z in range(nz):
        vals = avg[1:kx2 + 1, z].astype(np.complex128)
        mean_val = vals.sum() / kx2
        num = np.sum((xs - mean_kx) * (vals - mean_val))
        den = np.sum((xs - mean_kx) * (xs - mean_kx))
        slope = num / den
        avg[0, z] = np.complex64(mean_val - slope * mean_kx)
        if z == 0 and abs(complex(avg[0, z])) <= abs(complex(avg[1, z])):
            return False
    return True


def make_otf(psf: np.ndarray, interpkr: int = 10, background: float | None = None) -> np.ndarray:
    """Rotationally averaged OTF of a (nz, ny, nx) PSF stack (makeOTF with
    b3Dout = false and no cleanup).  Returns complex64 (nx//2+1, nz)."""
    psf = np.ascontiguousarray(psf, dtype=np.float32)
    nz, ny, nx = psf.shape
    (xc, yc, zc), est_bg = center_and_background(psf)
    bg = np.float32(est_bg if background is None else background)
    img = psf - bg
    bands = np.fft.rfftn(img.astype(np.float32)).astype(np.complex64)  # (nz, ny, nx//2+1)
    # shift_center (src/radialft_interface.cpp:286-319)
    kz = np.arange(nz)
    kz = np.where(kz > nz // 2, kz - nz, kz)
    ky = np.arange(ny)
    ky = np.where(ky > ny // 2, ky - ny, ky)
    kx = np.arange(nx // 2 + 1)
    dphiz = np.float32(2 * np.pi * float(zc) / nz)
    dphiy = np.float32(2 * np.pi * float(yc) / ny)
    dphix = np.float32(2 * np.pi * float(xc) / nx)
    phi = (dphiz * kz.astype(np.float32))[:, None, None] + (dphiy * ky.astype(np.float32))[None, :, None] + \
        (dphix * kx.astype(np.float32))[None, None, :]
    phi = phi.astype(np.float32)
    bands = (bands * (np.cos(phi) + 1j * np.sin(phi)).astype(np.complex64)).astype(np.complex64)
    # radialft (src/radialft_interface.cpp:321-381)
    nr = nx // 2 + 1
    rdist = np.sqrt((kx[None, :] ** 2 + ky[:, None] ** 2).astype(np.float32))  # (ny, nxh)
    ok = rdist < nr
    rbin = np.floor(rdist + np.float32(0.5)).astype(np.int64)  # rint of a non-negative value
    acc = np.zeros((nr, nz), np.complex128)
    cnt = np.zeros((nr, nz), np.int64)
    for z in range(nz):
        b = bands[z]
        sel = ok & (rbin * nz + z < nr * nz)
        np.add.at(acc[:, z], rbin[sel], b[sel].astype(np.complex128))
        np.add.at(cnt[:, z], rbin[sel], 1)
    avg = np.zeros((nr, nz), np.complex64)
    nzc = cnt > 0
    avg[nzc] = (acc[nzc] / cnt[nzc]).astype(np.complex64)
    for r in range(nr):
        avg[r, 0] = np.complex64(avg[r, 0].real)
        for k in range(1, nz // 2 + 1):
            v = (avg[r, k] + np.conj(avg[r, nz - k])) / np.float32(2)
            avg[r, k] = v
            avg[r, nz - k] = np.conj(v)
    if interpkr > 0:
        k2 = interpkr
        while not _fixorigin(avg, nz, k2):
            k2 -= 1
            if k2 < 4:
                raise RuntimeError("#pixels < 4 used in kr=0 extrapolation")
    avg = avg / np.float32(np.abs(avg).max())
    return avg.astype(np.complex64)


def write_tiff_f32(path: str, data: np.ndarray):
    """Uncompressed little-endian float32 TIFF, one page per leading index of a
    3-D array (2-D arrays are one page)."""
    a = np.ascontiguousarray(data, dtype="<f4")
    if a.ndim == 2:
        a = a[None]
    depth, height, width = a.shape
    page = width * height * 4
    ne = 10
    ifd_bytes = 2 + 12 * ne + 4
    out = bytearray(b"II" + struct.pack("<HI", 42, 8))
    pos = 8
    for z in range(depth):
        data_off = pos + ifd_bytes
        nxt = data_off + page if z + 1 < depth else 0
        ents = [(256, 4, width), (257, 4, height), (258, 3, 32), (259, 3, 1), (262, 3, 1), (273, 4, data_off),
                (277, 3, 1), (278, 4, height), (279, 4, page), (339, 3, 3)]
        out += struct.pack("<H", ne)
        for tag, typ, val in ents:
            if typ == 3:
                out += struct.pack("<HHIHH", tag, typ, 1, val, 0)
            else:
                out += struct.pack("<HHII", tag, typ, 1, val)
        out += struct.pack("<I", nxt)
        out += a[z].tobytes()
        pos = data_off + page
    with open(path, "wb") as f:
        f.write(out)


def write_otf_tiff(path: str, otf: np.ndarray):
    """OTF file in the reference's layout: an RA table (nr, nzotf) complex64
    becomes width 2*nzotf x height nr; a 3-D table (nzotf, nyotf, nxotf)
    becomes width 2*nxotf x height nyotf x nzotf pages."""
    t = np.ascontiguousarray(otf, dtype=np.complex64)
    f = t.view(np.float32)
    if t.ndim == 2:
        write_tiff_f32(path, f.reshape(t.shape[0], 2 * t.shape[1]))
    else:
        write_tiff_f32(path, f.reshape(t.shape[0], t.shape[1], 2 * t.shape[2]))

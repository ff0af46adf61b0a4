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


def _cleanup(avg: np.ndarray, nx: int, nz: int, dkr: float, dkz: float, lambdanm: int, krmax_user: int, NA: float,
             NIMM: float):
    # src/radialft_interface.cpp:383-411 (in place; float arithmetic)
    f = np.float32
    lamda = f(lambdanm * 0.001)
    sinalpha = f(f(NA) / f(NIMM))
    cosalpha = f(np.cos(np.arcsin(np.float64(sinalpha))))
    krmax = f(f(2 * NA) / lamda)
    if krmax_user * f(dkr) < krmax and krmax_user != 0:
        krmax = f(krmax_user * f(dkr))
    for ix in range(nx // 2 + 1):
        kr = f(ix * f(dkr))
        if kr <= krmax:
            beta = f(np.arcsin(np.float64(f(f(NA) - kr * lamda) / f(NIMM))))
            kzedge = f((f(NIMM) / lamda) * f(np.float64(np.cos(np.float64(beta))) - cosalpha))
            kzstart = int(np.rint(np.float64(f(kzedge / f(dkz)) + 1)))
            kzend = nz - kzstart
            if kzstart <= kzend:
                avg[ix, max(kzstart, 0):kzend + 1] = 0
        else:
            avg[ix, :] = 0


def _centred_bands(psf: np.ndarray, background):
    """R2C of PSF - background with the bead centre moved to the origin
    (determine_center_and_background + shift_center)."""
    psf = np.ascontiguousarray(psf, dtype=np.float32)
    nz, ny, nx = psf.shape
    (xc, yc, zc), est_bg = center_and_background(psf)
    bg = np.float32(est_bg if background is None else background)
    img = psf - bg
    bands = np.fft.rfftn(img.astype(np.float32)).astype(np.complex64)  # (nz, ny, nx//2+1)
    return bands, (xc, yc, zc)


def make_otf_3d(psf: np.ndarray, interpkr: int = 10, background: float | None = None) -> np.ndarray:
    """3-D OTF of a (nz, ny, nx) PSF stack (makeOTF with b3Dout = true,
    src/radialft_interface.cpp:157-162): the centred half spectrum with the
    (0, 0, 0) value extrapolated from the +-45 degree lines (:445-486) and
    divided by its max |.| (:506-518).  Returns complex64 (nz, ny, nx//2+1)."""
    bands, c = _centred_bands(psf, background)
    nz, ny, nh = bands.shape
    bands = _shift(bands, *c, psf.shape[2])
    if interpkr > 1:
        proj = np.zeros(nh, np.complex64)
        for z in range(nz):
            for x in range(1, min(nh, ny)):
                proj[x] += bands[z, x, x]
                proj[x] += bands[z, ny - x, x]
        xs = np.arange(1, interpkr + 1, dtype=np.float64)
        vals = proj[1:interpkr + 1].astype(np.complex128)
        mean_val = vals.sum() / interpkr
        mean_kx = (interpkr + 1) / 2.0
        slope = np.sum((xs - mean_kx) * (vals - mean_val)) / np.sum((xs - mean_kx) ** 2)
        bands[0, 0, 0] = np.complex64(mean_val - slope * mean_kx)
    m = np.float32(np.abs(bands).max())
    return (bands.view(np.float32) / m).view(np.complex64)


def _shift(bands, xc, yc, zc, nx):
    # shift_center (src/radialft_interface.cpp:286-319)
    nz, ny, nh = bands.shape
    kz = np.arange(nz)
    kz = np.where(kz > nz // 2, kz - nz, kz)
    ky = np.arange(ny)
    ky = np.where(ky > ny // 2, ky - ny, ky)
    kx = np.arange(nh)
    dphiz = np.float32(2 * np.pi * float(zc) / nz)
    dphiy = np.float32(2 * np.pi * float(yc) / ny)
    dphix = np.float32(2 * np.pi * float(xc) / nx)
    phi = (dphiz * kz.astype(np.float32))[:, None, None] + (dphiy * ky.astype(np.float32))[None, :, None] + \
        (dphix * kx.astype(np.float32))[None, None, :]
    phi = phi.astype(np.float32)
    return (bands * (np.cos(phi) + 1j * np.sin(phi)).astype(np.complex64)).astype(np.complex64)


def make_otf(psf: np.ndarray, interpkr: int = 10, background: float | None = None, cleanup: dict | None = None) -> np.ndarray:
    """Rotationally averaged OTF of a (nz, ny, nx) PSF stack (makeOTF with
    b3Dout = false).  cleanup: None, or dict(lambdanm, NA, NIMM, dr, dz,
    krmax) to zero the table outside the OTF support as bDoCleanup does.
    Returns complex64 (nx//2+1, nz)."""
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
    if cleanup is not None:
        c = dict(dict(lambdanm=520, NA=1.25, NIMM=1.3, dr=0.102, dz=0.102, krmax=0), **cleanup)
        dkr = np.float32(1.0 / (ny * np.float32(c["dr"])))
        dkz = np.float32(1.0 / (nz * np.float32(c["dz"])))
        _cleanup(avg, nx, nz, dkr, dkz, c["lambdanm"], c["krmax"], c["NA"], c["NIMM"])
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


def read_tiff_f32(path: str) -> np.ndarray:
    """Pages of an uncompressed 32-bit float TIFF (either byte order, any
    strip layout) as a (pages, height, width) float32 array."""
    buf = open(path, "rb").read()
    bo = {b"II": "<", b"MM": ">"}[buf[:2]]
    if struct.unpack(bo + "H", buf[2:4])[0] != 42:
        raise ValueError(f"{path}: not a TIFF file")
    size = {1: 1, 2: 1, 3: 2, 4: 4, 16: 8}
    fmt = {1: "B", 2: "B", 3: "H", 4: "I", 16: "Q"}
    pages = []
    off = struct.unpack(bo + "I", buf[4:8])[0]
    while off:
        n = struct.unpack(bo + "H", buf[off:off + 2])[0]
        tags = {}
        for i in range(n):
            e = off + 2 + 12 * i
            tag, typ, cnt = struct.unpack(bo + "HHI", buf[e:e + 8])
            sz = size[typ] * cnt
            data = buf[e + 8:e + 8 + sz] if sz <= 4 else \
                buf[struct.unpack(bo + "I", buf[e + 8:e + 12])[0]:][:sz]
            tags[tag] = struct.unpack(bo + fmt[typ] * cnt, data)
        w, h = tags[256][0], tags[257][0]
        if tags.get(258, (32,))[0] != 32 or tags.get(339, (3,))[0] != 3 or tags.get(259, (1,))[0] != 1:
            raise ValueError(f"{path}: only uncompressed 32-bit float pages are read here")
        raw = b"".join(buf[o:o + c] for o, c in zip(tags[273], tags[279]))
        pages.append(np.frombuffer(raw, dtype=bo + "f4", count=w * h).reshape(h, w))
        off = struct.unpack(bo + "I", buf[off + 2 + 12 * n:off + 6 + 12 * n])[0]
    return np.stack(pages).astype(np.float32)

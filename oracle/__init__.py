"""CPU oracle for the RL-Biggs-Andrews deconvolution path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the checker, and by ``bench.py``'s cpu_baseline leg.  The product
(``libcudaDecon.so`` and the ``cudadecon_amd`` package) never imports it.

``liboracle.so`` (built from ``oracle/oracle.cpp`` by ``oracle/Makefile``)
holds the FP64 arithmetic; this module adds ctypes bindings and restates the
host-side scalar logic of the reference:

* ``init_geometry``  - ``RL_interface_init`` sizes, deskew factor, rotation
  matrix and OTF k-scales (src/RL-Biggs-Andrews.cpp:506-655), with the
  float/double rounding points of the reference;
* ``rl_interface``   - the ``RL_interface`` / ``RichardsonLucy_GPU`` sequence
  (src/RL-Biggs-Andrews.cpp:56-450, 657-705);
* ``deskew_interface`` - ``Deskew_interface`` (src/RL-Biggs-Andrews.cpp:724-773).

Parity status of each function is listed in tests/test_oracle.py's module
docstring together with the check that pins it.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
FLT_EPS = float(np.finfo(np.float32).eps)


def build():
    """Compile liboracle.so with oracle/Makefile."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        i, d, f, l_ = C.c_int, C.c_double, C.c_float, C.c_long
        sig = {
            "orc_max_threads": (i, []),
            "orc_set_threads": (None, [i]),
            "orc_find_optimal_dimension": (C.c_uint, [C.c_uint, i]),
            "orc_fft1d_c2c_f64": (None, [_dp, i, i, i]),
            "orc_rfftn_f64": (None, [_dp, i, i, i, _dp]),
            "orc_irfftn_f64": (None, [_dp, i, i, i, _dp]),
            "orc_rfftn_f32": (None, [_fp, i, i, i, _fp]),
            "orc_irfftn_f32": (None, [_fp, i, i, i, _fp]),
            "orc_otf_interp_f64": (None, [_fp, i, i, i, i, i, i, d, d, d, _dp]),
            "orc_otf_interp_f32": (None, [_fp, i, i, i, i, i, i, f, f, f, _fp]),
            "orc_filter_f64": (None, [_dp, i, i, i, _dp, i]),
            "orc_rl_ba_f64": (None, [_dp, i, i, i, _dp, i, C.c_void_p, _dp, _dp]),
            "orc_rl_ba_f32": (None, [_fp, i, i, i, _fp, i, C.c_void_p, _fp, _dp]),
            "orc_rl_ba_f32_timed": (None, [_fp, i, i, i, _fp, i, _fp, _dp, _dp]),
            "orc_crop_u16": (None, [_u16p, i, i, i, i, i, i, _dp]),
            "orc_apodize": (None, [_dp, i, i, i, i]),
            "orc_bgsub": (None, [_dp, l_, d]),
            "orc_mean_above": (d, [_dp, l_]),
            "orc_rescale": (None, [_dp, l_, f]),
            "orc_zblend": (None, [_dp, i, i, i, i]),
            "orc_deskew": (None, [_dp, i, i, i, d, _dp, i, i, d]),
            "orc_rotate": (None, [_dp, i, i, i, _fp, _dp, i, i]),
            "orc_affine": (None, [_dp, i, i, i, _fp, _dp]),
            "orc_affine_ra": (None, [_dp, i, i, i, f, f, f, _fp, _dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# --------------------------------------------------------------------------
# FFT
# --------------------------------------------------------------------------
def fft1d(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """Complex FFT along the last axis (unnormalised inverse)."""
    x = _c(x, np.complex128)
    n = x.shape[-1]
    buf = x.reshape(-1).view(np.float64).copy()
    lib().orc_fft1d_c2c_f64(buf, n, x.size // n, int(inverse))
    return buf.view(np.complex128).reshape(x.shape)


def rfftn(v: np.ndarray) -> np.ndarray:
    """3-D R2C of a (nz, ny, nx) volume -> (nz, ny, nx//2+1), unnormalised."""
    v = _c(v, np.float64)
    nz, ny, nx = v.shape
    out = np.empty((nz, ny, nx // 2 + 1), np.complex128)
    lib().orc_rfftn_f64(v, nx, ny, nz, out.view(np.float64).reshape(-1))
    return out


def irfftn(s: np.ndarray, nx: int) -> np.ndarray:
    """3-D C2R (numpy irfftn semantics, unnormalised)."""
    s = _c(s, np.complex128)
    nz, ny, _ = s.shape
    out = np.empty((nz, ny, nx), np.float64)
    lib().orc_irfftn_f64(s.view(np.float64).reshape(-1), nx, ny, nz, out)
    return out


def rfftn_f32(v: np.ndarray) -> np.ndarray:
    v = _c(v, np.float32)
    nz, ny, nx = v.shape
    out = np.empty((nz, ny, nx // 2 + 1), np.complex64)
    lib().orc_rfftn_f32(v, nx, ny, nz, out.view(np.float32).reshape(-1))
    return out


# --------------------------------------------------------------------------
# OTF, filter, RL-BA
# --------------------------------------------------------------------------
def otf_interp(raw_otf: np.ndarray, otf_dims, grid, kscale) -> np.ndarray:
    """Interpolated OTF on the (nz, ny, nx//2+1) half grid, complex128.

    raw_otf: complex64 table in file order; otf_dims = (nxotf, nyotf, nzotf);
    grid = (nx, ny, nz); kscale = (kx, ky, kz) scale factors (float32 values).
    """
    nxotf, nyotf, nzotf = otf_dims
    nx, ny, nz = grid
    raw = _c(raw_otf, np.complex64).view(np.float32).reshape(-1)
    out = np.empty((nz, ny, nx // 2 + 1), np.complex128)
    lib().orc_otf_interp_f64(raw, nxotf, nyotf, nzotf, nx, ny, nz, float(kscale[0]), float(kscale[1]),
                             float(kscale[2]), out.view(np.float64).reshape(-1))
    return out


def filt(v: np.ndarray, H: np.ndarray, conj: bool = False) -> np.ndarray:
    """One filterGPU call: C2R(H . R2C(v / N)) (src/RLgpuImpl.cu:156-225)."""
    v = _c(v, np.float64).copy()
    _check_H(v.shape, H)
    nz, ny, nx = v.shape
    lib().orc_filter_f64(v, nx, ny, nz, _c(H, np.complex128).view(np.float64).reshape(-1), int(conj))
    return v


def _check_H(shape, H):
    """The half-spectrum OTF must match the volume (the C routines index it
    by the volume's shape)."""
    nz, ny, nx = shape
    if tuple(H.shape) != (nz, ny, nx // 2 + 1):
        raise ValueError(f"OTF shape {tuple(H.shape)} does not match the volume {tuple(shape)}: "
                         f"expected {(nz, ny, nx // 2 + 1)}")


def rl_ba(raw: np.ndarray, H: np.ndarray, n_iter: int, seed: np.ndarray | None = None):
    """FP64 RL-Biggs-Andrews; returns (X, lambdas)."""
    raw = _c(raw, np.float64)
    _check_H(raw.shape, H)
    nz, ny, nx = raw.shape
    X = np.empty_like(raw)
    lam = np.zeros(max(n_iter, 1), np.float64)
    Hd = _c(H, np.complex128).view(np.float64).reshape(-1)
    sp = None
    if seed is not None:
        seed = _c(seed, np.float64)
        sp = seed.ctypes.data
    lib().orc_rl_ba_f64(raw, nx, ny, nz, Hd, int(n_iter), sp, X, lam)
    return X, lam[:n_iter]


def rl_ba_f32(raw: np.ndarray, H: np.ndarray, n_iter: int, seed: np.ndarray | None = None):
    """FP32 flavour (timed CPU baseline); returns (X, lambdas)."""
    raw = _c(raw, np.float32)
    _check_H(raw.shape, H)
    nz, ny, nx = raw.shape
    X = np.empty_like(raw)
    lam = np.zeros(max(n_iter, 1), np.float64)
    Hf = _c(H, np.complex64).view(np.float32).reshape(-1)
    sp = None
    if seed is not None:
        seed = _c(seed, np.float32)
        sp = seed.ctypes.data
    lib().orc_rl_ba_f32(raw, nx, ny, nz, Hf, int(n_iter), sp, X, lam)
    return X, lam[:n_iter]


def rl_ba_f32_timed(raw: np.ndarray, H: np.ndarray, n_iter: int):
    """FP32 flavour with the wall time of each iteration; returns (X, lambdas, seconds)."""
    raw = _c(raw, np.float32)
    _check_H(raw.shape, H)
    nz, ny, nx = raw.shape
    X = np.empty_like(raw)
    lam = np.zeros(max(n_iter, 1), np.float64)
    sec = np.zeros(max(n_iter, 1), np.float64)
    lib().orc_rl_ba_f32_timed(raw, nx, ny, nz, _c(H, np.complex64).view(np.float32).reshape(-1), int(n_iter), X,
                              lam, sec)
    return X, lam[:n_iter], sec[:n_iter]


# --------------------------------------------------------------------------
# preprocessing and geometry
# --------------------------------------------------------------------------
def find_optimal_dimension(n: int, step: int = -1) -> int:
    return int(lib().orc_find_optimal_dimension(int(n), int(step)))


def apodize(img, napodize):
    img = _c(img, np.float64).copy()
    nz, ny, nx = img.shape
    lib().orc_apodize(img, nx, ny, nz, int(napodize))
    return img


def bgsub(img, bg):
    img = _c(img, np.float64).copy()
    lib().orc_bgsub(img.reshape(-1), img.size, float(bg))
    return img


def mean_above(img) -> float:
    img = _c(img, np.float64)
    return float(lib().orc_mean_above(img.reshape(-1), img.size))


def rescale(img, scale):
    img = _c(img, np.float64).copy()
    lib().orc_rescale(img.reshape(-1), img.size, float(np.float32(scale)))
    return img


def zblend(img, nzblend):
    img = _c(img, np.float64).copy()
    nz, ny, nx = img.shape
    lib().orc_zblend(img, nx, ny, nz, int(nzblend))
    return img


def deskew(img, deskew_factor, nx_out, extra_shift=0, pad_val=0.0):
    img = _c(img, np.float64)
    nz, ny, nx = img.shape
    out = np.empty((nz, ny, nx_out), np.float64)
    lib().orc_deskew(img, nx, ny, nz, float(deskew_factor), out, int(nx_out), int(extra_shift), float(pad_val))
    return out


def rotate(img, rot, nx_out, nz_out):
    img = _c(img, np.float64)
    nz, ny, nx = img.shape
    out = np.empty((nz_out, ny, nx_out), np.float64)
    lib().orc_rotate(img, nx, ny, nz, _c(rot, np.float32), out, int(nx_out), int(nz_out))
    return out


def affine(img, mat):
    img = _c(img, np.float64)
    nz, ny, nx = img.shape
    out = np.empty_like(img)
    lib().orc_affine(img, nx, ny, nz, _c(mat, np.float32).reshape(-1), out)
    return out


def affine_ra(img, dx, dy, dz, mat):
    img = _c(img, np.float64)
    nz, ny, nx = img.shape
    out = np.empty_like(img)
    lib().orc_affine_ra(img, nx, ny, nz, float(dx), float(dy), float(dz), _c(mat, np.float32).reshape(-1), out)
    return out


# --------------------------------------------------------------------------
# RL_interface_init restatement (src/RL-Biggs-Andrews.cpp:506-655)
# --------------------------------------------------------------------------
f32 = np.float32


def otf_dims(otf_width: int, otf_height: int, otf_depth: int, dr_psf, dz_psf):
    """determine_OTF_dimensions (src/RL-Biggs-Andrews.cpp:506-529), float math."""
    dr_psf, dz_psf = f32(dr_psf), f32(dz_psf)
    if otf_depth > 1:
        nxo, nyo, nzo = otf_width // 2, otf_height, otf_depth
    else:
        nxo, nyo, nzo = otf_height, 1, otf_width // 2
    dkx = f32(1) / (f32((nxo - 1) * 2) * dr_psf)
    dky = f32(1) / (f32(nyo) * dr_psf) if nyo > 1 else dkx
    dkz = f32(1) / (f32(nzo) * dz_psf)
    return (nxo, nyo, nzo), (dkx, dky, dkz)


def init_geometry(nx, ny, nz, dr, dz, dr_psf, dz_psf, deskew_angle, rotation_angle, output_width,
                  skewed_decon, otf_shape):
    """Scalars computed by RL_interface_init.

    otf_shape = (width, height, depth) of the OTF TIFF (width counts floats).
    Skewed decon follows the cudaDecon CLI (src/linearDecon.cpp:643,671,776-777,
    895-905): the FFT grid keeps the (optimal) raw width and the deskew runs
    after the iterations.
    """
    dr, dz = f32(dr), f32(dz)
    dr_psf, dz_psf = f32(dr_psf), f32(dz_psf)
    deskew_angle, rotation_angle = f32(deskew_angle), f32(rotation_angle)
    g = {}
    g["output_ny"] = find_optimal_dimension(ny)
    g["output_nz"] = find_optimal_dimension(nz)
    has_deskew = abs(float(deskew_angle)) > 0.0
    g["output_nx"] = nx if (has_deskew and not skewed_decon) else find_optimal_dimension(nx)
    g["crop"] = (g["output_nx"], g["output_ny"], g["output_nz"]) != (nx, ny, nz)
    if skewed_decon:
        dz_psf = f32(float(dz_psf) * abs(math.sin(float(deskew_angle) * math.pi / 180.0)))
    dims, dks = otf_dims(*otf_shape, dr_psf, dz_psf)
    g["otf_dims"], g["dk_otf"] = dims, dks
    deskew_factor = 0.0
    xdim = g["output_nx"]
    if has_deskew:
        if deskew_angle < 0:
            deskew_angle = f32(float(deskew_angle) + 180.0)
        c = math.cos(float(deskew_angle) * math.pi / 180.0)
        deskew_factor = c * float(dz) / float(dr)
        if output_width == 0:
            xdim = int(float(xdim) + math.floor(g["output_nz"] * float(dz) * abs(c) / float(dr)))
        else:
            xdim = int(output_width)
        xdim = find_optimal_dimension(xdim)
        dz = f32(float(dz) * abs(math.sin(float(deskew_angle) * math.pi / 180.0)))
    g["deskew_factor"] = deskew_factor
    g["deskewed_nx"] = xdim
    rot = None
    if abs(float(rotation_angle)) > 0.0:
        ra = f32(float(rotation_angle) * math.pi / 180.0)
        stretch = f32(dr / dz)
        cs, sn = np.cos(ra, dtype=f32), np.sin(ra, dtype=f32)
        rot = np.array([cs * stretch, sn * stretch, -sn, cs], np.float32)
    g["rot"] = rot
    fft_nx = g["output_nx"] if skewed_decon else xdim
    g["fft_dims"] = (fft_nx, g["output_ny"], g["output_nz"])
    dkx = f32(1.0 / float(dr * f32(fft_nx)))
    dky = f32(1.0 / float(dr * f32(g["output_ny"])))
    dkz = f32(1.0 / float(dz * f32(g["output_nz"])))
    g["kscale"] = (dkx / dks[0], dky / dks[1], dkz / dks[2])
    # sizes after the optional deskew-after-decon and rotation
    post_nx = xdim
    nx_o, nz_o = post_nx, g["output_nz"]
    if rot is not None:
        nz_o = int(f32(g["output_nz"]) * rot[3] / rot[0])
        nx_o = int(f32(post_nx) * rot[3] + f32(g["output_nz"]) * rot[2] * rot[2] / abs(rot[1]))
    g["out_dims"] = (nx_o, g["output_ny"], nz_o)
    return g


class Session:
    """Process-global state of the reference library (bleach statics at
    src/RL-Biggs-Andrews.cpp:46-47 plus the init scalars)."""

    def __init__(self, geom, raw_otf):
        self.g = geom
        self.raw_otf = raw_otf
        self.first = True
        self.i0 = 0.0
        self._H = None

    def H(self):
        if self._H is None:
            self._H = otf_interp(self.raw_otf, self.g["otf_dims"], self.g["fft_dims"], self.g["kscale"])
        return self._H


def rl_interface(s: Session, raw_u16: np.ndarray, background, do_rescale, save_deskewed, n_iters, extra_shift,
                 napodize=0, nzblend=0, pad_val=0.0, dup_rev_stack=False, skewed_decon=False, seed=None):
    """RL_interface -> RichardsonLucy_GPU (src/RL-Biggs-Andrews.cpp:56-450,657-705).

    Returns (result, raw_deskewed or None, lambdas).  Differences from the
    reference are the documented fixes: the deskew-only call (n_iters == 0 with
    save_deskewed) returns the deskewed raw as result instead of uninitialised
    memory; z blend is the intended taper; dupRevStack has no effect on the
    result (the library path's FFT plans and kernel sizes ignore the doubled
    half, src/RL-Biggs-Andrews.cpp:193-205 vs :622-642).
    """
    g = s.g
    nz, ny, nx = raw_u16.shape
    onx, ony, onz = g["output_nx"], g["output_ny"], g["output_nz"]
    X = np.empty((onz, ony, onx), np.float64)
    lib().orc_crop_u16(_c(raw_u16, np.uint16), nx, ny, nz, onx, ony, onz, X)
    deskewed = None
    lambdas = np.zeros(0)
    dfac = g["deskew_factor"]
    rot = g["rot"]
    if n_iters > 0 or (save_deskewed and dfac != 0) or rot is not None:
        X = apodize(X, napodize)
        X = bgsub(X, background)
        inten = mean_above(X)
        if do_rescale:
            if s.first:
                s.i0, s.first = inten, False
            else:
                X = rescale(X, s.i0 / inten)
        if ((not skewed_decon) or (save_deskewed and n_iters == 0)) and abs(dfac) > 0:
            X = deskew(X, dfac, g["deskewed_nx"], extra_shift, pad_val)
            if save_deskewed:
                deskewed = X.copy()
                if n_iters == 0:
                    return X, deskewed, lambdas
        if nzblend > 0:
            X = zblend(X, nzblend)
        if n_iters > 0:
            X, lambdas = rl_ba(X, s.H(), n_iters, seed)
    if skewed_decon and abs(dfac) > 0:
        X = deskew(X, dfac, g["deskewed_nx"], extra_shift, 0.0)
    if rot is not None:
        znz, _, xnx = X.shape
        nz_o = int(f32(znz) * rot[3] / rot[0])
        nx_o = int(f32(xnx) * rot[3] + f32(znz) * rot[2] * rot[2] / abs(rot[1]))
        X = rotate(X, rot, nx_o, nz_o)
    return X, deskewed, lambdas


def deskew_interface(raw, dz, dr, deskew_angle, output_width, extra_shift, pad_val=0.0, output_nz=0):
    """Deskew_interface (src/RL-Biggs-Andrews.cpp:724-773); output_nz is the
    library's global from the last RL_interface_init (0 if none)."""
    raw = _c(raw, np.float64)
    nz, ny, nx = raw.shape
    dz, dr, ang = f32(dz), f32(dr), f32(deskew_angle)
    fac = 0.0
    xdim = nx
    if abs(float(ang)) > 0:
        if ang < 0:
            ang = f32(float(ang) + 180.0)
        c = math.cos(float(ang) * math.pi / 180.0)
        fac = c * float(dz) / float(dr)
        if output_width == 0:
            xdim = int(float(xdim) + math.floor(output_nz * float(dz) * abs(c) / float(dr)) / 4.0)
        else:
            xdim = int(output_width)
    return deskew(raw, fac, xdim, extra_shift, pad_val)


def camcor(raw_u16: np.ndarray, cam: np.ndarray):
    """camcor_kernel (src/camcor.cu:80-108) in FP64: out = v - off - 0.9 a
    (1 - exp(-b (v_prev - off))) with v_prev the plane z-1 (plane nz-1 for
    z = 0), truncated toward zero and saturated to [0, 65535].  cam is
    (3, ny, nx) = (a, b, offset).  Returns (uint16 result, FP64 value before
    the conversion)."""
    v = raw_u16.astype(np.float64)
    prev = np.roll(v, 1, axis=0)
    a, b, off = (cam[i].astype(np.float64)[None] for i in range(3))
    r = v - off - 0.9 * a * (1.0 - np.exp(-b * (prev - off)))
    out = np.clip(np.trunc(r), 0, 65535).astype(np.uint16)
    return out, r


def wiener(raw: np.ndarray, raw_otf: np.ndarray, otf_dims, kscale, dkxy, rcutoff, wiener_const, background=0.0):
    """wienerfilter + its call (src/linearDecon.cpp:116-168, 1116-1128) in
    FP64: G = R2C(raw - background); for kr = sqrt(kx^2 dkx^2 + ky^2 dky^2) <=
    rcutoff, G <- conj(H) G / (|H|^2 + w^2) (1 - kr / rcutoff), else 0;
    result C2R(G) / N.  kx runs over the half axis 0..nx/2, ky signed."""
    nz, ny, nx = raw.shape
    H = otf_interp(raw_otf, otf_dims, (nx, ny, nz), kscale)
    G = rfftn(raw.astype(np.float64) - background)
    kx = np.arange(nx // 2 + 1, dtype=np.float64)
    ky = np.arange(ny, dtype=np.float64)
    ky = np.where(ky > ny // 2, ky - ny, ky)
    kr = np.sqrt((kx[None, :] * dkxy[0]) ** 2 + (ky[:, None] * dkxy[1]) ** 2)
    w2 = float(wiener_const) ** 2
    apo = np.where(kr <= rcutoff, 1.0 - kr / rcutoff, 0.0)[None]
    G = np.conj(H) * G / (np.abs(H) ** 2 + w2) * apo
    return irfftn(G, nx) / raw.size

"""cudadecon_amd - MI355X (gfx950) build of cudaDecon's RL-Biggs-Andrews path.

The product is ``libcudaDecon.so`` in this directory: a C-ABI library whose
exported symbols are those of the reference library (src/linearDecon.h:243-317)
plus the ``cudaDecon_*`` extensions declared in ``include/cudaDecon.h``.  This
module binds them with ctypes the way pycudadecon binds the reference
library, and adds thin numpy conveniences.

The library is loaded lazily; when it is missing every call raises
``RuntimeError`` (there is no CPU fallback).  Build it with
``python -c "import __graft_entry__ as g; g.build()"`` from the repository
root, or ``make -C cudadecon_amd/csrc``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

__all__ = [
    "lib", "lib_path", "RLEngine", "rl_init", "rl_decon", "rl_cleanup", "deskew", "affine", "affine_ra",
    "fft_r2c", "fft_c2r", "filter_volume", "output_shape", "last_error", "Batch", "WarmNet",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib_path() -> str:
    return os.path.join(_HERE, "libcudaDecon.so")


_f = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u16 = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")

# Exported symbols and their ctypes signatures (include/cudaDecon.h).
SIGNATURES = {
    # reference API
    "RL_interface_init": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                    C.c_float, C.c_int, C.c_bool, C.c_char_p]),
    "RL_interface": (C.c_int, [_u16, C.c_int, C.c_int, C.c_int, _f, C.c_void_p, C.c_float, C.c_bool, C.c_bool,
                               C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_bool, C.c_bool]),
    "Deskew_interface": (C.c_int, [_f, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, _f, C.c_int,
                                   C.c_int, C.c_float]),
    "Affine_interface": (C.c_int, [_f, C.c_int, C.c_int, C.c_int, _f, _f]),
    "Affine_interface_RA": (C.c_int, [_f, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, _f, _f]),
    "camcor_interface_init": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_void_p]),
    "camcor_interface": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]),
    "RL_cleanup": (None, []),
    "cuda_reset": (None, []),
    "get_output_nx": (C.c_uint, []),
    "get_output_ny": (C.c_uint, []),
    "get_output_nz": (C.c_uint, []),
    "get_version": (C.c_char_p, []),
    # extensions
    "cudaDecon_last_error": (C.c_char_p, []),
    "cudaDecon_fft_supported": (C.c_int, [C.c_int]),
    "cudaDecon_find_optimal_dimension": (C.c_uint, [C.c_uint, C.c_int]),
    "cudaDecon_deskew_width": (C.c_uint, [C.c_int, C.c_float, C.c_float, C.c_float, C.c_int]),
    "cudaDecon_set_flat_start": (C.c_int, [C.c_int, C.c_float]),
    "cudaDecon_get_session_dims": (C.c_uint, [C.c_int]),
    "cudaDecon_get_session_kscale": (C.c_int, [_f]),
    "cudaDecon_fft_r2c": (C.c_int, [_f, C.c_int, C.c_int, C.c_int, _f]),
    "cudaDecon_fft_c2r": (C.c_int, [_f, C.c_int, C.c_int, C.c_int, _f]),
    "cudaDecon_filter": (C.c_int, [_f, C.c_int, C.c_int, C.c_int, _f, C.c_int, C.c_int, C.c_int, C.c_float,
                                   C.c_float, C.c_float, C.c_int]),
    "cudaDecon_rl_create": (C.c_void_p, [C.c_int, C.c_int, C.c_int, _f, C.c_int, C.c_int, C.c_int, C.c_float,
                                         C.c_float, C.c_float]),
    "cudaDecon_rl_create_stacks": (C.c_void_p, [C.c_int, C.c_int, C.c_int, C.c_int, _f, C.c_int, C.c_int, C.c_int,
                                                C.c_float, C.c_float, C.c_float]),
    "cudaDecon_rl_stacks": (C.c_int, [C.c_void_p]),
    "cudaDecon_rl_set_raw_stack": (C.c_int, [C.c_void_p, C.c_int, _f]),
    "cudaDecon_rl_get_stack": (C.c_int, [C.c_void_p, C.c_int, _f]),
    "cudaDecon_rl_set_raw": (C.c_int, [C.c_void_p, _f]),
    "cudaDecon_rl_set_seed": (C.c_int, [C.c_void_p, _f]),
    "cudaDecon_rl_iterate": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "cudaDecon_rl_get": (C.c_int, [C.c_void_p, _f]),
    "cudaDecon_rl_timing": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "cudaDecon_rl_geometry": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "cudaDecon_rl_copy_probe": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "cudaDecon_rl_set_raw_block": (C.c_int, [C.c_void_p, _f, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                             C.c_float]),
    "cudaDecon_rl_get_block": (C.c_int, [C.c_void_p, _f, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "cudaDecon_rl_stats": (C.c_int, [C.c_void_p, _d]),
    "cudaDecon_rl_destroy": (None, [C.c_void_p]),
    "cudaDecon_batch_create": (C.c_void_p, [C.c_int, C.c_void_p, _f, C.c_int, C.c_int, C.c_int, _f, C.c_int]),
    "cudaDecon_batch_set_raw": (C.c_int, [C.c_void_p, C.c_int, _f]),
    "cudaDecon_batch_set_raw_block": (C.c_int, [C.c_void_p, C.c_int, _f, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                                C.c_int, C.c_float]),
    "cudaDecon_batch_run": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "cudaDecon_batch_get": (C.c_int, [C.c_void_p, C.c_int, _f]),
    "cudaDecon_batch_destroy": (None, [C.c_void_p]),
    "cudaDecon_batch_create_tables": (C.c_void_p, [C.c_int, C.c_void_p, C.c_int, _f, C.c_void_p, C.c_void_p, _f,
                                                   C.c_int]),
    "cudaDecon_batch_lambdas": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _d]),
    "cudaDecon_batch_set_grid_cap": (C.c_int, [C.c_void_p, C.c_long]),
    "cudaDecon_batch_launches": (C.c_int, [C.c_void_p]),
    "cudaDecon_batch_set_partitions": (C.c_int, [C.c_void_p, C.c_int]),
    "cudaDecon_rl_snapshot": (C.c_int, [C.c_void_p]),
    "cudaDecon_rl_set_raw_u16": (C.c_int, [C.c_void_p, _u16, C.c_float]),
    "cudaDecon_rl_deskew": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, C.c_float,
                                      C.c_void_p]),
    "cudaDecon_rl_affine": (C.c_int, [C.c_void_p, _f]),
    "cudaDecon_rl_post_dims": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "cudaDecon_rl_post_get": (C.c_int, [C.c_void_p, _f]),
    "cudaDecon_rl_post_stats": (C.c_int, [C.c_void_p, _d]),
    "cudaDecon_rl_post_get_block": (C.c_int, [C.c_void_p, C.c_int, _f, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                              C.c_int]),
    "cudaDecon_rl_buffer_stats": (C.c_int, [C.c_void_p, C.c_int, _d]),
    "cudaDecon_autotune": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_int]),
    "cudaDecon_autotune_clear": (C.c_int, []),
    "cudaDecon_fft_variant": (C.c_int, [C.c_int, C.c_void_p, C.c_int]),
    "cudaDecon_wiener": (C.c_int, [_f, C.c_int, C.c_int, C.c_int, _f, C.c_int, C.c_int, C.c_float, C.c_float,
                                   C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float]),
    "cudaDecon_rl_distance": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "cudaDecon_net_create": (C.c_void_p, [C.c_ulonglong]),
    "cudaDecon_net_num_params": (C.c_long, []),
    "cudaDecon_net_get_params": (C.c_int, [C.c_void_p, _f]),
    "cudaDecon_net_set_params": (C.c_int, [C.c_void_p, _f]),
    "cudaDecon_net_forward": (C.c_int, [C.c_void_p, _f, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _f]),
    "cudaDecon_net_fetch": (C.c_long, [C.c_void_p, C.c_int, C.c_void_p]),
    "cudaDecon_net_loss_grad": (C.c_int, [C.c_void_p, _f, _f, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                          C.c_void_p]),
    "cudaDecon_net_step": (C.c_int, [C.c_void_p, _f, _f, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                                     C.c_void_p]),
    "cudaDecon_net_add_samples": (C.c_int, [C.c_void_p, _f, _f, C.c_int, C.c_int, C.c_int, C.c_int]),
    "cudaDecon_net_train": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_ulonglong, C.c_void_p,
                                      C.c_void_p]),
    "cudaDecon_net_destroy": (None, [C.c_void_p]),
    "cudaDecon_rl_warm_start": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "cudaDecon_device_sync": (C.c_int, []),
    "cudaDecon_device_malloc": (C.c_void_p, [C.c_size_t]),
    "cudaDecon_device_free": (C.c_int, [C.c_void_p]),
    "cudaDecon_device_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    "cudaDecon_rl_run_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, _f, C.c_int, C.c_int,
                                           C.c_int, C.c_float, C.c_float, C.c_float, C.c_int, C.c_void_p]),
}

PASS_NAMES = ["x_r2c", "y_fwd", "z_filter", "y_inv", "x_lrcore", "x_update", "lambda"]


def lib():
    """Load libcudaDecon.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; the HIP extension is required (no CPU fallback)")
        L = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def last_error() -> str:
    return lib().cudaDecon_last_error().decode()


def _check(rc, what):
    if not rc:
        raise RuntimeError(f"{what} failed: {last_error()}")


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ---------------------------------------------------------------------------
# reference-style session API (pycudadecon's RL_interface usage)
# ---------------------------------------------------------------------------
def rl_init(rawdata_shape, otfpath, dzdata=0.5, dxdata=0.1, dzpsf=0.1, dxpsf=0.1, deskew=0.0, rotate=0.0,
            width=0, skewed_decon=False):
    """RL_interface_init for a (nz, ny, nx) stack; returns nothing, raises on failure."""
    nz, ny, nx = rawdata_shape
    rc = lib().RL_interface_init(int(nx), int(ny), int(nz), float(dxdata), float(dzdata), float(dxpsf),
                                 float(dzpsf), float(deskew), float(rotate), int(width), bool(skewed_decon),
                                 os.fsencode(otfpath))
    _check(rc, "RL_interface_init")


def output_shape():
    L = lib()
    return (int(L.get_output_nz()), int(L.get_output_ny()), int(L.get_output_nx()))


def rl_decon(im, background=100.0, n_iters=10, shift=0, save_deskewed=False, rescale=False, napodize=0,
             nz_blend=0, pad_val=0.0, dup_reversed_z=False, skewed_decon=False):
    """RL_interface on a uint16 (nz, ny, nx) stack; returns result (and the
    deskewed raw when save_deskewed)."""
    im = np.ascontiguousarray(im, dtype=np.uint16)
    nz, ny, nx = im.shape
    out = np.empty(output_shape(), np.float32)
    L = lib()
    desk = None
    dptr = None
    if save_deskewed:
        desk = np.empty(deskewed_shape(), np.float32)
        dptr = desk.ctypes.data
    rc = L.RL_interface(im, nx, ny, nz, out, dptr, float(background), bool(rescale), bool(save_deskewed),
                        int(n_iters), int(shift), int(napodize), int(nz_blend), float(pad_val),
                        bool(dup_reversed_z), bool(skewed_decon))
    _check(rc, "RL_interface")
    return (out, desk) if save_deskewed else out


def session_kscale():
    """OTF k-scales of the current session's iteration grid."""
    ks = np.zeros(3, np.float32)
    _check(lib().cudaDecon_get_session_kscale(ks), "cudaDecon_get_session_kscale")
    return tuple(ks)


def deskewed_shape():
    """(nz, ny, nx) of the deskewed raw stack of the current session."""
    L = lib()
    return (int(L.cudaDecon_get_session_dims(0)), int(L.get_output_ny()), int(L.cudaDecon_get_session_dims(1)))


def rl_cleanup():
    lib().RL_cleanup()


def deskew(im, angle, dz, dx, width=0, shift=0, pad_val=0.0, out_width=None):
    """Deskew_interface on a float (nz, ny, nx) stack.  The result's width is
    the one the library writes (cudaDecon_deskew_width: `width`, or with
    width == 0 the reference's rule from the session's output nz,
    src/RL-Biggs-Andrews.cpp:736-744); out_width, if given, must equal it."""
    im = _f32(im)
    nz, ny, nx = im.shape
    w = int(lib().cudaDecon_deskew_width(nx, float(dz), float(dx), float(angle), int(width)))
    if out_width is not None and int(out_width) != w:
        raise ValueError(f"deskew: out_width {out_width} differs from the width Deskew_interface writes ({w})")
    if w <= 0:
        raise ValueError("deskew: zero output width")
    out = np.empty((nz, ny, w), np.float32)
    _check(lib().Deskew_interface(im, nx, ny, nz, float(dz), float(dx), float(angle), out, int(width), int(shift),
                                  float(pad_val)), "Deskew_interface")
    return out


def affine(im, mat):
    im = _f32(im)
    nz, ny, nx = im.shape
    out = np.empty_like(im)
    _check(lib().Affine_interface(im, nx, ny, nz, out, _f32(np.asarray(mat).reshape(16))), "Affine_interface")
    return out


def affine_ra(im, mat, dx, dy, dz):
    im = _f32(im)
    nz, ny, nx = im.shape
    out = np.empty_like(im)
    _check(lib().Affine_interface_RA(im, nx, ny, nz, float(dx), float(dy), float(dz), out,
                                     _f32(np.asarray(mat).reshape(16))), "Affine_interface_RA")
    return out


# ---------------------------------------------------------------------------
# extensions
# ---------------------------------------------------------------------------
def fft_r2c(v):
    v = _f32(v)
    nz, ny, nx = v.shape
    out = np.empty((nz, ny, nx // 2 + 1), np.complex64)
    _check(lib().cudaDecon_fft_r2c(v, nx, ny, nz, out.view(np.float32).reshape(-1)), "cudaDecon_fft_r2c")
    return out


def fft_c2r(s, nx):
    s = np.ascontiguousarray(s, dtype=np.complex64)
    nz, ny, _ = s.shape
    out = np.empty((nz, ny, nx), np.float32)
    _check(lib().cudaDecon_fft_c2r(s.view(np.float32).reshape(-1), nx, ny, nz, out), "cudaDecon_fft_c2r")
    return out


def otf_image(otf_table: np.ndarray) -> np.ndarray:
    """OTF image as stored in its TIFF: an RA table (nr, nzotf) complex64 is
    width 2*nzotf x height nr; a 3-D table (nzotf, nyotf, nxotf) complex64 is
    width 2*nxotf x height nyotf x depth nzotf."""
    t = np.ascontiguousarray(otf_table, dtype=np.complex64)
    img = t.view(np.float32)
    if t.ndim == 2:
        return img.reshape(1, t.shape[0], 2 * t.shape[1])
    return img.reshape(t.shape[0], t.shape[1], 2 * t.shape[2])


def filter_volume(v, otf_table, kscale, conj=False):
    """One filterGPU call: C2R(H . R2C(v / N)) with H interpolated from otf_table."""
    v = _f32(v).copy()
    nz, ny, nx = v.shape
    img = otf_image(otf_table)
    d, h, w = img.shape
    _check(lib().cudaDecon_filter(v, nx, ny, nz, np.ascontiguousarray(img).reshape(-1), w, h, d,
                                  float(kscale[0]), float(kscale[1]), float(kscale[2]), int(conj)),
           "cudaDecon_filter")
    return v


def autotune(shapes_zyx, reps=5, use_cache=True):
    """Tune the FFT launch shapes for the grids (list of (nz, ny, nx)); the
    choice applies to every later call in this process.  Returns the report."""
    import json
    dims = np.ascontiguousarray(np.array(shapes_zyx, dtype=np.int32).reshape(-1, 3))
    buf = C.create_string_buffer(1 << 20)
    _check(lib().cudaDecon_autotune(dims.ctypes.data, dims.shape[0], int(reps), int(bool(use_cache)), buf, len(buf)),
           "cudaDecon_autotune")
    return json.loads(buf.value.decode())


def autotune_clear():
    _check(lib().cudaDecon_autotune_clear(), "cudaDecon_autotune_clear")


def fft_variant(n):
    """None for the built-in shape of length n, else its
    {cols, xthreads, col_nt_load, xcd_remap, col_tpt, zf_waves, zf_rows}."""
    v = np.zeros(7, np.int32)
    tuned = lib().cudaDecon_fft_variant(int(n), v.ctypes.data, 7)
    names = ("cols", "xthreads", "col_nt_load", "xcd_remap", "col_tpt", "zf_waves", "zf_rows")
    return dict(zip(names, map(int, v))) if tuned else None


def wiener_filter(v, otf_table, kscale, dkxy, rcutoff, wiener, background=0.0):
    """One-step Wiener filter of a (nz, ny, nx) volume (cudaDecon_wiener);
    otf_table is the rotationally averaged (nr, nzotf) complex table."""
    out = _f32(v).copy()
    nz, ny, nx = out.shape
    img = otf_image(otf_table)
    d, h, w = img.shape
    flat = np.ascontiguousarray(img).reshape(-1)
    _check(lib().cudaDecon_wiener(out, nx, ny, nz, flat, w, h, float(kscale[0]), float(kscale[1]), float(kscale[2]),
                                  float(dkxy[0]), float(dkxy[1]), float(rcutoff), float(wiener), float(background)),
           "cudaDecon_wiener")
    return out


def camcor(raw_u16, camparam):
    """Camera correction (camcor_interface_init + camcor_interface) of a
    (nz, ny, nx) uint16 stack; camparam (3, ny, nx) = (a, b, offset)."""
    raw = np.ascontiguousarray(raw_u16, dtype=np.uint16)
    cam = np.ascontiguousarray(camparam, dtype=np.float32)
    nz, ny, nx = raw.shape
    if cam.shape != (3, ny, nx):
        raise ValueError("camparam must be (3, ny, nx)")
    _check(lib().camcor_interface_init(nx, ny, nz, cam.ctypes.data), "camcor_interface_init")
    out = np.empty_like(raw)
    _check(lib().camcor_interface(raw.ctypes.data, nx, ny, nz, out.ctypes.data), "camcor_interface")
    return out


class DeviceBuffer:
    """Device memory for the entry points that take device pointers
    (cudaDecon_device_malloc / copy / free)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = lib().cudaDecon_device_malloc(self.nbytes)
        if not self.ptr:
            raise RuntimeError(f"cudaDecon_device_malloc failed: {last_error()}")

    def upload(self, a):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        _check(lib().cudaDecon_device_copy(self.ptr, a.ctypes.data, a.nbytes, 1), "cudaDecon_device_copy")

    def download(self, shape, dtype=np.float32):
        out = np.empty(shape, dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        _check(lib().cudaDecon_device_copy(out.ctypes.data, self.ptr, out.nbytes, 2), "cudaDecon_device_copy")
        return out

    def close(self):
        if self.ptr:
            lib().cudaDecon_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rl_run_noalloc(d_raw, d_out, shape, otf_table, kscale, n_iters, stream=None):
    """cudaDecon_rl_run_noalloc on device buffers (DeviceBuffer or raw
    pointers): n_iters RL-BA iterations of the stack in d_raw into d_out,
    with no device allocation."""
    nz, ny, nx = shape
    img = otf_image(otf_table)
    d, h, w = img.shape
    flat = np.ascontiguousarray(img).reshape(-1)
    src = d_raw.ptr if isinstance(d_raw, DeviceBuffer) else d_raw
    dst = d_out.ptr if isinstance(d_out, DeviceBuffer) else d_out
    _check(lib().cudaDecon_rl_run_noalloc(src, dst, nx, ny, nz, flat, w, h, d, float(kscale[0]), float(kscale[1]),
                                          float(kscale[2]), int(n_iters), stream), "cudaDecon_rl_run_noalloc")


class RLEngine:
    """RL-Biggs-Andrews on HBM-resident state (cudaDecon_rl_* handle).  With
    stacks > 1 the engine holds that many stacks of `shape` and iterates
    them together (cudaDecon_rl_create_stacks): set_raw / get then take and
    return (stacks, nz, ny, nx) arrays, iterate's factors are (n, stacks)."""

    def __init__(self, shape, otf_table, kscale, stacks=1):
        nz, ny, nx = shape
        self.shape = (nz, ny, nx)
        self.stacks = int(stacks)
        img = otf_image(otf_table)
        d, h, w = img.shape
        self._img = np.ascontiguousarray(img).reshape(-1)
        ks = (float(kscale[0]), float(kscale[1]), float(kscale[2]))
        if self.stacks == 1:
            self.h = lib().cudaDecon_rl_create(nx, ny, nz, self._img, w, h, d, *ks)
        else:
            self.h = lib().cudaDecon_rl_create_stacks(nx, ny, nz, self.stacks, self._img, w, h, d, *ks)
        if not self.h:
            raise RuntimeError(f"cudaDecon_rl_create failed: {last_error()}")

    def _all(self):
        return self.shape if self.stacks == 1 else (self.stacks,) + self.shape

    def set_raw(self, raw):
        raw = _f32(raw)
        if raw.size != int(np.prod(self._all())):
            raise ValueError(f"set_raw: expected {self._all()} voxels")
        _check(lib().cudaDecon_rl_set_raw(self.h, raw), "cudaDecon_rl_set_raw")

    def set_raw_stack(self, s, raw):
        raw = _f32(raw)
        if raw.shape != self.shape:
            raise ValueError(f"set_raw_stack: expected {self.shape}")
        _check(lib().cudaDecon_rl_set_raw_stack(self.h, int(s), raw), "cudaDecon_rl_set_raw_stack")

    def set_seed(self, seed):
        _check(lib().cudaDecon_rl_set_seed(self.h, _f32(seed)), "cudaDecon_rl_set_seed")

    def iterate(self, n, k0=0, seeded=False):
        """Run n iterations; returns (lambdas, device ms)."""
        lam = np.zeros(max(int(n), 1) * self.stacks, np.float64)
        ms = np.zeros(1, np.float64)
        _check(lib().cudaDecon_rl_iterate(self.h, int(n), int(k0), int(seeded), lam.ctypes.data, ms.ctypes.data),
               "cudaDecon_rl_iterate")
        lam = lam[:n * self.stacks]
        return (lam if self.stacks == 1 else lam.reshape(n, self.stacks)), float(ms[0])

    def get(self):
        out = np.empty(self._all(), np.float32)
        _check(lib().cudaDecon_rl_get(self.h, out), "cudaDecon_rl_get")
        return out

    def get_stack(self, s):
        out = np.empty(self.shape, np.float32)
        _check(lib().cudaDecon_rl_get_stack(self.h, int(s), out), "cudaDecon_rl_get_stack")
        return out

    def timing(self, enable=-1):
        ms = np.zeros(7, np.float64)
        calls = np.zeros(7, np.int64)
        _check(lib().cudaDecon_rl_timing(self.h, int(enable), ms.ctypes.data, calls.ctypes.data),
               "cudaDecon_rl_timing")
        return dict(zip(PASS_NAMES, ms)), dict(zip(PASS_NAMES, calls))

    def geometry(self):
        v = (C.c_long * 3)()
        _check(lib().cudaDecon_rl_geometry(self.h, C.byref(v, 0), C.byref(v, 8), C.byref(v, 16)),
               "cudaDecon_rl_geometry")
        return int(v[0]), int(v[1]), int(v[2])

    def set_raw_block(self, block, offset_zyx, background=0.0):
        """Stack = background everywhere, with `block` (bz, by, bx) at offset (oz, oy, ox)."""
        b = _f32(block)
        bz, by, bx = b.shape
        oz, oy, ox = offset_zyx
        _check(lib().cudaDecon_rl_set_raw_block(self.h, b, bx, by, bz, int(ox), int(oy), int(oz),
                                                float(background)), "cudaDecon_rl_set_raw_block")

    def get_block(self, shape_zyx, offset_zyx):
        bz, by, bx = shape_zyx
        oz, oy, ox = offset_zyx
        out = np.empty((bz, by, bx), np.float32)
        _check(lib().cudaDecon_rl_get_block(self.h, out, bx, by, bz, int(ox), int(oy), int(oz)),
               "cudaDecon_rl_get_block")
        return out

    def stats(self):
        """{'raw': (sum, min, max), 'x': (sum, min, max)} computed on the device."""
        o = np.zeros(6, np.float64)
        _check(lib().cudaDecon_rl_stats(self.h, o), "cudaDecon_rl_stats")
        return {"raw": tuple(o[:3]), "x": tuple(o[3:])}

    def copy_probe(self, n=10):
        """Average ms of one plain streaming copy of the stack (read + write)."""
        ms = np.zeros(1, np.float64)
        _check(lib().cudaDecon_rl_copy_probe(self.h, int(n), ms.ctypes.data), "cudaDecon_rl_copy_probe")
        return float(ms[0])

    def set_raw_u16(self, raw_u16, offset):
        """Stage a uint16 stack as max(v - offset, 0) (camera offset subtraction on the device)."""
        r = np.ascontiguousarray(raw_u16, dtype=np.uint16)
        _check(lib().cudaDecon_rl_set_raw_u16(self.h, r, float(offset)), "cudaDecon_rl_set_raw_u16")

    def deskew(self, dz, dr, angle, width=0, shift=0, pad_val=0.0):
        """Deskew the current X on the device; returns the output width."""
        nx = C.c_int(0)
        _check(lib().cudaDecon_rl_deskew(self.h, float(dz), float(dr), float(angle), int(width), int(shift),
                                         float(pad_val), C.byref(nx)), "cudaDecon_rl_deskew")
        return nx.value

    def affine(self, mat):
        """Affine resampling of the last deskew result (else of X) on the device."""
        m = _f32(np.asarray(mat, np.float32).reshape(-1))
        if m.size != 16:
            raise ValueError("affine matrix must have 16 entries (row-major 4x4)")
        _check(lib().cudaDecon_rl_affine(self.h, m), "cudaDecon_rl_affine")

    def post_shape(self):
        v = (C.c_int * 3)()
        _check(lib().cudaDecon_rl_post_dims(self.h, C.byref(v, 0), C.byref(v, 4), C.byref(v, 8)),
               "cudaDecon_rl_post_dims")
        return int(v[2]), int(v[1]), int(v[0])

    def post_get(self):
        out = np.empty(self.post_shape(), np.float32)
        _check(lib().cudaDecon_rl_post_get(self.h, out), "cudaDecon_rl_post_get")
        return out

    def post_get_block(self, buf, shape_zyx, offset_zyx):
        """Block of the deskew (buf 0) or affine (buf 1) result."""
        bz, by, bx = shape_zyx
        oz, oy, ox = offset_zyx
        out = np.empty((bz, by, bx), np.float32)
        _check(lib().cudaDecon_rl_post_get_block(self.h, int(buf), out, bx, by, bz, int(ox), int(oy), int(oz)),
               "cudaDecon_rl_post_get_block")
        return out

    def buffer_stats(self, which):
        """(sum, min, max) of an engine buffer: 0 staged stack, 1 X, 2 Y (seed), 3 G."""
        o = np.zeros(3, np.float64)
        _check(lib().cudaDecon_rl_buffer_stats(self.h, int(which), o), "cudaDecon_rl_buffer_stats")
        return tuple(o)

    def post_stats(self):
        o = np.zeros(3, np.float64)
        _check(lib().cudaDecon_rl_post_stats(self.h, o), "cudaDecon_rl_post_stats")
        return tuple(o)

    def snapshot(self):
        """Keep a device copy of the current X as the reference of distance()."""
        _check(lib().cudaDecon_rl_snapshot(self.h), "cudaDecon_rl_snapshot")

    def distance(self, seed=False):
        """||X - snapshot|| / ||snapshot|| (or of the Y_0 seed) computed on the device."""
        r = np.zeros(1, np.float64)
        _check(lib().cudaDecon_rl_distance(self.h, int(bool(seed)), r.ctypes.data), "cudaDecon_rl_distance")
        return float(r[0])

    def warm_start(self, net):
        """Seed Y_0 with the network's output on the staged stack; returns the
        device ms of the network.  Follow with iterate(n, seeded=True)."""
        ms = np.zeros(1, np.float64)
        _check(lib().cudaDecon_rl_warm_start(self.h, net.h, ms.ctypes.data), "cudaDecon_rl_warm_start")
        return float(ms[0])

    def close(self):
        if self.h:
            lib().cudaDecon_rl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many independent stacks of uneven shape (cudaDecon_batch_* handle).

    shapes: list of (nz, ny, nx); kscales: one (kx, ky, kz) per stack;
    otf_table: one OTF table for every stack, or a list of tables with
    table_of giving each stack's (channels)."""

    MODES = {"sequential": 0, "streams": 1, "graph": 2, "grouped": 3, "grouped_graph": 4}

    def __init__(self, shapes, otf_table, kscales, nstreams=8, table_of=None):
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        dims = np.ascontiguousarray(np.array(self.shapes, np.int32).reshape(-1))
        ks = _f32(np.array(kscales, np.float32).reshape(-1))
        if isinstance(otf_table, (list, tuple)):
            imgs = [otf_image(t) for t in otf_table]
            self._img = np.ascontiguousarray(np.concatenate([i.reshape(-1) for i in imgs]))
            whd = np.ascontiguousarray(np.array([(i.shape[2], i.shape[1], i.shape[0]) for i in imgs], np.int32))
            tof = np.ascontiguousarray(np.array(table_of if table_of is not None else [0] * len(self.shapes),
                                                np.int32))
            if tof.size != len(self.shapes):
                raise ValueError("table_of needs one table index per stack")
            self.h = lib().cudaDecon_batch_create_tables(len(self.shapes), dims.ctypes.data, len(imgs), self._img,
                                                         whd.ctypes.data, tof.ctypes.data, ks, int(nstreams))
        else:
            img = otf_image(otf_table)
            d, h, w = img.shape
            self._img = np.ascontiguousarray(img).reshape(-1)
            self.h = lib().cudaDecon_batch_create(len(self.shapes), dims.ctypes.data, self._img, w, h, d, ks,
                                                  int(nstreams))
        if not self.h:
            raise RuntimeError(f"cudaDecon_batch_create failed: {last_error()}")

    def lambdas(self, i, iters):
        """Acceleration factors of stack i in the last run (0 for k < 2)."""
        out = np.zeros(int(iters), np.float64)
        _check(lib().cudaDecon_batch_lambdas(self.h, int(i), int(iters), out), "cudaDecon_batch_lambdas")
        return out

    def set_grid_cap(self, max_workgroups):
        """Grouped modes: persistent column-pass grids of at most this many workgroups (0: one per tile)."""
        _check(lib().cudaDecon_batch_set_grid_cap(self.h, int(max_workgroups)), "cudaDecon_batch_set_grid_cap")

    def set_partitions(self, n):
        """Grouped modes: run the stacks as n ranges of about equal voxel count, each grouped on its own stream
        (at most nstreams)."""
        _check(lib().cudaDecon_batch_set_partitions(self.h, int(n)), "cudaDecon_batch_set_partitions")

    def launches(self):
        """Kernel launches per iteration of the grouped modes."""
        return int(lib().cudaDecon_batch_launches(self.h))

    def set_raw(self, i, raw):
        _check(lib().cudaDecon_batch_set_raw(self.h, int(i), _f32(raw)), "cudaDecon_batch_set_raw")

    def set_raw_block(self, i, block, offset_zyx, background=0.0):
        b = _f32(block)
        bz, by, bx = b.shape
        oz, oy, ox = offset_zyx
        _check(lib().cudaDecon_batch_set_raw_block(self.h, int(i), b, bx, by, bz, int(ox), int(oy), int(oz),
                                                   float(background)), "cudaDecon_batch_set_raw_block")

    def run(self, iters, mode="graph"):
        """Run `iters` iterations of every stack; returns device ms."""
        ms = np.zeros(1, np.float64)
        _check(lib().cudaDecon_batch_run(self.h, int(iters), self.MODES[mode], ms.ctypes.data),
               "cudaDecon_batch_run")
        return float(ms[0])

    def get(self, i):
        out = np.empty(self.shapes[i], np.float32)
        _check(lib().cudaDecon_batch_get(self.h, int(i), out), "cudaDecon_batch_get")
        return out

    def close(self):
        if self.h:
            lib().cudaDecon_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WarmNet:
    """Warm-start network (cudaDecon_net_* handle): BF16 3-D U-Net on the
    matrix cores whose output seeds Y_0 of RL-Biggs-Andrews."""

    FETCH = ("e1", "p1", "e2", "p2", "e3", "d3", "a2", "d2", "dy", "dPre5", "dA2", "dPre4", "dE3", "dPre3", "dP2",
             "dPre2", "dP1", "dPre1")

    def __init__(self, seed=0, params=None):
        self.h = lib().cudaDecon_net_create(int(seed))
        if not self.h:
            raise RuntimeError(f"cudaDecon_net_create failed: {last_error()}")
        if params is not None:
            self.set_params(params)

    @staticmethod
    def num_params():
        return int(lib().cudaDecon_net_num_params())

    def get_params(self):
        p = np.empty(self.num_params(), np.float32)
        _check(lib().cudaDecon_net_get_params(self.h, p), "cudaDecon_net_get_params")
        return p

    def set_params(self, p):
        p = _f32(p).reshape(-1)
        if p.size != self.num_params():
            raise ValueError(f"expected {self.num_params()} parameters, got {p.size}")
        _check(lib().cudaDecon_net_set_params(self.h, p), "cudaDecon_net_set_params")

    @staticmethod
    def _batch(raw):
        r = _f32(raw)
        if r.ndim == 3:
            r = r[None]
        return np.ascontiguousarray(r)

    def forward(self, raw, seed=False, inference=False):
        """y (normalised output), or the RL seed when seed=True; raw (nb,nz,ny,nx)
        or (nz,ny,nx).  inference=True runs the path of RLEngine.warm_start
        (full-size e1 recomputed in the head, not stored)."""
        r = self._batch(raw)
        out = np.empty_like(r)
        nb, nz, ny, nx = r.shape
        mode = int(bool(seed)) + (2 if inference else 0)
        _check(lib().cudaDecon_net_forward(self.h, r, nb, nz, ny, nx, mode, out), "cudaDecon_net_forward")
        return out.reshape(np.shape(raw))

    def fetch(self, name):
        which = self.FETCH.index(name)
        n = lib().cudaDecon_net_fetch(self.h, which, None)
        if n < 0:
            raise RuntimeError(last_error())
        out = np.empty(n, np.float32)
        if lib().cudaDecon_net_fetch(self.h, which, out.ctypes.data) != n:
            raise RuntimeError(last_error())
        return out

    def loss_grad(self, raw, target):
        r, t = self._batch(raw), self._batch(target)
        nb, nz, ny, nx = r.shape
        loss = np.zeros(1, np.float64)
        g = np.empty(self.num_params(), np.float32)
        _check(lib().cudaDecon_net_loss_grad(self.h, r, t, nb, nz, ny, nx, loss.ctypes.data, g.ctypes.data),
               "cudaDecon_net_loss_grad")
        return float(loss[0]), g

    def step(self, raw, target, lr):
        r, t = self._batch(raw), self._batch(target)
        nb, nz, ny, nx = r.shape
        loss = np.zeros(1, np.float64)
        _check(lib().cudaDecon_net_step(self.h, r, t, nb, nz, ny, nx, float(lr), loss.ctypes.data),
               "cudaDecon_net_step")
        return float(loss[0])

    def add_samples(self, raw, target):
        r, t = self._batch(raw), self._batch(target)
        nb, nz, ny, nx = r.shape
        _check(lib().cudaDecon_net_add_samples(self.h, r, t, nb, nz, ny, nx), "cudaDecon_net_add_samples")

    def train(self, steps, batch, lr, seed=0):
        """(losses per step, device ms)."""
        losses = np.zeros(max(int(steps), 1), np.float32)
        ms = np.zeros(1, np.float64)
        _check(lib().cudaDecon_net_train(self.h, int(steps), int(batch), float(lr), int(seed), losses.ctypes.data,
                                         ms.ctypes.data), "cudaDecon_net_train")
        return losses[:steps], float(ms[0])

    def close(self):
        if self.h:
            lib().cudaDecon_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

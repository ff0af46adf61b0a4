"""ctypes binding of libradialft.so, the C entry point makeOTF
(include/radialft.h; src/radialft_interface.cpp:52-196 in the reference),
the way pycudadecon binds the reference's libradialft."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

SIGNATURES = {
    "makeOTF": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_float, C.c_int, C.c_bool, C.c_float, C.c_float,
                          C.c_float, C.c_float, C.c_int, C.c_bool, C.c_bool]),
    "radialft_last_error": (C.c_char_p, []),
}


def lib_path() -> str:
    return os.path.join(_HERE, "libradialft.so")


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(lib_path()):
            raise RuntimeError(f"{lib_path()} is not built (make -C cudadecon_amd/csrc)")
        L = C.CDLL(lib_path())
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def make_otf_file(psf_path, otf_path, lambdanm=520, dz=0.102, interpkr=10, background=None, NA=1.25, NIMM=1.3,
                  dr=0.102, krmax=0, cleanup=False, three_d=False):
    """makeOTF(psf_path -> otf_path); background None estimates it from the
    PSF as the reference does.  Raises RuntimeError when makeOTF returns 1."""
    rc = lib().makeOTF(os.fsencode(psf_path), os.fsencode(otf_path), int(lambdanm), float(dz), int(interpkr),
                       background is not None, float(background or 0.0), float(NA), float(NIMM), float(dr),
                       int(krmax), bool(cleanup), bool(three_d))
    if rc != 0:
        raise RuntimeError("makeOTF failed: " + lib().radialft_last_error().decode())

"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/cudaDecon.h declares, and its host-only entry points behave (no GPU
call is made here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# exported symbols of the reference library (src/linearDecon.h:243-317)
REFERENCE_SYMBOLS = [
    "RL_interface_init", "RL_interface", "Deskew_interface", "Affine_interface", "Affine_interface_RA",
    "camcor_interface_init", "camcor_interface", "RL_cleanup", "cuda_reset", "get_output_nx", "get_output_ny",
    "get_output_nz", "get_version",
]


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "cudaDecon.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"CDN_DEFAULT", "if", "defined"}
    return sorted({n for n in names if n not in skip and not n.isupper()})


def test_header_declares_reference_api():
    decl = declared_symbols()
    for s in REFERENCE_SYMBOLS:
        assert s in decl, s


def test_library_exports_every_declared_symbol(cdn):
    L = ctypes.CDLL(cdn.lib_path())
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_python_bindings_cover_header(cdn):
    assert set(declared_symbols()) == set(cdn.SIGNATURES)


def test_version(cdn):
    # src/CMakeLists.txt:17 (project version 0.7.0)
    assert cdn.lib().get_version() == b"0.7.0"


def test_no_session_outputs_zero(cdn):
    L = cdn.lib()
    L.RL_cleanup()
    assert L.get_output_nx() == 0 and L.get_output_ny() == 0 and L.get_output_nz() == 0


def test_rl_interface_before_init_is_an_error(cdn):
    import numpy as np
    L = cdn.lib()
    L.RL_cleanup()
    raw = np.zeros((4, 4, 4), np.uint16)
    out = np.zeros((4, 4, 4), np.float32)
    rc = L.RL_interface(raw, 4, 4, 4, out, None, 100.0, False, False, 1, 0, 0, 0, 0.0, False, False)
    assert rc == 0
    assert "before RL_interface_init" in cdn.last_error()


def test_init_with_missing_otf_returns_zero(cdn, tmp_path):
    L = cdn.lib()
    rc = L.RL_interface_init(128, 128, 64, 0.104, 0.2, 0.104, 0.2, 0.0, 0.0, 0, False,
                             str(tmp_path / "missing.tif").encode())
    assert rc == 0
    assert "cannot open" in cdn.last_error()


def test_camcor_argument_errors(cdn):
    L = cdn.lib()
    assert L.camcor_interface_init(1, 1, 1, None) == 0
    assert "bad argument" in cdn.last_error()
    assert L.camcor_interface(None, 1, 1, 1, None) == 0


def test_find_optimal_dimension_matches_oracle(cdn, orc):
    L = cdn.lib()
    for n in list(range(1, 3000)) + [3727, 4096, 8191, 8192]:
        assert L.cudaDecon_find_optimal_dimension(n, -1) == orc.find_optimal_dimension(n, -1), n
        assert L.cudaDecon_find_optimal_dimension(n, 1) == orc.find_optimal_dimension(n, 1), n


@pytest.mark.parametrize("n", [64, 96, 128, 192, 256, 384, 512, 768, 810, 1024, 2048])
def test_supported_fft_lengths(cdn, n):
    assert cdn.lib().cudaDecon_fft_supported(n) == 1


def test_unsupported_fft_length_is_reported(cdn):
    L = cdn.lib()
    assert L.cudaDecon_fft_supported(11) == 0
    assert L.cudaDecon_fft_supported(20000) == 0
    # any other 2,3,5,7-smooth length is compiled at run time
    assert L.cudaDecon_fft_supported(1000) == 1 and L.cudaDecon_fft_supported(150) == 1


def test_init_rejects_too_long_fft(cdn, tmp_path):
    import numpy as np
    from cudadecon_amd.otf import write_otf_tiff
    p = tmp_path / "otf.tif"
    write_otf_tiff(str(p), np.ones((17, 33), np.complex64))
    L = cdn.lib()
    # 20000 = 2^5 5^4 is smooth but beyond the 16384 kernel limit
    rc = L.RL_interface_init(20000, 128, 64, 0.104, 0.2, 0.104, 0.2, 0.0, 0.0, 0, False, str(p).encode())
    assert rc == 0
    assert "20000" in cdn.last_error()


def test_init_compiles_other_smooth_lengths_at_run_time(cdn, tmp_path):
    """150 x 72 x 70 (2*3*5^2, 2^3*3^2, 2*5*7) is not precompiled: init
    compiles its kernels with hipRTC (no GPU needed) and succeeds."""
    import numpy as np
    from cudadecon_amd.otf import write_otf_tiff
    p = tmp_path / "otf.tif"
    write_otf_tiff(str(p), np.ones((17, 33), np.complex64))
    L = cdn.lib()
    rc = L.RL_interface_init(150, 72, 70, 0.104, 0.2, 0.104, 0.2, 0.0, 0.0, 0, False, str(p).encode())
    assert rc == 1, cdn.last_error()
    assert (L.get_output_nx(), L.get_output_ny(), L.get_output_nz()) == (150, 72, 70)
    L.RL_cleanup()


# RL_interface_init scalars: library against the oracle's restatement of
# src/RL-Biggs-Andrews.cpp:531-655 (no GPU work happens before RL_interface)
INIT_CASES = [
    # nx, ny, nz, dr, dz, dr_psf, dz_psf, angle, rot, width, skewed
    (128, 128, 64, 0.104, 0.2, 0.104, 0.2, 0.0, 0.0, 0, False),
    (512, 512, 128, 0.104, 0.3, 0.104, 0.15675, 31.5, 31.5, 0, False),
    (100, 128, 64, 0.104, 0.3, 0.104, 0.3, -32.8, -32.8, 0, False),  # deskewed 255 -> 252 (run-time kernels)
    (100, 128, 64, 0.104, 0.3, 0.104, 0.3, 31.5, 0.0, 256, False),
    (130, 130, 66, 0.1, 0.25, 0.1, 0.25, 0.0, 15.0, 0, False),  # crops to 128 x 128 x 64
    (1030, 520, 260, 0.104, 0.2, 0.104, 0.2, 0.0, 0.0, 0, False),  # crops to 1029 x 512 x 256
    (512, 512, 128, 0.104, 0.3, 0.104, 0.3, 31.5, 0.0, 0, True),
]


@pytest.mark.parametrize("case", INIT_CASES, ids=[str(i) for i in range(len(INIT_CASES))])
def test_init_geometry_matches_oracle(cdn, orc, tmp_path, case):
    import numpy as np
    from cudadecon_amd.otf import write_otf_tiff
    nx, ny, nz, dr, dz, drp, dzp, ang, rot, width, skewed = case
    tab = np.ones((17, 33), np.complex64)
    p = tmp_path / "otf.tif"
    write_otf_tiff(str(p), tab)
    L = cdn.lib()
    L.RL_cleanup()
    rc = L.RL_interface_init(nx, ny, nz, dr, dz, drp, dzp, ang, rot, width, skewed, str(p).encode())
    assert rc == 1, cdn.last_error()
    g = orc.init_geometry(nx, ny, nz, dr, dz, drp, dzp, ang, rot, width, skewed, (66, 17, 1))
    assert (L.get_output_nx(), L.get_output_ny(), L.get_output_nz()) == g["out_dims"]
    assert L.cudaDecon_get_session_dims(1) == g["deskewed_nx"]
    assert L.cudaDecon_get_session_dims(2) == g["output_nx"]
    assert L.cudaDecon_get_session_dims(3) == g["fft_dims"][0]
    ks = cdn.session_kscale()
    assert all(np.float32(a) == np.float32(b) for a, b in zip(ks, g["kscale"])), (ks, g["kscale"])
    L.RL_cleanup()


def test_negative_rotation_sizes_use_abs(cdn, orc, tmp_path):
    """get_output_nx of the reference omits fabs(p[1]) (src/RL-Biggs-Andrews.cpp:479)
    while the result is sized with it (:411); this library uses fabs in both."""
    import numpy as np
    from cudadecon_amd.otf import write_otf_tiff
    p = tmp_path / "otf.tif"
    write_otf_tiff(str(p), np.ones((17, 33), np.complex64))
    L = cdn.lib()
    assert L.RL_interface_init(100, 128, 64, 0.104, 0.3, 0.104, 0.3, -31.5, -31.5, 0, False, str(p).encode()) == 1
    g = orc.init_geometry(100, 128, 64, 0.104, 0.3, 0.104, 0.3, -31.5, -31.5, 0, False, (66, 17, 1))
    assert g["rot"][1] < 0
    assert L.get_output_nx() == g["out_dims"][0] > 0
    L.RL_cleanup()


def test_init_rejects_nan_otf(cdn, tmp_path):
    """One NaN in the OTF would make every filtered voxel NaN: init refuses
    the file and says why (no GPU work happens before RL_interface)."""
    import numpy as np
    from cudadecon_amd.otf import write_otf_tiff
    tab = np.ones((17, 33), np.complex64)
    tab[3, 4] = np.complex64(complex(np.nan, 0.0))
    p = tmp_path / "otf_nan.tif"
    write_otf_tiff(str(p), tab)
    L = cdn.lib()
    L.RL_cleanup()
    assert L.RL_interface_init(128, 128, 64, 0.104, 0.2, 0.104, 0.2, 0.0, 0.0, 0, False, str(p).encode()) == 0
    assert "NaN or Inf" in cdn.last_error()
    assert L.get_output_nx() == 0  # no session was set up


def test_deskew_width_rule(cdn, tmp_path):
    """cudaDecon_deskew_width gives the width Deskew_interface writes
    (src/RL-Biggs-Andrews.cpp:736-744): the caller's width, nx at a zero
    angle, else nx + floor(output_nz dz |cos| / dr) / 4 from the session."""
    import math

    import numpy as np
    from cudadecon_amd.otf import write_otf_tiff
    L = cdn.lib()
    assert L.cudaDecon_deskew_width(512, 0.3, 0.104, 31.5, 700) == 700
    assert L.cudaDecon_deskew_width(512, 0.3, 0.104, 0.0, 0) == 512
    assert L.cudaDecon_deskew_width(512, float("nan"), 0.104, 31.5, 0) == 0  # no valid width
    p = tmp_path / "otf.tif"
    write_otf_tiff(str(p), np.ones((17, 33), np.complex64))
    assert L.RL_interface_init(512, 512, 128, 0.104, 0.3, 0.104, 0.3, 31.5, 0.0, 0, False, str(p).encode()) == 1
    L.RL_cleanup()  # output_nz survives cleanup, as in the reference
    want = int(512 + math.floor(np.float32(128 * np.float32(0.3)) * abs(math.cos(math.radians(31.5)))
                                / np.float32(0.104)) / 4.0)
    got = L.cudaDecon_deskew_width(512, 0.3, 0.104, 31.5, 0)
    assert abs(got - want) <= 1 and got > 512, (got, want)
    with pytest.raises(ValueError):
        cdn.deskew(np.zeros((4, 4, 512), np.float32), 31.5, 0.3, 0.104, out_width=got + 1)

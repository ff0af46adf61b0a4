"""Host code that reads user files, built with AddressSanitizer and
UndefinedBehaviorSanitizer on the CPU (tests/native/host_inputs_check.cpp):
the TIFF reader, the OTF loader, the session geometry of RL_interface_init,
the Deskew_interface width rule and makeOTF.  Malformed files (truncated,
zero-count, oversized-count, cyclic IFD chains, strips outside the file,
image data larger than the file, NaN OTFs) must be rejected with a message
and no sanitizer report; well-formed ones must read back exactly."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "cudadecon_amd", "csrc")


def _ifd(entries, next_off):
    """Little-endian IFD bytes: entries are (tag, type, count, value)."""
    b = struct.pack("<H", len(entries))
    for tag, typ, cnt, val in entries:
        if typ == 3 and cnt == 1:
            b += struct.pack("<HHIHH", tag, typ, cnt, val, 0)
        else:
            b += struct.pack("<HHII", tag, typ, cnt, val)
    return b + struct.pack("<I", next_off)


def _page_entries(w, h, bps, data_off, nbytes, fmt=1, comp=1):
    return [(256, 4, 1, w), (257, 4, 1, h), (258, 3, 1, bps), (259, 3, 1, comp), (273, 4, 1, data_off),
            (277, 3, 1, 1), (279, 4, 1, nbytes), (339, 3, 1, fmt)]


def _single_page(w, h, bps, pixels: bytes, fmt=1, comp=1, extra=None):
    data_off = 8
    ifd_off = 8 + len(pixels)
    ents = _page_entries(w, h, bps, data_off, len(pixels), fmt, comp)
    if extra:
        ents = sorted(ents + extra)
    return b"II" + struct.pack("<HI", 42, ifd_off) + pixels + _ifd(ents, 0)


def malformed_files(d):
    files = {}
    px = np.arange(12, dtype=np.uint16).tobytes()
    good = _single_page(4, 3, 16, px)
    files["good_u16.tif"] = good
    otf = np.linspace(-1, 1, 2 * 6 * 5, dtype=np.float32)  # RA table nr 5 x nzotf 6
    files["good_otf.tif"] = _single_page(12, 5, 32, otf.tobytes(), fmt=3)
    otf_nan = otf.copy()
    otf_nan[7] = np.nan
    files["nan_otf.tif"] = _single_page(12, 5, 32, otf_nan.tobytes(), fmt=3)
    files["tiny_otf.tif"] = _single_page(1, 5, 32, otf[:5].tobytes(), fmt=3)
    files["short.tif"] = b"II*\x00"
    files["bad_bom.tif"] = b"XX" + good[2:]
    files["bigtiff.tif"] = b"II" + struct.pack("<H", 43) + good[4:]
    files["ifd_past_end.tif"] = b"II" + struct.pack("<HI", 42, 1 << 30) + px
    files["truncated_ifd.tif"] = good[:-20]
    # zero-count width, and a strip-offset array claiming 2^32-1 entries
    b = bytearray(good)
    ifd_off = 8 + len(px)
    struct.pack_into("<I", b, ifd_off + 2 + 4, 0)
    files["zero_count.tif"] = bytes(b)
    b = bytearray(good)
    struct.pack_into("<I", b, ifd_off + 2 + 12 * 4 + 4, 0xFFFFFFFF)
    files["huge_count.tif"] = bytes(b)
    # IFD chain that points back at itself, and a two-page loop
    b = bytearray(good)
    struct.pack_into("<I", b, len(b) - 4, ifd_off)
    files["cyclic_self.tif"] = bytes(b)
    ents = _page_entries(4, 3, 16, 8, len(px))
    i1 = 8 + len(px)
    i2 = i1 + len(_ifd(ents, 0))
    files["cyclic_pair.tif"] = (b"II" + struct.pack("<HI", 42, i1) + px + _ifd(ents, i2) + _ifd(ents, i1))
    # strip outside the file, and a page far larger than the file
    files["strip_past_end.tif"] = _single_page(4, 3, 16, px)[:8] + px + _ifd(
        _page_entries(4, 3, 16, 1 << 20, len(px)), 0)
    files["huge_page.tif"] = b"II" + struct.pack("<HI", 42, 8 + len(px)) + px + _ifd(
        _page_entries(65535, 65535, 32, 8, len(px)), 0)
    # 300 IFDs that all point at the same strip: more image than file
    n = 300
    big = np.ones(64 * 64, np.uint16).tobytes()
    ents = _page_entries(64, 64, 16, 8, len(big))
    size = len(_ifd(ents, 0))
    body = b""
    for k in range(n):
        off = 8 + len(big) + k * size
        body += _ifd(ents, off + size if k + 1 < n else 0)
    files["shared_strips.tif"] = b"II" + struct.pack("<HI", 42, 8 + len(big)) + big + body
    files["bps12.tif"] = _single_page(4, 3, 12, px)
    files["compressed.tif"] = _single_page(4, 3, 16, px, comp=5)
    files["float16.tif"] = _single_page(4, 3, 16, px, fmt=3)
    files["empty_page.tif"] = _single_page(0, 3, 16, px)
    for name, data in files.items():
        with open(os.path.join(d, name), "wb") as f:
            f.write(data)
    return sorted(files)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_host_inputs_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "host_inputs_check")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
                           "-fsanitize=address,undefined,float-cast-overflow", "-fno-sanitize-recover=all",
                           "-I", CSRC, os.path.join(ROOT, "tests", "native", "host_inputs_check.cpp"),
                           os.path.join(CSRC, "tiff.cpp"), os.path.join(CSRC, "geometry.cpp"),
                           os.path.join(CSRC, "radialft.cpp"), "-o", exe])
    names = malformed_files(str(tmp_path))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=66",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=67")
    r = subprocess.run([exe, str(tmp_path / "out_otf.tif")] + [str(tmp_path / n) for n in names],
                       capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout
    assert r.returncode == 0, out + r.stderr
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    lines = dict(line.split(": ", 1) for line in out.splitlines() if ": read" in line or ": rejected" in line)
    assert lines["good_u16.tif"].startswith("read 4x3x1 sum 66"), lines["good_u16.tif"]
    assert lines["good_otf.tif"].startswith("read 12x5x1"), lines["good_otf.tif"]
    assert "good_otf.tif: otf 5x1x6" in out
    assert "nan_otf.tif: otf rejected: OTF table holds a NaN" in out
    assert "tiny_otf.tif: otf rejected" in out
    expect = {
        "short.tif": "file too short", "bad_bom.tif": "bad byte-order mark", "bigtiff.tif": "not a classic TIFF",
        "ifd_past_end.tif": "truncated", "truncated_ifd.tif": "truncated", "zero_count.tif": "empty field",
        "huge_count.tif": "truncated", "cyclic_self.tif": "loops", "cyclic_pair.tif": "loops",
        "strip_past_end.tif": "truncated", "huge_page.tif": "larger than the file",
        "shared_strips.tif": "larger than the file", "bps12.tif": "bits per sample",
        "compressed.tif": "compressed", "float16.tif": "sample format", "empty_page.tif": "empty page",
    }
    for name, why in expect.items():
        assert lines[name].startswith("rejected:") and why in lines[name], (name, lines[name])
        assert f"{name}: makeOTF 1" in out, name
    assert "geometry:" in out

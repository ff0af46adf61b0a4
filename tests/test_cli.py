"""The cudaDecon command-line caller (cudadecon_amd/cudaDecon, built from
cudadecon_amd/csrc/cli.cpp; reference src/linearDecon.cpp:170-1431,
src/boostfs.cpp).

CPU: options, file discovery (pattern, --skip, --no_overwrite), the
">>>file_finished" marker, and errors, with no GPU work (-i 0 runs nothing).
GPU: a folder of synthetic uint16 TIFF stacks (one of another size) through
the CLI; every output is bitwise what RL_interface gives when called
directly on the same stack (flat start, bleach correction across files,
deskew with the deskewed raw stack, uint16 output, MIPs, a final crop, Y
tiles stitched as the reference does)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cudadecon_amd", "cudaDecon")


def _cli(args, check=True):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    r = subprocess.run([EXE] + [str(a) for a in args], capture_output=True, text=True, timeout=600)
    if check and r.returncode != 0:
        raise AssertionError(f"cudaDecon {args} failed ({r.returncode}):\n{r.stdout}\n{r.stderr}")
    return r


def _write_stack(path, a):
    """uint16 (nz, ny, nx) stack as a TIFF, one page per plane."""
    import struct
    a = np.ascontiguousarray(a, dtype=np.uint16)
    nz, ny, nx = a.shape
    with open(path, "wb") as f:
        f.write(b"II" + struct.pack("<HI", 42, 8))
        pos = 8
        for z in range(nz):
            n_ent = 8
            data_off = pos + 2 + 12 * n_ent + 4
            nxt = data_off + a[z].nbytes if z + 1 < nz else 0
            ents = [(256, 4, nx), (257, 4, ny), (258, 3, 16), (259, 3, 1), (262, 3, 1), (273, 4, data_off),
                    (277, 3, 1), (279, 4, a[z].nbytes)]
            f.write(struct.pack("<H", n_ent))
            for tag, typ, val in ents:
                if typ == 3:
                    f.write(struct.pack("<HHIHH", tag, typ, 1, val, 0))
                else:
                    f.write(struct.pack("<HHII", tag, typ, 1, val))
            f.write(struct.pack("<I", nxt))
            f.write(a[z].tobytes())
            pos = data_off + a[z].nbytes


def _read(path):
    """Pages of an uncompressed little-endian float32 or uint16 TIFF as a
    float32 (pages, height, width) array; other tag types are skipped."""
    import struct
    buf = open(path, "rb").read()
    assert buf[:4] == b"II*\x00", path
    pages, off = [], struct.unpack("<I", buf[4:8])[0]
    while off:
        n = struct.unpack("<H", buf[off:off + 2])[0]
        tags = {}
        for i in range(n):
            e = off + 2 + 12 * i
            tag, typ, cnt = struct.unpack("<HHI", buf[e:e + 8])
            if typ in (3, 4) and cnt == 1:
                tags[tag] = struct.unpack("<H" if typ == 3 else "<I", buf[e + 8:e + (10 if typ == 3 else 12)])[0]
        w, h, bits, fmt = tags[256], tags[257], tags[258], tags.get(339, 1)
        dt = {(32, 3): "<f4", (16, 1): "<u2"}[(bits, fmt)]
        pages.append(np.frombuffer(buf, dtype=dt, count=w * h, offset=tags[273]).reshape(h, w))
        off = struct.unpack("<I", buf[off + 2 + 12 * n:off + 6 + 12 * n])[0]
    return np.stack(pages).astype(np.float32)


def _folder(tmp_path, shapes, seed0=70):
    from cudadecon_amd.synth import make_stack
    d = tmp_path / "data"
    d.mkdir(parents=True)
    stacks = []
    for i, s in enumerate(shapes):
        raw, _, _ = make_stack("beads_64x128x128", shape=s, seed=seed0 + i)
        _write_stack(str(d / f"cell1_ch0_stack{i:04d}.tif"), raw)
        stacks.append(raw)
    (d / "notes.txt").write_text("not an image")
    return d, stacks


def _otf(tmp_path):
    from cudadecon_amd.otf import make_otf, write_otf_tiff
    from cudadecon_amd.synth import gaussian_psf
    tab = make_otf(gaussian_psf(1.5, 3.0))
    p = str(tmp_path / "otf.tif")
    write_otf_tiff(p, tab)
    return p


def test_cli_help_version_and_bad_options():
    r = _cli(["--version"])
    assert r.stdout.strip() == "0.7.0"
    assert "usage: cudaDecon" in _cli(["--help"]).stdout
    r = _cli(["--nonsense", "a", "b", "c"], check=False)
    assert r.returncode == 1 and "unknown option" in r.stderr
    r = _cli(["-i", "x", "a", "b", "c"], check=False)
    assert r.returncode == 1 and "needs an integer" in r.stderr
    r = _cli(["--LSC", "f.tif", "a", "b", "c"], check=False)
    assert r.returncode == 1 and "not supported" in r.stderr
    r = _cli(["only_one"], check=False)
    assert r.returncode == 1 and "expected input-dir" in r.stderr


def test_cli_file_discovery_skip_and_no_overwrite(tmp_path):
    """-i 0 with no deskew or rotation performs nothing (no GPU work): the
    run lists, skips and marks files only."""
    d, _ = _folder(tmp_path, [(64, 64, 64)] * 3)
    (d / "other_ch0_stack0000.tif").write_bytes((d / "cell1_ch0_stack0000.tif").read_bytes())
    otf = _otf(tmp_path)
    r = _cli(["-i", 0, d, "cell1_ch0", otf])
    assert "Found 3 file(s)" in r.stdout and r.stdout.count(">>>file_finished") == 3
    r = _cli(["-i", 0, "--skip", 2, d, "ch0", otf])
    assert "Found 4 file(s)" in r.stdout and r.stdout.count(">>>file_finished") == 2
    r = _cli(["-i", 0, d, "stack0001.tif", otf])
    assert "Found 1 file(s)" in r.stdout
    # an output in GPUdecon/ whose name starts with the stem marks the stack as done
    (d / "GPUdecon").mkdir(exist_ok=True)
    (d / "GPUdecon" / "cell1_ch0_stack0001_decon.tif").write_bytes(b"")
    r = _cli(["-i", 0, "--no_overwrite", d, "cell1", otf])
    assert "Found 2 file(s)" in r.stdout
    r = _cli(["-i", 0, "--skip", 5, d, "cell1", otf])
    assert "No files need processing" in r.stdout
    r = _cli(["-i", 0, d, "cell1", str(tmp_path / "missing_otf.tif")], check=False)
    assert r.returncode == 1 and "cannot open TIFF file" in r.stderr


def _direct(cdn, raw, otf, iters, bg=90.0, napod=15, flat=None, rescale=False, **geo):
    """RL_interface called directly (the session set up for this stack)."""
    L = cdn.lib()
    cdn.rl_init(raw.shape, otf, dzdata=geo.get("dz", 0.25), dxdata=0.104, dzpsf=geo.get("dzpsf", 0.1), dxpsf=0.104,
                deskew=geo.get("deskew", 0.0), rotate=geo.get("rotate", 0.0))
    if flat is not None:
        assert L.cudaDecon_set_flat_start(1, float(flat)) == 1
    else:
        L.cudaDecon_set_flat_start(0, 0.0)
    out = cdn.rl_decon(raw, background=bg, n_iters=iters, napodize=napod, rescale=rescale,
                       save_deskewed=geo.get("save_deskewed", False))
    return out


@pytest.mark.gpu
def test_cli_outputs_equal_rl_interface(cdn, tmp_path):
    shapes = [(64, 96, 128), (64, 96, 128), (48, 64, 96)]
    d, stacks = _folder(tmp_path, shapes)
    otf = _otf(tmp_path)
    print("cli: plain RL, MIPs, crop, flat start", flush=True)
    r = _cli(["-i", 6, "-M", 0, 1, 1, "--FlatStart", d, "cell1_ch0_stack", otf])
    assert r.stdout.count(">>>file_finished") == 3 and "Processed 3 images" in r.stdout
    cdn.rl_cleanup()
    for i, raw in enumerate(stacks):
        got = _read(str(d / "GPUdecon" / f"cell1_ch0_stack{i:04d}_decon.tif"))
        ref = _direct(cdn, raw, otf, 6, flat=float(np.median(raw.astype(np.float64))))
        assert got.shape == ref.shape and np.array_equal(got, ref), i
        assert np.array_equal(_read(str(d / "GPUdecon" / "MIPs" / f"cell1_ch0_stack{i:04d}_MIP_z.tif"))[0],
                              ref.max(axis=0))
        assert np.array_equal(_read(str(d / "GPUdecon" / "MIPs" / f"cell1_ch0_stack{i:04d}_MIP_y.tif"))[0],
                              ref.max(axis=1))
        assert not (d / "GPUdecon" / "MIPs" / f"cell1_ch0_stack{i:04d}_MIP_x.tif").exists()
    cdn.rl_cleanup()

    print("cli: bleach correction across files, uint16, final crop", flush=True)
    r = _cli(["-i", 4, "-p", "-u", "-C", 2, 101, 3, 60, 1, 40, "-b", 100, "-e", 0, d, "stack000[01]", otf])
    refs = []
    for i in range(2):
        refs.append(_direct(cdn, stacks[i], otf, 4, bg=100.0, napod=0, rescale=True))
    cdn.rl_cleanup()
    for i in range(2):
        got = _read(str(d / "GPUdecon" / f"cell1_ch0_stack{i:04d}_decon.tif"))
        want = np.clip(refs[i][1:41, 3:61, 2:102], 0, 65535).astype(np.uint16).astype(np.float32)
        assert got.shape == want.shape and np.array_equal(got, want), i

    print("cli: deskew with the deskewed raw stack", flush=True)
    d2, st2 = _folder(tmp_path / "b", [(64, 128, 96)], seed0=90)
    r = _cli(["-i", 3, "-D", 31.5, "-z", 0.3, "-Z", 0.3, "-S", "-m", 1, 0, 0, d2, "stack", otf])
    got = _read(str(d2 / "GPUdecon" / "cell1_ch0_stack0000_decon.tif"))
    desk = _read(str(d2 / "Deskewed" / "cell1_ch0_stack0000_deskewed.tif"))
    ref, ref_desk = _direct(cdn, st2[0], otf, 3, dz=0.3, dzpsf=0.3, deskew=31.5, save_deskewed=True)
    cdn.rl_cleanup()
    assert np.array_equal(got, ref) and np.array_equal(desk, ref_desk)
    mipx = _read(str(d2 / "Deskewed" / "MIPs" / "cell1_ch0_stack0000_MIP_x.tif"))[0]
    assert np.array_equal(mipx, ref_desk.max(axis=2))

    print("cli: a result that cannot be written ends the run with an error", flush=True)
    d3, _ = _folder(tmp_path / "c", [(32, 64, 64), (32, 64, 64)], seed0=7)
    (d3 / "GPUdecon" / "cell1_ch0_stack0000_decon.tif").mkdir(parents=True)  # a folder where the file goes
    r = _cli(["-i", 2, d3, "stack", otf], check=False)
    assert r.returncode == 1 and "!!Error occurred" in r.stderr, (r.returncode, r.stdout, r.stderr)


@pytest.mark.gpu
def test_cli_y_tiles_stitch_as_the_reference(cdn, tmp_path):
    """--tile 64 --TileOverlap 8 on a 160-row stack: ceil(160 / (64 - 8)) = 3
    tiles of 64 rows, overlap floor((3 64 - 160) / 2) = 16, offsets 0, 48, 96
    (src/linearDecon.cpp:514-537; the last tile repeats the stack's last row
    past its end), each through RL_interface, stitched without blending:
    rows [0, 56) from tile 0, [56, 104) from tile 1, [104, 160) from tile 2
    (src/linearDecon.cpp:1219-1226)."""
    d, stacks = _folder(tmp_path, [(32, 160, 64)], seed0=95)
    otf = _otf(tmp_path)
    r = _cli(["-i", 3, "--tile", 64, "--TileOverlap", 8, d, "stack", otf])
    assert "# of tiles: 3." in r.stdout and "Overlap: 16 Y pix" in r.stdout
    got = _read(str(d / "GPUdecon" / "cell1_ch0_stack0000_decon.tif"))
    raw = stacks[0]
    tiles = []
    for off in (0, 48, 96):
        rows = np.clip(np.arange(off, off + 64), 0, raw.shape[1] - 1)
        tiles.append(_direct(cdn, np.ascontiguousarray(raw[:, rows, :]), otf, 3))
    cdn.rl_cleanup()
    want = np.zeros_like(got)
    want[:, 0:56] = tiles[0][:, 0:56]
    want[:, 56:104] = tiles[1][:, 8:56]
    want[:, 104:160] = tiles[2][:, 8:64]
    assert got.shape == (32, 160, 64) and np.array_equal(got, want)

"""Film output (SaveImage, basic.hpp:506-672) and the nanogi command line."""
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

import nanogi_amd as ngi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "nanogi_amd", "nanogi")
CORNELL = os.path.join(ROOT, "tests", "golden", "cornelbox", "scene.yml")


def _film(w=37, h=23, seed=0):
    rng = np.random.default_rng(seed)
    f = rng.gamma(1.0, 0.5, (h, w, 3))
    f[0, 0] = [0, 0, 0]
    f[h - 1, w - 1] = [100.0, 0.001, 3.0]
    return f


def _read_pfm(path):
    data = open(path, "rb").read()
    parts = data.split(b"\n", 3)
    assert parts[0] == b"PF"
    w, h = map(int, parts[1].split())
    assert float(parts[2]) < 0   # little endian
    return np.frombuffer(parts[3], dtype="<f4").reshape(h, w, 3)


def _read_hdr(path):
    data = open(path, "rb").read()
    head, _, rest = data.partition(b"\n\n")
    assert head.startswith(b"#?RADIANCE")
    line, _, body = rest.partition(b"\n")
    _, h, _, w = line.split()
    h, w = int(h), int(w)
    out = np.zeros((h, w, 3))
    pos = 0
    for y in range(h):
        if w >= 8 and body[pos] == 2 and body[pos + 1] == 2:
            assert (body[pos + 2] << 8 | body[pos + 3]) == w
            pos += 4
            comps = []
            for _c in range(4):
                vals = []
                while len(vals) < w:
                    n = body[pos]
                    pos += 1
                    if n > 128:
                        vals += [body[pos]] * (n - 128)
                        pos += 1
                    else:
                        vals += list(body[pos:pos + n])
                        pos += n
                comps.append(vals)
            px = np.array(comps, dtype=np.float64).T
        else:
            px = np.frombuffer(body[pos:pos + 4 * w], dtype=np.uint8).reshape(w, 4).astype(np.float64)
            pos += 4 * w
        e = px[:, 3]
        scale = np.where(e > 0, np.ldexp(1.0, (e - 136).astype(int)), 0.0)
        out[y] = (px[:, :3] + 0.5) * scale[:, None] * (e[:, None] > 0)
    return out[::-1]   # file is top-down; film row 0 is the bottom


def _read_exr(path):
    data = open(path, "rb").read()
    assert struct.unpack("<I", data[:4])[0] == 20000630
    pos = 8
    attrs = {}
    while data[pos] != 0:
        name_end = data.index(b"\0", pos)
        name = data[pos:name_end].decode()
        type_end = data.index(b"\0", name_end + 1)
        size = struct.unpack("<I", data[type_end + 1:type_end + 5])[0]
        attrs[name] = data[type_end + 5:type_end + 5 + size]
        pos = type_end + 5 + size
    pos += 1
    comp = attrs["compression"][0]
    assert comp in (0, 3)   # none, or ZIP over blocks of 16 scanlines
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    lines = 16 if comp == 3 else 1
    nblocks = (h + lines - 1) // lines
    offsets = struct.unpack(f"<{nblocks}Q", data[pos:pos + 8 * nblocks])
    img = np.zeros((h, w, 3), dtype=np.float32)
    for off in offsets:
        y, size = struct.unpack("<iI", data[off:off + 8])
        n = min(lines, h - y)
        buf = data[off + 8:off + 8 + size]
        if comp == 3 and size < n * 3 * 4 * w:
            buf = _exr_unzip(buf)
        planes = np.frombuffer(buf, dtype="<f4").reshape(n, 3, w)
        img[y:y + n] = planes[:, ::-1].transpose(0, 2, 1)   # B, G, R -> R, G, B
    return img[::-1]


def _exr_unzip(buf):
    """OpenEXR ZIP block (file-format specification): inflate, undo the
    byte-delta predictor (d = t[i] - t[i-1] + 128), then interleave the
    first half (even-indexed bytes) with the second (odd-indexed bytes)."""
    d = np.frombuffer(zlib.decompress(buf), dtype=np.uint8).astype(np.int64)
    d[1:] -= 128
    t = (np.cumsum(d) & 0xFF).astype(np.uint8)
    out = np.empty_like(t)
    half = (len(t) + 1) // 2
    out[0::2], out[1::2] = t[:half], t[half:]
    return out.tobytes()


def _read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos = 8
    idat = b""
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert zlib.crc32(typ + body) & 0xFFFFFFFF == crc
        if typ == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert depth == 8 and ctype == 2
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, dtype=np.uint8).reshape(h, 1 + 3 * w)
    assert np.all(rows[:, 0] == 0)
    return rows[:, 1:].reshape(h, w, 3)[::-1]


def test_pfm_round_trip(tmp_path):
    f = _film()
    ngi.save_image(tmp_path / "a.pfm", f)
    np.testing.assert_array_equal(_read_pfm(tmp_path / "a.pfm"), f.astype(np.float32))


def test_exr_round_trip(tmp_path):
    f = _film()
    ngi.save_image(tmp_path / "sub" / "a.exr", f)   # parent directory is created
    np.testing.assert_array_equal(_read_exr(tmp_path / "sub" / "a.exr"), f.astype(np.float32))


@pytest.mark.parametrize("h", [16, 37, 1])
def test_exr_zip_blocks_of_a_smooth_image(tmp_path, h):
    """A smooth film compresses: ZIP blocks of 16 scanlines (the last one
    partial) decode to the exact FP32 values, and the file is smaller than
    the uncompressed pixels."""
    w = 64
    yy, xx = np.mgrid[0:h, 0:w]
    f = np.stack([0.01 * xx + 0.02 * yy, np.full((h, w), 0.5), 1.0 / (1.0 + xx + yy)], axis=2)
    path = tmp_path / "s.exr"
    ngi.save_image(path, f)
    np.testing.assert_array_equal(_read_exr(path), f.astype(np.float32))
    if h > 1:
        assert path.stat().st_size < 12 * w * h


def test_png_is_deflated(tmp_path):
    f = np.full((40, 50, 3), 0.25)
    ngi.save_image(tmp_path / "c.png", f)
    assert (tmp_path / "c.png").stat().st_size < 40 * 50 * 3 // 4
    assert np.all(_read_png(tmp_path / "c.png") == int(0.25 ** (1 / 2.2) * 255))


def test_png_reads_back_with_pillow(tmp_path):
    """An independent decoder sees the same upright, gamma-encoded pixels."""
    Image = pytest.importorskip("PIL.Image")
    h, w = 37, 50
    yy, xx = np.mgrid[0:h, 0:w]
    f = np.stack([xx / w, yy / h, np.full((h, w), 0.3)], axis=2)
    ngi.save_image(tmp_path / "g.png", f)
    ref = np.clip((np.power(f.astype(np.float32).astype(np.float64), 1 / 2.2) * 255).astype(int), 0, 255)[::-1]
    np.testing.assert_array_equal(np.asarray(Image.open(tmp_path / "g.png")), ref)


def test_hdr_round_trip(tmp_path):
    for w in (5, 37):   # flat and run-length encoded scanlines
        f = _film(w=w)
        ngi.save_image(tmp_path / "a.hdr", f)
        got = _read_hdr(tmp_path / "a.hdr")
        mx = f.max(axis=2, keepdims=True)
        assert np.all(np.abs(got - f) <= mx / 128 + 1e-30)


def test_png_gamma(tmp_path):
    f = _film()
    ngi.save_image(tmp_path / "a.png", f)
    expect = np.clip((np.power(f.astype(np.float32).astype(np.float64), 1 / 2.2) * 255).astype(int), 0, 255)
    np.testing.assert_array_equal(_read_png(tmp_path / "a.png"), expect)


def test_bad_extension(tmp_path):
    with pytest.raises(ngi.NgiError) as e:
        ngi.save_image(tmp_path / "a.bmp", _film())
    assert e.value.status == ngi.ERR_INVALID_ARG


def _cli(*args):
    return subprocess.run([CLI] + list(args), capture_output=True, text=True, timeout=120)


def test_cli_usage_exits_zero():
    r = _cli()
    assert r.returncode == 0 and "Usage: nanogi [options] <renderer> <scene> <result> <width> <height>" in r.stdout
    assert _cli("--help").returncode == 0


@pytest.mark.parametrize("args", [("-r", "ptmnee", CORNELL, "x.hdr", "8", "8"), ("ptmnee", CORNELL), ("-i", CORNELL),
                                  ("pt", "/nonexistent.yml", "x.hdr", "8", "8"), ("--bogus", "1")])
def test_cli_errors(args, tmp_path):
    r = _cli(*args)
    assert r.returncode == 1 and "ERROR" in r.stderr


@pytest.mark.gpu
def test_cli_renders_and_matches_library(tmp_path):
    out = tmp_path / "img.pfm"
    r = _cli("-r", "ptdirect", "-i", CORNELL, "-o", str(out), "-w", "40", "-h", "30", "--spp", "16", "-m", "6",
             "--seed", "3", "--rr-depth", "3")
    assert r.returncode == 0, r.stderr
    assert "# of samples: 19200" in r.stdout
    img = _read_pfm(out)
    s = ngi.Scene.load_yaml(CORNELL, 40 / 30)
    f, _ = ngi.render(s, renderer="ptdirect", width=40, height=30, sampling=ngi.SAMPLING_PER_PIXEL, spp=16,
                      max_num_vertices=6, seed=3, rr_depth=3, engine=ngi.ENGINE_PERSISTENT)
    np.testing.assert_allclose(img, f.astype(np.float32), rtol=1e-6, atol=1e-9)


@pytest.mark.gpu
def test_cli_autotune(tmp_path):
    out = tmp_path / "img.pfm"
    r = _cli("-r", "ptdirect", "-i", CORNELL, "-o", str(out), "-w", "40", "-h", "30", "--spp", "16", "-m", "6",
             "--seed", "3", "--autotune", "100000")
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("Autotune:") == 9 and "# of samples: 19200" in r.stdout
    img = _read_pfm(out)
    assert img.shape == (30, 40, 3) and np.isfinite(img).all() and img.mean() > 0


@pytest.mark.gpu
def test_cli_reference_sampling_and_progress_images(tmp_path):
    out = tmp_path / "img.hdr"
    fmt = str(tmp_path / "prog" / "{{count}}.pfm")
    r = _cli("pt", CORNELL, str(out), "32", "24", "-n", "20000", "--progress-image-update-interval", "0.0001",
             "--progress-image-update-format", fmt)
    assert r.returncode == 0, r.stderr
    assert out.exists() and (tmp_path / "prog" / "0000000001.pfm").exists()
    img = _read_hdr(out)
    assert img.shape == (24, 32, 3) and np.isfinite(img).all() and img.mean() > 0


@pytest.mark.gpu
def test_cli_time_budget_renders_whole_chunks(tmp_path):
    """-t renders chunks of GrainSize * 1000 paths until the time is up
    (src/nanogi.cpp:243-418): the sample count is a whole number of chunks
    and the elapsed time reaches the budget."""
    out = tmp_path / "img.pfm"
    r = _cli("pt", CORNELL, str(out), "32", "24", "-t", "0.3", "--grain-size", "10")
    assert r.returncode == 0, r.stderr
    n = int(r.stdout.split("# of samples: ")[1].split()[0])
    t = float(r.stdout.split("Elapsed time: ")[1].split()[0])
    assert n >= 10000 and n % 10000 == 0 and t >= 0.3
    img = _read_pfm(out)
    assert img.shape == (24, 32, 3) and np.isfinite(img).all() and img.mean() > 0

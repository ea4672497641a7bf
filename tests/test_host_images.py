"""Image writers of the pass layer (Image::save_to_file, texture.cu:96-160,
:219-230): the film is flipped in x and y, PFM keeps float32, EXR is an
OpenEXR 2.0 scanline file with HALF channels, read back here by an
independent parser of the published file layout (magic, attribute header,
offset table, per-line channel planes in alphabetical channel order); HDR is
Radiance RGBE, read back by a parser of the published format (header, flat and
run-length scanlines); PNG, BMP and JPEG hold the reference's 8-bit values
(clamp to [0, 1], times 255, truncated: Image::f_to_u, texture.cu:11-21) and are
decoded with Pillow (exact for PNG and BMP, within a few levels for JPEG).
CPU only."""
import ctypes
import os
import struct

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOSTLIB = os.path.join(ROOT, "srt_amd", "libsrt_host.so")


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(HOSTLIB):
        pytest.fail("libsrt_host.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(HOSTLIB)
    lib.srt_host_save_image.restype = ctypes.c_int
    lib.srt_host_save_image.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    return lib


def save(host, path, rgba, W, H):
    a = np.ascontiguousarray(rgba, np.float32)
    assert host.srt_host_save_image(str(path).encode(), W, H, a.ctypes.data) == 0


def read_exr(path):
    b = open(path, "rb").read()
    assert struct.unpack_from("<I", b, 0)[0] == 20000630
    assert struct.unpack_from("<I", b, 4)[0] & 0xff == 2
    pos, attrs = 8, {}
    while b[pos] != 0:
        end = b.index(b"\0", pos); name = b[pos:end].decode(); pos = end + 1
        end = b.index(b"\0", pos); typ = b[pos:end].decode(); pos = end + 1
        size = struct.unpack_from("<i", b, pos)[0]; pos += 4
        attrs[name] = (typ, b[pos:pos + size]); pos += size
    pos += 1
    chl, channels = attrs["channels"][1], []
    q = 0
    while chl[q] != 0:
        end = chl.index(b"\0", q)
        channels.append((chl[q:end].decode(), struct.unpack_from("<i", chl, end + 1)[0]))
        q = end + 1 + 16
    assert attrs["compression"][1] == b"\0"
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    W, H = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from(f"<{H}Q", b, pos)
    img = {}
    for c, _ in channels:
        img[c] = np.zeros((H, W), np.float32)
    for off in offsets:
        y, size = struct.unpack_from("<ii", b, off)
        data = np.frombuffer(b, "<f2", count=size // 2, offset=off + 8).reshape(len(channels), W)
        for k, (c, _) in enumerate(channels):
            img[c][y] = data[k]
    return channels, img


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = map(int, f.readline().split())
        assert float(f.readline()) < 0
        return np.frombuffer(f.read(), "<f4").reshape(h, w, 3)[::-1]


def test_exr_half_channels_and_orientation(host, tmp_path):
    W, H = 7, 5
    rng = np.random.default_rng(1)
    film = rng.uniform(0, 8, size=(W * H, 4)).astype(np.float32)
    film[:, 3] = rng.uniform(0, 1, size=W * H)
    film[0, :3] = [65519.0, 1e-8, 6.0e-5]          # largest finite half, flushed, subnormal
    save(host, tmp_path / "a.exr", film, W, H)
    save(host, tmp_path / "a.pfm", film, W, H)
    channels, img = read_exr(tmp_path / "a.exr")
    assert channels == [("A", 1), ("B", 1), ("G", 1), ("R", 1)]
    expect = film.reshape(H, W, 4)[::-1, ::-1]                   # flip y and x
    for k, c in enumerate("RGBA"):
        assert np.array_equal(img[c], expect[..., k].astype(np.float16).astype(np.float32)), c
    rgb = np.stack([img[c] for c in "RGB"], -1)
    assert np.allclose(rgb, read_pfm(tmp_path / "a.pfm"), rtol=1e-3, atol=1e-7)
    assert img["R"][-1, -1] == np.float32(np.float16(65519.0))


def test_unknown_extension_fails(host, tmp_path):
    film = np.zeros((4, 4), np.float32)
    assert host.srt_host_save_image(str(tmp_path / "a.tga").encode(), 2, 2, film.ctypes.data) == -1


def ldr(film):
    """The reference's 8-bit conversion of the flipped film (float32 arithmetic)."""
    return (np.clip(film, 0.0, 1.0).astype(np.float32) * np.float32(255.0)).astype(np.uint8)


def ldr_film(W, H, seed):
    rng = np.random.default_rng(seed)
    film = rng.uniform(-0.3, 1.3, size=(W * H, 4)).astype(np.float32)
    film[:4, 0] = [0.5, 1.0, 0.999, 1.0 / 255.0]                 # 127.5 -> 127, 255, 254.7 -> 254, 1
    film[:, 3] = 1.0
    return film


def test_png_and_bmp_hold_the_reference_8bit_values(host, tmp_path):
    Image = pytest.importorskip("PIL.Image")
    W, H = 13, 6
    film = ldr_film(W, H, 3)
    film[5, 3] = 0.25                                              # an alpha below 1 survives
    expect = ldr(film.reshape(H, W, 4)[::-1, ::-1])
    assert expect[-1, -1, 0] == 127 and expect[-1, -2, 0] == 255 and expect[-1, -3, 0] == 254
    save(host, tmp_path / "a.png", film, W, H)
    png = np.asarray(Image.open(tmp_path / "a.png").convert("RGBA"))
    assert png.shape == (H, W, 4)
    assert np.array_equal(png, expect)
    save(host, tmp_path / "a.bmp", film, W, H)
    im = Image.open(tmp_path / "a.bmp")
    bmp = np.asarray(im.convert("RGBA" if "A" in im.mode else "RGB"))
    assert np.array_equal(bmp[..., :3], expect[..., :3])
    if bmp.shape[-1] == 4:
        assert np.array_equal(bmp[..., 3], expect[..., 3])


@pytest.mark.parametrize("W,H", [(37, 23), (8, 8), (1, 1)])
def test_jpeg_decodes_close_to_the_8bit_values(host, tmp_path, W, H):
    Image = pytest.importorskip("PIL.Image")
    y, x = np.mgrid[0:H, 0:W].astype(np.float32)
    film = np.ones((H, W, 4), np.float32)
    film[..., 0] = x / max(W - 1, 1)
    film[..., 1] = y / max(H - 1, 1)
    film[..., 2] = 0.5 + 0.4 * np.sin(x * 0.7 + y * 0.3)
    film[H // 2:, W // 2:, :3] = [0.9, 0.1, 0.2]                   # a sharp-edged patch
    film = film.reshape(-1, 4)
    save(host, tmp_path / "a.jpg", film, W, H)
    im = Image.open(tmp_path / "a.jpg")
    assert im.format == "JPEG" and im.size == (W, H)
    got = np.asarray(im.convert("RGB")).astype(np.int32)
    expect = ldr(film.reshape(H, W, 4)[::-1, ::-1])[..., :3].astype(np.int32)
    err = np.abs(got - expect)
    assert err.mean() < 1.0, err.mean()     # observed 0.43 on 37 x 23
    assert err.max() <= 4, err.max()         # observed 2


def read_hdr(path):
    b = open(path, "rb").read()
    assert b.startswith(b"#?RADIANCE\n")
    pos = 0
    lines = []
    while True:
        end = b.index(b"\n", pos)
        line = b[pos:end]
        pos = end + 1
        if line == b"" and lines:
            break
        lines.append(line)
    assert b"FORMAT=32-bit_rle_rgbe" in lines
    end = b.index(b"\n", pos)
    ys, h, xs, w = b[pos:end].split()
    pos = end + 1
    assert ys == b"-Y" and xs == b"+X"
    H, W = int(h), int(w)
    out = np.zeros((H, W, 4), np.uint8)
    for row in range(H):
        if 8 <= W <= 0x7fff and b[pos] == 2 and b[pos + 1] == 2 and b[pos + 2] < 128:
            assert (b[pos + 2] << 8 | b[pos + 3]) == W
            pos += 4
            for c in range(4):
                x = 0
                while x < W:
                    n = b[pos]
                    pos += 1
                    if n > 128:                                     # run of one value
                        out[row, x:x + n - 128, c] = b[pos]
                        pos += 1
                        x += n - 128
                    else:                                           # literal bytes
                        out[row, x:x + n, c] = np.frombuffer(b, np.uint8, n, pos)
                        pos += n
                        x += n
        else:
            out[row] = np.frombuffer(b, np.uint8, 4 * W, pos).reshape(W, 4)
            pos += 4 * W
    assert pos == len(b)
    e = out[..., 3].astype(np.int32)
    scale = np.where(e > 0, np.ldexp(1.0, e - 136), 0.0)
    return out[..., :3].astype(np.float64) * scale[..., None]


@pytest.mark.parametrize("W,H", [(19, 4), (5, 3)])
def test_hdr_rgbe_round_trip(host, tmp_path, W, H):
    rng = np.random.default_rng(W)
    film = np.ones((W * H, 4), np.float32)
    film[:, :3] = np.exp(rng.uniform(-6, 6, size=(W * H, 3))).astype(np.float32)
    film[0, :3] = 0.0                                              # black
    film[1, :3] = [3.0, 0.0, 1e-30]                                # one channel only
    save(host, tmp_path / "a.hdr", film, W, H)
    got = read_hdr(tmp_path / "a.hdr")
    expect = film.reshape(H, W, 4)[::-1, ::-1, :3].astype(np.float64)
    # a shared exponent: each channel is truncated to 1/256 of the largest channel's power of two
    tol = np.ldexp(1.0, np.frexp(expect.max(-1))[1] - 8)[..., None]
    assert np.all(got <= expect * (1 + 1e-6) + 1e-30)
    assert np.all(expect - got <= tol * (1 + 1e-6))
    assert got[-1, -1].tolist() == [0.0, 0.0, 0.0]
    assert got[-1, -2, 0] == 3.0


def test_unknown_extension_creates_no_file(host, tmp_path):
    film = np.zeros((4, 4), np.float32)
    assert host.srt_host_save_image(str(tmp_path / "a.webp").encode(), 2, 2, film.ctypes.data) == -1
    assert not (tmp_path / "a.webp").exists()

"""Image writers of the pass layer (Image::save_to_file, texture.cu:96-117,
:219-230): the film is flipped in x and y, PFM keeps float32, and EXR is an
OpenEXR 2.0 scanline file with HALF channels, read back here by an
independent parser of the published file layout (magic, attribute header,
offset table, per-line channel planes in alphabetical channel order).
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

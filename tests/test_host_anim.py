"""Keyframed node animation of the pass layer (the scene config's
"animation" list, renderer.cu:71-121; animation.cpp; quaternion.h Slerp;
transform.h Compose = T * R * S) against a float64 restatement: linear
translation and scale, spherical rotation, repeat extrapolation, and the
channel defaults.  CPU only."""
import ctypes
import json
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOSTLIB = os.path.join(ROOT, "srt_amd", "libsrt_host.so")


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(HOSTLIB):
        pytest.fail("libsrt_host.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(HOSTLIB)
    lib.srt_host_node_transform.restype = ctypes.c_int
    lib.srt_host_node_transform.argtypes = [ctypes.c_char_p, ctypes.c_float, ctypes.c_void_p, ctypes.c_char_p,
                                            ctypes.c_int]
    return lib


def transform(host, anim, t):
    out = np.zeros(12, np.float32)
    err = ctypes.create_string_buffer(256)
    r = host.srt_host_node_transform(json.dumps(anim).encode(), t, out.ctypes.data, err, 256)
    return r, out.reshape(3, 4), err.value.decode()


def quat_matrix(q):
    x, y, z, w = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def slerp(a, b, f):
    a, b = np.asarray(a, float), np.asarray(b, float)
    c = np.dot(a, b)
    if c < 0:
        b, c = -b, -c
    th = np.arccos(min(c, 1.0))
    return (np.sin((1 - f) * th) * a + np.sin(f * th) * b) / np.sin(th)


def expect(keys, tr, rot, sc, t):
    t = np.fmod(t, keys[-1])
    i = max(0, min(np.searchsorted(keys, t) - 1, len(keys) - 2))
    f = (t - keys[i]) / (keys[i + 1] - keys[i])
    T = np.asarray(tr[i]) * (1 - f) + np.asarray(tr[i + 1]) * f
    R = quat_matrix(slerp(rot[i], rot[i + 1], f))
    S = np.asarray(sc[i]) * (1 - f) + np.asarray(sc[i + 1]) * f
    M = np.zeros((3, 4))
    M[:, :3] = R * S[None, :]
    M[:, 3] = T
    return M


def test_trs_keyframes_match_float64(host):
    s2 = np.sqrt(0.5)
    keys = [0.0, 1.0, 3.0]
    tr = [[0, 0, 0], [1, 2, 3], [-1, 0, 2]]
    rot = [[0, 0, 0, 1], [0, s2, 0, s2], [s2, 0, 0, s2]]
    sc = [[1, 1, 1], [2, 2, 2], [1, 0.5, 1]]
    anim = {"name": "box", "keys": keys, "translations": tr, "rotations": rot, "scales": sc}
    for t in [0.0, 0.25, 0.5, 0.999, 1.5, 2.75, 3.5, 7.25]:      # 3.5 and 7.25 wrap (repeat)
        r, M, _ = transform(host, anim, t)
        assert r == 0
        assert np.abs(M - expect(np.array(keys), tr, rot, sc, t)).max() < 2e-5, t


def test_channel_defaults_and_single_keys(host):
    r, M, _ = transform(host, {"name": "a", "keys": [0.0, 2.0], "translations": [[0, 0, 0], [4, 0, 0]]}, 1.0)
    assert r == 0 and np.allclose(M, [[1, 0, 0, 2], [0, 1, 0, 0], [0, 0, 1, 0]], atol=1e-7)
    r, M, _ = transform(host, {"name": "a", "keys": [0.0], "scales": [[3, 3, 3]]}, 0.7)
    assert r == 0 and np.allclose(M, [[3, 0, 0, 0], [0, 3, 0, 0], [0, 0, 3, 0]], atol=1e-7)


def test_bad_animation_entries(host):
    assert transform(host, {"name": "a", "keys": []}, 0.0)[0] == -1
    r, _, err = transform(host, {"name": "a", "keys": [0, 1], "translations": [[0, 0, 0]]}, 0.0)
    assert r == -1 and "one entry per key" in err

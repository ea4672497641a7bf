"""The C++ pass layer mirrors the reference's RenderPass interface: pass names,
JSON keys and defaults (path_tracer.h:34-38, tonemapping.h:39-41,
wavefront.h:66-70, simple.h:33-39), unknown-pass errors, and config parsing with comments
(main.cu:5-17).  CPU only: no device is touched."""
import ctypes
import json
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOSTLIB = os.path.join(ROOT, "srt_amd", "libsrt_host.so")


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(HOSTLIB):
        pytest.fail("libsrt_host.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(HOSTLIB)
    lib.srt_host_pass_params.restype = ctypes.c_int
    lib.srt_host_pass_params.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    return lib


def params(host, name, p=None):
    buf = ctypes.create_string_buffer(4096)
    r = host.srt_host_pass_params(name.encode(), (json.dumps(p) if p is not None else "").encode(), buf, 4096)
    return r, buf.value.decode()


@pytest.mark.parametrize("name", ["PathTracer", "Wavefront"])
def test_path_tracer_defaults(host, name):
    r, out = params(host, name)
    assert r > 0
    assert json.loads(out) == {"samples_per_pixel": 1, "max_depth": 16, "rr_depth": 4, "use_nee": True,
                               "use_mis": True, "seed": 0}
    r, out = params(host, name, {"samples_per_pixel": 16, "max_depth": 5, "use_mis": False})
    assert json.loads(out) == {"samples_per_pixel": 16, "max_depth": 5, "rr_depth": 4, "use_nee": True,
                               "use_mis": False, "seed": 0}


def test_tonemapping_defaults_and_names(host):
    r, out = params(host, "ToneMapping")
    assert json.loads(out) == {"type": "None", "exposure": 0, "use_gamma": True}
    for t in ["Clamp", "Reinhard", "Uncharted2", "ACES"]:
        assert json.loads(params(host, "ToneMapping", {"type": t, "exposure": 1.5})[1])["type"] == t
    # unknown operator names fall back to None (tonemapping.h:27-35)
    assert json.loads(params(host, "ToneMapping", {"type": "Filmic"})[1])["type"] == "None"


def test_light_tracer_defaults(host):
    r, out = params(host, "LightTracer")
    assert r > 0
    assert json.loads(out) == {"samples_per_pixel": 1, "max_depth": 16, "rr_depth": 4, "seed": 0}
    assert json.loads(params(host, "LightTracer", {"samples_per_pixel": 8, "max_depth": 5})[1]) == \
        {"samples_per_pixel": 8, "max_depth": 5, "rr_depth": 4, "seed": 0}


def test_simple_defaults_and_names(host):
    r, out = params(host, "Simple")
    assert r > 0
    assert json.loads(out) == {"type": "Depth", "samples_per_pixel": 16, "seed": 0}
    for t in ["Normal", "BaseColor", "Ambient", "FaceOrientation"]:
        assert json.loads(params(host, "Simple", {"type": t, "samples_per_pixel": 4})[1]) == \
            {"type": t, "samples_per_pixel": 4, "seed": 0}
    # unknown names read as Depth (simple.h:24-31)
    assert json.loads(params(host, "Simple", {"type": "Curvature"})[1])["type"] == "Depth"


def test_accumulate_denoise_and_unknown_pass(host):
    assert params(host, "Accumulate")[0] >= 0
    assert params(host, "Denoise")[0] >= 0
    r, out = params(host, "Bloom")
    assert r == -1 and "Render pass Bloom not found!" in out


def test_config_comments_and_bad_json(host):
    r, out = params(host, "PathTracer", None)
    assert r > 0
    buf = ctypes.create_string_buffer(1024)
    txt = b'{ // comment\n "max_depth": 3 /* block */ }'
    assert host.srt_host_pass_params(b"PathTracer", txt, buf, 1024) > 0
    assert json.loads(buf.value)["max_depth"] == 3
    assert host.srt_host_pass_params(b"PathTracer", b'{"max_depth": }', buf, 1024) == -1
    assert b"JSON parse error" in buf.value

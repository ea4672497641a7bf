"""SRT path-tracing hot path for AMD Instinct MI355X (gfx950).

The product is the HIP library srt_amd/libsrt_hip.so behind the C ABI in
include/srt_hip.h, and the C++ render passes in srt_amd/host that mirror the
reference's RenderPass interface.  This package holds the Python bindings
(abi.py), the procedural workloads (scenes.py) and the build script.
"""
from . import scenes  # noqa: F401

__all__ = ["scenes"]

"""Host-side checks of the size limits in srt_amd/csrc/limits.h, compiled with
g++ alone (no GPU): the 4-wide node array's byte size and per-lane offsets stay
below 2^31 up to kMaxNodes4 and not beyond it (the 32-bit buffer-resource
addressing of the traversal), the triangle limit fits the 28-bit leaf field,
and the path-pool range fits 32-bit slot indices."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "srt_amd", "csrc", "limits.h")

PROG = r'''
#include "limits.h"
#include <cstdio>
#include <cstdint>
using namespace srt;
int main() {
    int bad = 0;
    // byte size of the array, as the traversal's buffer resource computes it (int32)
    if (!(kMaxNodes4 * kNode4Bytes <= INT32_MAX)) bad |= 1;
    if (!((kMaxNodes4 + 1) * kNode4Bytes > INT32_MAX)) bad |= 2;
    // the last node's last byte as a 32-bit lane offset
    if (!((uint64_t)(kMaxNodes4 - 1) * 128u + 127u < 0x80000000ull)) bad |= 4;
    if (!nodes4_fit(kMaxNodes4) || nodes4_fit(kMaxNodes4 + 1) || nodes4_fit(-1)) bad |= 8;
    if (!nodes4_fit(1000, 1000) || nodes4_fit(1001, 1000) || nodes4_fit(kMaxNodes4 + 1, INT64_MAX)) bad |= 16;
    // leaf refs: bit 31 | first << 3 | (count - 1) with first <= kMaxTriangles - 1
    if (!(((uint64_t)(kMaxTriangles - 1) << 3 | 7u) < 0x80000000ull)) bad |= 32;
    if (!path_pool_ok(64) || path_pool_ok(63) || !path_pool_ok(0x7fffffffll) || path_pool_ok(0x80000000ll)) bad |= 64;
    std::printf("%d %lld\n", bad, (long long)kMaxNodes4);
    return bad;
}
'''


def test_limits_compile_and_hold(tmp_path):
    src = tmp_path / "limits_test.cpp"
    src.write_text(PROG)
    exe = tmp_path / "limits_test"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.dirname(HDR), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert r.stdout.split() == ["0", "16777215"]


def test_header_documents_the_limits():
    text = open(os.path.join(ROOT, "include", "srt_hip.h")).read()
    assert "2^28 - 1 triangles" in text and "16777215" in text and "[64, 2^31 - 1]" in text

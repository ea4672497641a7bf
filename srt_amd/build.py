"""Build the HIP library (srt_amd/libsrt_hip.so) and the host C++ passes for gfx950.

Every HIP source is compiled with strict IEEE single precision (no FMA
contraction, correctly rounded division and square root, denormals kept) so
the kernels reproduce a host execution of the same formulas bit for bit, and
with packed FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsrt_hip.so")
HOSTLIB = os.path.join(HERE, "libsrt_host.so")
SRTRACER = os.path.join(HERE, "srtracer")

HIP_SOURCES = ["srt_api.hip", "wavefront.hip", "bvh.hip", "mlp.hip", "serve.hip", "instances.hip"]
# per-source code-generation options: the MLP kernels keep MFMA accumulators in
# ordinary VGPRs (their epilogues read every element; the AGPR form would add a
# v_accvgpr_read per element)
HIP_SOURCE_FLAGS = {"mlp.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}
HOST_SOURCES = ["json.cpp", "scene_io.cpp", "image_out.cpp", "passes.cpp", "renderer.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SRT_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"),
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
             "-I", os.path.join(ROOT, "include"), "-I", HOST]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd[:4])} ...")
    if verbose and r.stdout.strip():
        # hipcc's host pass warns about the packed-fp32 target feature; keep other output
        lines = [l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l]
        if lines:
            print("\n".join(lines))


def build_hip(verbose=False, jobs=4):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(BUILD, s.replace(".hip", ".o")) for s in srcs]

    def compile_one(pair):
        src, obj = pair
        _run([HIPCC, *HIP_FLAGS, *HIP_SOURCE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj], verbose)

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(compile_one, zip(srcs, objs)))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_host(verbose=False):
    srcs = [os.path.join(HOST, s) for s in HOST_SOURCES if os.path.exists(os.path.join(HOST, s))]
    if not srcs:
        return None
    tmp = HOSTLIB + ".tmp"
    _run(["g++", *CXX_FLAGS, "-shared", "-o", tmp, *srcs, "-L", HERE, "-lsrt_hip",
          f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN", "-lpthread"], verbose)
    os.replace(tmp, HOSTLIB)
    main = os.path.join(HOST, "main.cpp")
    if os.path.exists(main):
        tmp = SRTRACER + ".tmp"
        _run(["g++", *CXX_FLAGS, "-o", tmp, main, "-L", HERE, "-lsrt_host", "-lsrt_hip",
              "-Wl,-rpath,$ORIGIN", "-lpthread"], verbose)
        os.replace(tmp, SRTRACER)
    return HOSTLIB


def build_oracle(verbose=False):
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "-s"], verbose)
    return os.path.join(odir, "build", "libsrt_oracle.so")


def build_all(verbose=False):
    build_hip(verbose)
    build_host(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIB)

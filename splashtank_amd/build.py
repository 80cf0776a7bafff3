"""Build recipe of the MI355X FLIP library (and its test-side checker).

``build_hip()`` compiles the HIP sources for gfx950 into
``splashtank_amd/libsplashtank_hip.so`` (the C ABI of ``include/splashtank.h``).
``build_napi()`` compiles the Node N-API addon that binds the same C ABI for the
JS ``FlipSim`` facade.  ``build_oracle()`` compiles the CPU oracle used by the
tests.  Every call compiles, also when an output already exists.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libsplashtank_hip.so")
ADDON = os.path.join(HERE, "flip", "splashtank_hip.node")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# contraction off: every f32 rounding is the one the source writes; divide and
# sqrt correctly rounded; denormals kept; no packed-f32 instructions.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", INCLUDE, "-I", CSRC,
]
SOURCES = ["flip_kernels.hip", "flip_adjoint.hip", "flip_norms.hip", "flip_sdf.hip", "flip_ocean.hip", "flip_capi.hip"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def build_hip(verbose: bool = False) -> str:
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(bdir, src.replace(".hip", ".o"))
        out = _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        if verbose and out.strip():
            print(out)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_napi() -> str | None:
    """N-API addon for Node (headers in /usr/include/node); skipped without them."""
    node_inc = "/usr/include/node"
    if not os.path.isfile(os.path.join(node_inc, "node_api.h")):
        return None
    src = os.path.join(CSRC, "napi_addon.cpp")
    if not os.path.isfile(src):
        return None
    tmp = ADDON + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-DNODE_GYP_MODULE_NAME=splashtank_hip",
          "-I", node_inc, "-I", INCLUDE, src,
          "-L", HERE, "-lsplashtank_hip", "-Wl,-rpath,$ORIGIN/..", "-o", tmp])
    os.replace(tmp, ADDON)
    return ADDON


def build_checks() -> str:
    """Device self-checks the GPU tests run (tests/c/*.hip), as a library beside them."""
    src = os.path.join(ROOT, "tests", "c", "convert_check.hip")
    out = os.path.join(ROOT, "tests", "c", "libconvert_check.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", src, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_oracle() -> str:
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-B", "-C", odir, "liboracle.so"])
    return os.path.join(odir, "liboracle.so")


REF_SRC = "/root/reference/native/splash.cpp"


def build_ref() -> str | None:
    """The reference's CPU baker (oracle/_ref/splash), where its source is present."""
    if not os.path.exists(REF_SRC):
        return None
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-B", "-C", odir, "ref", f"REF_SRC={REF_SRC}"])
    return os.path.join(odir, "_ref", "splash")


if __name__ == "__main__":
    print(build_hip(verbose=True))
    print(build_napi())
    print(build_checks())
    print(build_oracle())
    print(build_ref())

"""Builds liblbm (HIP, gfx950) and the Node addon in this directory.

    python build.py            # lib/liblbm.so and lib/lbm.node

hipcc cross-compiles gfx950 code without a GPU.  Device code is built with
-ffp-contract=off (every multiply and add rounds separately, as the reference's
double arithmetic does) and without packed FP32 instructions."""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result",
]


def run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "liblbm.so")
    tmp = out + ".tmp"
    run([HIPCC, *HIP_FLAGS, "-I" + INCLUDE, "-I" + CSRC, "-o", tmp,
         os.path.join(CSRC, "lbm.hip")])
    os.replace(tmp, out)
    return out


def build_node_addon():
    """N-API addon (lib/lbm.node) over liblbm; needs Node's headers."""
    inc = "/usr/include/node"
    src = os.path.join(HERE, "node", "lbm_addon.cc")
    if not (os.path.exists(os.path.join(inc, "node_api.h")) and os.path.exists(src)):
        print("node addon: skipped (no node_api.h or no source)")
        return None
    out = os.path.join(LIB, "lbm.node")
    cxx = shutil.which("g++") or "g++"
    run([cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + inc, "-I" + INCLUDE,
         "-DNAPI_VERSION=8", "-o", out, src, "-L" + LIB, "-llbm",
         "-Wl,-rpath,$ORIGIN"])
    return out


def main():
    build_lib()
    build_node_addon()


if __name__ == "__main__":
    sys.exit(main())

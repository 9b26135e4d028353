"""Build recipe for the gfx950 routing library (hipcc, in-tree)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libiitmgr.so")
SOURCES = [os.path.join(HERE, "csrc", "gr_route.hip"), os.path.join(HERE, "csrc", "gr_mlp.hip"),
           os.path.join(HERE, "csrc", "gr_eval.hip"),
           os.path.join(HERE, "csrc", "gr_host.cpp"), os.path.join(HERE, "csrc", "gr_io.cpp")]
HEADERS = [os.path.join(HERE, "csrc", "gr_kernels.h"), os.path.join(HERE, "csrc", "gr_internal.h"), os.path.join(HERE, "csrc", "gr_device.h"),
           os.path.join(HERE, "..", "include", "iitmgr.h")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    # packed fp32 instructions stay off on this hardware
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    cmd = [hipcc] + HIPCC_FLAGS + ["-o", tmp] + SOURCES
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))

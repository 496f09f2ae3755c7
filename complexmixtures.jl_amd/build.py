"""Build recipe of libcmgpu.so (gfx950).

Compiles csrc/cmgpu.hip with hipcc into the package directory, where the
ctypes loader (_lib.py) finds it.  Flags:
  -ffp-contract=off        expressions evaluate as written; fused multiply-adds
                           only where fma()/fmaf() are spelled out (bit
                           reproducibility of the FP64 path)
  -packed-fp32-ops off     no packed FP32 instructions
"""
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
SRC = os.path.join(PKG_DIR, "csrc", "cmgpu.hip")
OUT = os.path.join(PKG_DIR, "libcmgpu.so")
ARCH = os.environ.get("CM_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if p and (os.path.exists(p) or p == "hipcc"):
            return p
    raise RuntimeError("hipcc not found")


def command(out=OUT, extra=()):
    return [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            # a local that shadows another (e.g. an LDS pointer hidden by a
            # private array of the same name) is an error, not a warning
            "-Werror=shadow",
            "-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG_DIR, "csrc"),
            *extra, "-o", out, SRC]


def build(verbose=False):
    tmp = OUT + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libcmgpu.so")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))

"""Build recipe of the HIP library misc_amd/libastro_hip.so (gfx950).

hipcc compiles the kernels, the host orchestration / C ABI and the host
book-keeping into one shared library in-tree, so it travels with the source
tree.  Packed FP32 instructions are disabled for all device code.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libastro_hip.so")
SOURCES = ["astro_kernels.hip", "astro_abi.hip", "astro_gnn.hip", "astro_batch.hip", "astro_design.hip", "astro_prep.cpp",
           "astro_shell.cpp"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc] + FLAGS + [os.path.join(SRC, s) for s in SOURCES] + ["-o", OUT + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libastro_hip.so")
    if verbose:
        sys.stdout.write(r.stdout)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))

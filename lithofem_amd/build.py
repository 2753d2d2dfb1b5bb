"""Build the gfx950 assembly library (lithofem_amd/liblfasm.so) with hipcc.

Run ``python -m lithofem_amd.build``.  The library is built in-tree so it
travels with the source tree; nothing is JIT-compiled at import time.
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = [HERE / "csrc" / "lfasm_api.hip"]
DEPS = [HERE / "csrc" / "lfasm_device.hpp", HERE.parent / "include" / "lithofem_asm.h"]
OUT = HERE / "liblfasm.so"
ARCH = os.environ.get("LFASM_ARCH", "gfx950")

FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
         # no packed-FP32 instructions anywhere in the device code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-result", "-lpthread"]


# the device-ceiling probe of bench.py --mode peaks: its own library
PEAKS_SRC = HERE / "csrc" / "peaks.hip"
PEAKS_OUT = HERE / "libpeaks.so"


def _hipcc(src, out, verbose):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = out.with_suffix(".so.tmp")
    cmd = [hipcc, *FLAGS, "-o", str(tmp), *map(str, src)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)


def build(verbose: bool = False) -> Path:
    _hipcc(SRC, OUT, verbose)
    _hipcc([PEAKS_SRC], PEAKS_OUT, verbose)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    print(OUT, file=sys.stderr)

"""Compile the HIP library ``libwfx_xas.so`` for gfx950 (in-tree).

``python -m waveformx_amd.build`` or ``waveformx_amd.build.build()``.
Floating-point contraction is off for the whole library (the setup's
arithmetic is the reference's unfused one; the FP32 kernels write their
fused multiply-adds explicitly), and packed FP32 instructions are disabled.
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
SOURCES = [PACKAGE / "csrc" / "xas_kernels.hip"]
HEADERS = sorted((PACKAGE / "csrc").glob("*.h")) + sorted((PACKAGE / "csrc").glob("*.inc"))
LIBRARY = PACKAGE / "libwfx_xas.so"
ARCH = os.environ.get("WFX_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    path = Path(rocm) / "bin" / "hipcc"
    return str(path) if path.exists() else "hipcc"


def build(verbose: bool = False) -> Path:
    """Compile the library (always) and return its path."""
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [hipcc(), *FLAGS, "-I", str(PACKAGE / "csrc"), *map(str, SOURCES), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)

"""Build the gfx950 shared library of the stage kernels in-tree.

The library is plain C-ABI (include/snapy_amd.h): hipcc compiles every
csrc/*.hip file for gfx950 and links them into snapy_amd/_snapy_amd.so.  It
links against the HIP runtime that torch ships (same soname), so loading it
after ``import torch`` shares torch's runtime, streams and allocator.
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
LIB = HERE / "_snapy_amd.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: FP64 parity with the reference CPU path needs every
# product and sum rounded separately.  packed-fp32-ops off: some packed FP32
# instruction forms have given wrong results on this hardware.
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _torch_lib_dir() -> str | None:
    try:
        import importlib.util

        spec = importlib.util.find_spec("torch")
        if spec is None or spec.origin is None:
            return None
        return str(Path(spec.origin).parent / "lib")
    except Exception:  # pragma: no cover
        return None


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def build(verbose: bool = False) -> Path:
    """Compile every kernel source into LIB; raises on a compiler error."""
    srcs = [str(s) for s in sources()]
    cmd = [HIPCC, *FLAGS, "-I", str(ROOT / "include"), *srcs, "-o", str(LIB) + ".tmp"]
    tl = _torch_lib_dir()
    if tl:
        cmd += [f"-Wl,-rpath,{tl}"]
    cmd += ["-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with status {res.returncode}")
    os.replace(str(LIB) + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))

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


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with status {res.returncode}")


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every kernel source into LIB; raises on a compiler error.

    Each source compiles to its own object in a private temporary folder
    (in parallel, at most ``jobs`` at once), then one link makes the
    library.  Every source is recompiled on each call."""
    import tempfile
    from concurrent.futures import ThreadPoolExecutor

    srcs = sources()
    if jobs is None:
        jobs = min(len(srcs), max(1, min(16, len(os.sched_getaffinity(0)))))
    compile_flags = [f for f in FLAGS if f != "-shared"]
    with tempfile.TemporaryDirectory(prefix="snapy_amd_build_") as tmp:
        objs = [str(Path(tmp) / (s.stem + ".o")) for s in srcs]
        cmds = [[HIPCC, *compile_flags, "-I", str(ROOT / "include"), "-c", str(s), "-o", o]
                for s, o in zip(srcs, objs)]
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            for f in [ex.submit(_run, c, verbose) for c in cmds]:
                f.result()
        link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(LIB) + ".tmp"]
        tl = _torch_lib_dir()
        if tl:
            link += [f"-Wl,-rpath,{tl}"]
        link += ["-Wl,-rpath,/opt/rocm/lib"]
        _run(link, verbose)
    os.replace(str(LIB) + ".tmp", LIB)
    return LIB


HOST_SRC = ROOT / "tests" / "cxx" / "sod_host.cpp"
HOST_BIN = ROOT / "tests" / "cxx" / "sod_host"


def build_host_example(verbose: bool = False) -> Path | None:
    """Compile the C++ host that drives the library through its C-ABI alone
    (no Python, no torch) next to its source; it links the built LIB, found
    at run time through its rpath.  None when the source is absent."""
    if not HOST_SRC.exists():
        return None
    rel = os.path.relpath(LIB.parent, HOST_BIN.parent)
    cmd = [HIPCC, "-O2", "-std=c++17", "-I", str(ROOT / "include"), str(HOST_SRC),
           "-L", str(LIB.parent), f"-l:{LIB.name}", f"-Wl,-rpath,$ORIGIN/{rel}",
           "-Wl,-rpath,/opt/rocm/lib", "-o", str(HOST_BIN) + ".tmp"]
    _run(cmd, verbose)
    os.replace(str(HOST_BIN) + ".tmp", HOST_BIN)
    return HOST_BIN


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_host_example(verbose=True))

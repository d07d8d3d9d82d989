"""Builds libfbp.so (HIP kernels + C ABI + C++ drop-in classes) for gfx950.

Every source is compiled with hipcc directly into ``_build/`` next to this
file and linked into ``libfbp.so`` in the package directory, so the library
travels with the source tree.  ``python -m fasterbackprojection_amd.build``
rebuilds everything; ``build(force=False)`` skips up-to-date objects.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
DROPIN = PKG / "dropin"
BUILD = PKG / "_build"
LIB = PKG / "libfbp.so"
TOOLS = PKG / "tools"
CLI = PKG / "fbp_reconstruct"

ARCH = os.environ.get("FBP_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    # IEEE order of operations everywhere; fused multiply-adds are written out
    # explicitly where they are wanted.
    "-ffp-contract=off",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
    f"-I{DROPIN}",
]
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: libfbp needs ROCm's hipcc to build")


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip")) + sorted(DROPIN.glob("*.cpp"))


def _headers() -> list[Path]:
    return (sorted(CSRC.glob("*.h")) + sorted(DROPIN.glob("*.h"))
            + sorted((ROOT / "include").glob("*.h")))


def _compile(src: Path, force: bool) -> Path:
    obj = BUILD / (src.stem + ("_hip" if src.suffix == ".hip" else "_cpp") + ".o")
    newest_dep = max([src.stat().st_mtime] + [h.stat().st_mtime for h in _headers()])
    if not force and obj.exists() and obj.stat().st_mtime >= newest_dep:
        return obj
    cmd = [hipcc()] + COMMON
    if src.suffix == ".hip":
        cmd += HIP_FLAGS + ["-x", "hip"]
    else:
        cmd += ["-x", "c++", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]
    cmd += ["-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(force: bool = True, jobs: int | None = None) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp)]
           + [str(o) for o in objs]
           + ["-L/opt/rocm/lib", "-lrocfft", "-lz", "-Wl,-rpath,/opt/rocm/lib"])
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    # Command-line driver of the drop-in classes, linked against libfbp.so.
    cmd = ([hipcc(), "-O2", "-std=c++17", f"-I{ROOT / 'include'}", f"-I{DROPIN}",
            "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-x", "c++",
            str(TOOLS / "fbp_reconstruct.cpp"), f"-L{PKG}", "-lfbp",
            "-L/opt/rocm/lib", "-lamdhip64", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN",
            "-Wl,-rpath,/opt/rocm/lib", "-o", str(CLI)])
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"fbp_reconstruct link failed\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return LIB


if __name__ == "__main__":
    force = "--incremental" not in sys.argv
    print(build(force=force))

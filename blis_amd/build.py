"""Build the MI355X BLIS library (libblis_amd.so) in-tree.

Every HIP source under csrc/ is compiled for gfx950 with hipcc and linked into
blis_amd/lib/libblis_amd.so. The build always recompiles (it is the "does it
build" check); objects go to blis_amd/_build/ (ignored by git).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libblis_amd.so"
ROOT = PKG.parent

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BLIS_AMD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-command-line-argument",
    # Keep the compiler from emitting packed FP32 arithmetic.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
    "-DBLIS_IS_BUILDING_LIBRARY",
]

SOURCES = [
    "runtime.cpp",
    "api.cpp",
    "gemm.hip",
    "trsm.hip",
    "level12.hip",
    "batch.hip",
    "split.hip",
    "mlp.hip",
    "select.hip",
    "struc.hip",
    "api_struc.cpp",
    "complex.hip",
    "api_complex.cpp",
    "cblas.cpp",
    "cntx.cpp",
    "sup.hip",
    "crt.hip",
]


def _compile(src: Path, obj: Path) -> tuple[str, int, str]:
    lang = ["-x", "hip"] if src.suffix in (".hip",) else []
    cmd = [HIPCC, *COMMON, *lang, "-c", str(src), "-o", str(obj)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(
        line for line in (p.stdout + p.stderr).splitlines()
        if "packed-fp32-ops" not in line
    )
    return str(src.name), p.returncode, out


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in SOURCES]
    objs = [BUILD / (s.name + ".o") for s in srcs]
    jobs = jobs or min(8, os.cpu_count() or 4)
    failed = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, s, o) for s, o in zip(srcs, objs)]
        for f in futs:
            name, rc, out = f.result()
            if out.strip() and (verbose or rc != 0):
                print(f"[{name}]\n{out}", file=sys.stderr)
            if rc != 0:
                failed.append(name)
    if failed:
        raise RuntimeError(f"hipcc failed for: {', '.join(failed)}")
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *[str(o) for o in objs], "-lpthread"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

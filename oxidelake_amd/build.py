"""Builds liboxide_rocm.so (the gfx950 HIP library behind include/oxide_rocm.h).

Every HIP source compiles to an object with hipcc for gfx950 only, in
parallel, and the objects link into one shared library in this package
directory. Device code is built without packed-FP32 instructions.
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
ROOT = PKG.parent
LIB = PKG / "liboxide_rocm.so"
OBJ_DIR = PKG / "_build"

SOURCES = [
    "runtime.hip",
    "expr.hip",
    "filter.hip",
    "aggregate.hip",
    "join.hip",
    "vector.hip",
    "datagen.hip",
    "jit.hip",
]

# Device headers the generated kernels include; embedded into the library
# as strings (jit_sources.inc) for hipRTC.
JIT_HEADERS = ["divlit.h", "scalar.h", "compact.h", "filter_jit.h", "jtable.h", "aggregate_jit.h"]

# Sources whose float arithmetic must not be contracted into FMA: the
# filter and aggregate evaluate SQL expressions, where a*b+c is two roundings.
NO_CONTRACT = {"expr.hip", "filter.hip", "aggregate.hip", "join.hip"}

# Sources that accumulate FP64 sums with hardware atomics (global and LDS
# add-f64 instead of compare-and-swap loops).
FP_ATOMICS = {"aggregate.hip", "join.hip"}

ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP backend cannot be built")


def common_flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-Wall",
        "-Wno-unused-function",
        "-Wno-unused-variable",
        "-Wno-unused-but-set-variable",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        f"-I{ROOT / 'include'}",
    ]


def write_jit_sources() -> Path:
    """Writes _build/jit_sources.inc: the JIT headers as raw string literals."""
    out = OBJ_DIR / "jit_sources.inc"
    delim = "OXRJIT"
    parts = [
        "// generated by oxidelake_amd/build.py from " + ", ".join(JIT_HEADERS),
        f"static const int kJitHeaderCount = {len(JIT_HEADERS)};",
        "static const char *const kJitHeaderNames[] = {"
        + ", ".join(f'"{h}"' for h in JIT_HEADERS) + "};",
        "static const char *const kJitHeaderText[] = {",
    ]
    for h in JIT_HEADERS:
        text = (CSRC / h).read_text()
        if f"){delim}\"" in text:
            raise RuntimeError(f"{h} contains the raw-string delimiter")
        parts.append(f'R"{delim}({text}){delim}",')
    parts.append("};")
    body = "\n".join(parts) + "\n"
    if not out.exists() or out.read_text() != body:
        out.write_text(body)
    return out


def _compile(src: str) -> Path:
    obj = OBJ_DIR / (src + ".o")
    flags = common_flags() + [f"-I{OBJ_DIR}"]
    if src in NO_CONTRACT:
        flags.append("-ffp-contract=off")
    if src in FP_ATOMICS:
        flags.append("-munsafe-fp-atomics")
    cmd = [hipcc(), *flags, "-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr}")
    warn = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if warn.strip() and os.environ.get("OXR_BUILD_VERBOSE"):
        print(warn, file=sys.stderr)
    return obj


def build(verbose: bool = False) -> Path:
    """Compiles every HIP source and links liboxide_rocm.so; returns its path."""
    OBJ_DIR.mkdir(exist_ok=True)
    write_jit_sources()
    srcs = [s for s in SOURCES if (CSRC / s).exists()]
    workers = min(len(srcs), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
           "-L/opt/rocm/lib", "-lhiprtc"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)

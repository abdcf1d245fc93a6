"""Build the native libraries in-tree.

- ``libglint_gen.so``: the host-side seeded trace generator (gcc).
- ``libglint.so``: the HIP kernels and the C-ABI declared in include/glint.h
  (hipcc, --offload-arch=gfx950).

Every HIP object is compiled with packed-FP32 instruction selection disabled and
floating-point contraction off, so the FP64 distance tests round exactly as the
reference's expressions are written.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"

HIP_SOURCES = ["glint.hip", "walk.hip", "kv.hip", "serve.hip", "trace_io.hip"]
HIP_HEADERS = ["glint_device.h", "kv_engine.h", "ctx_view.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GLINT_OFFLOAD_ARCH", "gfx950")


def source_digest() -> str:
    """First 16 hex digits of the SHA-256 of the kernel sources and headers
    libglint.so is built from: stamps measurements (rocprof counter
    summaries) with the code they measured."""
    import hashlib
    h = hashlib.sha256()
    for f in [CSRC / s for s in HIP_SOURCES + HIP_HEADERS] + sorted(INCLUDE.glob("*.h")):
        h.update(f.name.encode())
        h.update(f.read_bytes())
    return h.hexdigest()[:16]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_gen() -> Path:
    out = PKG / "libglint_gen.so"
    _run(["gcc", "-O3", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off", "-fno-fast-math",
          "-fopenmp", "-Wall", "-Wextra", f"-I{INCLUDE}", "-o", str(out), str(CSRC / "gen.c"), "-lm"])
    return out


def build_hip() -> Path:
    out = PKG / "libglint.so"
    tmp = PKG / "libglint.so.tmp"
    bdir = PKG / "_build"
    bdir.mkdir(exist_ok=True)
    gen_obj = bdir / "gen.o"
    _run(["gcc", "-O3", "-fPIC", "-c", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
          f"-I{INCLUDE}", "-o", str(gen_obj), str(CSRC / "gen.c")])
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wall", "-Wno-unused-function", f"-I{INCLUDE}", f"-I{CSRC}",
           "-o", str(tmp)]
    cmd += [str(CSRC / s) for s in HIP_SOURCES] + ["-x", "none", str(gen_obj), "-fopenmp"]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_host() -> Path:
    """The C++ driver over the C ABI (links libglint.so)."""
    out = PKG / "host" / "glint_pipeline"
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", f"-I{PKG / 'host'}", "-o", str(out),
          str(PKG / "host" / "glint_pipeline.cpp"), f"-L{PKG}", "-lglint", f"-Wl,-rpath,$ORIGIN/.."])
    return out


def build_all() -> None:
    build_gen()
    build_hip()
    build_host()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"gen": build_gen, "hip": build_hip, "host": build_host, "all": build_all}[w]()

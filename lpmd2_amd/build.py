"""In-tree build of liblpmdx.so for gfx950 with hipcc (no CMake, no JIT cache)."""
from __future__ import annotations

import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(_HERE)
# engine_kinds_*.hip instantiate the force-kernel launchers of a few potential
# kinds each, so that the kernel variants compile in parallel
SOURCES = [os.path.join(_HERE, "csrc", f) for f in (
    "batch.hip", "engine_kinds_eam_a.hip", "engine_kinds_pair_a.hip", "engine_kinds_eam_b.hip",
    "engine_kinds_pair_b.hip", "engine_kinds_pair_c.hip", "engine_kinds_pair_d.hip", "engine_kinds_eam_c.hip",
    "engine.hip", "server.cpp", "formats.cpp")]
HEADERS = [os.path.join(_HERE, "csrc", f) for f in ("common.hpp", "sort.hpp", "cells.hpp", "forces.hpp",
                                                     "batch.hpp", "host_util.hpp", "table.hpp",
                                                     "engine_system.hpp", "engine_launch.hpp")] + [
    os.path.join(ROOT, "include", "lpmdx.h")
]
OUT = os.path.join(_HERE, "liblpmdx.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # FP64 geometry must round like the reference (no contraction, no reassociation)
    "-ffp-contract=off",
    "-fvisibility=hidden",
    # no packed FP32 instructions
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def build_library(force: bool = True) -> str:
    """Compiles every source to an object in parallel, then links the shared
    library next to this file (written to a temporary name, then renamed)."""
    if not force and os.path.exists(OUT):
        newest = max(os.path.getmtime(p) for p in SOURCES + HEADERS)
        if os.path.getmtime(OUT) >= newest:
            return OUT
    objdir = os.path.join(_HERE, "_obj")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in FLAGS if f != "-shared"]
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        procs.append((src, subprocess.Popen([HIPCC, *compile_flags, *lang, "-c", "-o", obj, src])))
        objs.append(obj)
    failed = [src for src, p in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, f"hipcc {' '.join(failed)}")
    tmp = OUT + ".tmp"
    subprocess.check_call([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread", "-lz"])
    os.replace(tmp, OUT)
    return OUT

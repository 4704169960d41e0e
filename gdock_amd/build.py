"""Builds libgdock_hip.so (HIP, gfx950) in-tree with hipcc.

Run as ``python -m gdock_amd.build``.  The library holds every kernel and the
C ABI declared in include/gdock_hip.h; nothing is JIT-compiled at run time.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "libgdock_hip.so"
SOURCES = ["api.cpp", "pdb.cpp", "score.hip", "fcc.hip", "eval.hip", "batch.cpp", "lanes.cpp", "ga.cpp", "ga_device.hip", "serve.cpp", "output.cpp"]

# -ffp-contract=off: FP64 sums and products round as in the reference (no
# implicit FMA); the FP32 pair loop spells its FMAs out.  Packed FP32
# instruction selection is disabled for the device code.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if c and Path(c).exists():
            return c
    return "hipcc"


def sources() -> list[str]:
    return [str(PKG / "csrc" / s) for s in SOURCES]


def _run(cmd: list[str]) -> subprocess.CompletedProcess:
    return subprocess.run(cmd, capture_output=True, text=True)


def build(verbose: bool = True, lib: Path = LIB, objdir: Path | None = None, extra_flags=()) -> Path:
    """One object per source (compiled in parallel into gdock_amd/_build, or
    objdir), then one link into lib.  extra_flags go to every compile and to
    the link (the host-sanitizer test builds a copy with them)."""
    from concurrent.futures import ThreadPoolExecutor
    out = objdir or PKG / "_build"
    out.mkdir(parents=True, exist_ok=True)
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"] + list(extra_flags)
    jobs = []
    for s in sources():
        obj = out / (Path(s).name + ".o")
        jobs.append([hipcc(), *compile_flags, "-c", *(["-x", "hip", s] if s.endswith(".cpp") else [s]),
                     "-o", str(obj)])
    if verbose:
        print(f"hipcc {' '.join(compile_flags)} -c <{len(jobs)} sources> -> {out}", flush=True)
    with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 4, 16)) as pool:
        results = list(pool.map(_run, jobs))
    notes = []
    for cmd, res in zip(jobs, results):
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"hipcc failed compiling {cmd[-3]}")
        notes += [l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    tmp = lib.with_suffix(".so.tmp")
    link = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *extra_flags, *[j[-1] for j in jobs], "-o", str(tmp)]
    res = _run(link)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libgdock_hip.so")
    if notes and verbose:
        sys.stderr.write("\n".join(notes) + "\n")
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    build()

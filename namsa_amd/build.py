"""Build libnamsa_hip.so in-tree with hipcc for gfx950.

    python -m namsa_amd.build          # compile every HIP source, link the .so

The library lands next to this file (namsa_amd/libnamsa_hip.so) so it ships
with the source tree.  Flags: -O3, gfx950 only, packed FP32 ops disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops), position-independent.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
OBJDIR = PKG / "_build"
LIB = PKG / "libnamsa_hip.so"
SOURCES = ["abi_common.hip", "multislice.hip", "slices.hip", "passes.hip", "gemm_mfma.hip", "fno_ops.hip",
           "fno_fused.hip", "records.hip"]
PASS_SIZES = [128, 192, 256, 384, 512, 768, 1024, 2048]
ARCH = os.environ.get("NAMSA_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the HIP toolchain is required to build namsa_amd")


def common_flags() -> list[str]:
    return [
        "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-result", "-Wno-unknown-pragmas",
        f"-I{INCLUDE}", f"-I{CSRC}",
    ]


def _compile(job) -> Path:
    src, defines = job
    tag = "".join("_" + d.split("=")[-1] for d in defines)
    obj = OBJDIR / (Path(src).stem + tag + ".o")
    cmd = [hipcc(), *common_flags(), *(f"-D{d}" for d in defines), "-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    """Compile all sources (always) and link the shared library."""
    OBJDIR.mkdir(exist_ok=True)
    jobs = [(s, []) for s in SOURCES] + [("passes_n.hip", [f"NAMSA_PASS_N={n}"]) for n in PASS_SIZES]
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, jobs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)

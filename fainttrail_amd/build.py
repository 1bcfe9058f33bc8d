"""Build the gfx950 backend library (libftk_hip.so) and the CPU oracle in-tree.

hipcc compiles for gfx950 only, with -ffp-contract=off (no fused multiply-add, the
reference's choice for bit-identical results, CMakeLists.txt:56-60 and 86-91),
denormals kept, packed-FP32 instruction selection disabled, and wave-uniform branches
left unstructured.
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
LIB = PKG / "libftk_hip.so"
OBJ = PKG / "_build"

HIP_SOURCES = ["stack_inst_f16.hip", "stack_inst_bf16.hip", "stack_inst_f32.hip", "stack_kernel.hip",
               "probe_kernel.hip", "prep_kernels.hip", "search_kernels.hip", "search_device.hip", "ftk_api.hip",
               "ftk_prep_api.hip", "ftk_search_api.hip", "store_kernels.hip", "serve.hip", "screen_kernels.hip",
               "refine_kernels.hip", "fuse.hip", "stack_plan.cpp", "candidates.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # branches on wave-uniform values stay plain scalar branches (the lattice stacking
    # loop's per-row-class switch): structurizing them made the compiler copy the
    # accumulators at every join
    "-mllvm", "-structurizecfg-skip-uniform-regions=true",
    "-Wall",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the gfx950 backend)")


def _run(cmd: list[str], cwd: Path) -> None:
    res = subprocess.run(cmd, cwd=str(cwd), capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd[:3])} ... (exit {res.returncode})")


def build_backend(jobs: int | None = None) -> Path:
    """Compiles every source to an object of its own (in parallel), then links the .so."""
    OBJ.mkdir(exist_ok=True)
    inc = ["-I", str(ROOT / "include"), "-I", str(CSRC)]

    def compile_one(s: str) -> Path:
        p = CSRC / s
        o = OBJ / (p.name + ".o")
        # .cpp files are host code; hipcc compiles them as HIP too (-x hip) for one .so
        src = ["-x", "hip", str(p)] if p.suffix == ".cpp" else [str(p)]
        _run([hipcc(), *HIPCC_FLAGS, *inc, "-c", *src, "-o", str(o)], CSRC)
        return o

    jobs = jobs or max(1, min(len(HIP_SOURCES), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)], CSRC)
    os.replace(tmp, LIB)
    return LIB


def build_oracle(with_reference: bool = True) -> None:
    _run(["make", "-s", "-C", str(ROOT / "oracle"), "-j8"], ROOT)
    if with_reference and Path("/root/reference/src").is_dir():
        _run(["make", "-s", "-C", str(ROOT / "oracle"), "-j8", "ref", "cli", "cli-hip"], ROOT)


def build_all() -> None:
    build_backend()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", LIB)

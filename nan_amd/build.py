"""Builds ``nan_amd/libnan_hip.so`` for gfx950 with hipcc (in-tree).

The host BVH builder is compiled with g++, the kernels and the C ABI with
hipcc; both with ``-ffp-contract=off`` (the arithmetic contract of the
kernels), and hipcc with packed-FP32 instructions disabled.
"""

from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libnan_hip.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NAN_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wno-unused-result",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + proc.stdout)


def build(verbose: bool = False) -> Path:
    build_dir = PKG / "_build"
    build_dir.mkdir(exist_ok=True)
    bvh_obj = build_dir / "nan_bvh.o"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fvisibility=hidden", "-c",
          str(CSRC / "nan_bvh.cpp"), "-o", str(bvh_obj)])
    sources = ("nan_api.hip", "nan_wavefront.hip", "nan_nrc.hip", "nan_nrc_fused.hip", "nan_nrc_noalloc.hip", "nan_learned.hip",
               "nan_filter.hip", "nan_atmosphere.hip")
    objs = [str(build_dir / (Path(src).stem + ".o")) for src in sources]
    cmds = [[HIPCC, *HIP_FLAGS, "-I", str(ROOT / "include"), "-c", str(CSRC / src), "-o", obj]
            for src, obj in zip(sources, objs)]
    with ThreadPoolExecutor(max_workers=len(cmds)) as pool:
        list(pool.map(_run, cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, str(bvh_obj), "-o", str(tmp)])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)

"""Build libssmhip.so for gfx950 with hipcc (no cmake/ninja involved).

Usage: python ssm-simulators_amd/csrc/build.py [--jobs N]
Objects go to ssm-simulators_amd/csrc/_build/, the library to
ssm-simulators_amd/ssmhip/_native/libssmhip.so.
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

CSRC = Path(__file__).resolve().parent
PKG = CSRC.parent
REPO = PKG.parent
OUT_DIR = PKG / "ssmhip" / "_native"
OBJ_DIR = CSRC / "_build"
LIB_NAME = "libssmhip.so"

# Sources and the extra flags each one needs.
SOURCES = {
    "capi.hip": [],
    "em_kernel.hip": ["-ffp-contract=off"],
    "lba_kernel.hip": ["-ffp-contract=off"],
    "kde_kernel.hip": [],
    "aux_kernel.hip": [],
    "filter_kernel.hip": [],
    # MFMA results in VGPRs (no accumulator-register copies before tanh)
    "mlp_kernel.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
    "probe_kernel.hip": [],
}

ARCH = os.environ.get("SSMHIP_ARCH", "gfx950")
COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
    # packed FP32 instructions stay off (plain scalar FP32 VALU only)
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    f"-I{REPO / 'include'}",
    f"-I{CSRC}",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found; a ROCm toolchain is required to build libssmhip")
    return path


def _compile(src: str, extra: list[str]) -> Path:
    obj = OBJ_DIR / (Path(src).stem + ".o")
    cmd = [hipcc(), *COMMON, *extra, "-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(jobs: int = 4, verbose: bool = True) -> Path:
    OBJ_DIR.mkdir(parents=True, exist_ok=True)
    OUT_DIR.mkdir(parents=True, exist_ok=True)
    srcs = {s: f for s, f in SOURCES.items() if (CSRC / s).exists()}
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        futs = {ex.submit(_compile, s, f): s for s, f in srcs.items()}
        objs = []
        for fut in cf.as_completed(futs):
            objs.append(fut.result())
            if verbose:
                print(f"[ssmhip build] compiled {futs[fut]}", flush=True)
    lib = OUT_DIR / LIB_NAME
    tmp = OUT_DIR / (LIB_NAME + ".tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, sorted(objs))]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, lib)
    if verbose:
        print(f"[ssmhip build] wrote {lib}", flush=True)
    return lib


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    args = ap.parse_args()
    build(args.jobs)
    return 0


if __name__ == "__main__":
    sys.exit(main())

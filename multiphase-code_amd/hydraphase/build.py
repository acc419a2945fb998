"""Build libhydra_hip.so in-tree with hipcc for gfx950.

Flags: FP64 results must be bit-identical to the reference's serial CPU
evaluation, so multiply-adds are never contracted (-ffp-contract=off); the
packed-FP32 instruction forms are disabled for every HIP translation unit.
"""
from __future__ import annotations

import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
LIB = os.path.join(PKG, "libhydra_hip.so")
SOURCES = ["hydra_kernels.hip", "hydra_capi.cpp", "hydra_noalloc.hip", "hydra_stream.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs = []
    for src in SOURCES:
        obj = os.path.join(PKG, "_build", src + ".o")
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        cmd = [hipcc, *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-x", "hip", "-c",
               os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
        if verbose and r.stderr:
            print(r.stderr)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


FORTRAN = os.path.join(ROOT, "fortran")
DRIVER = os.path.join(PKG, "hydra_driver")


def build_driver() -> str:
    """The Fortran drop-in driver, linked against libhydra_hip.so."""
    fc = os.environ.get("FC", "/opt/rocm/bin/amdflang")
    mod_dir = os.path.join(PKG, "_build", "fortran")
    os.makedirs(mod_dir, exist_ok=True)
    cmd = [fc, "-O2", "-cpp", f"-J{mod_dir}", os.path.join(FORTRAN, "hydra_hip_mod.f90"),
           os.path.join(FORTRAN, "hydra_driver.F90"), f"-L{PKG}", "-lhydra_hip",
           "-Wl,-rpath,$ORIGIN", "-o", DRIVER + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=mod_dir)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed:\n{r.stderr}")
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_driver())

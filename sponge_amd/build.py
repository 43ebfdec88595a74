"""Builds libsponge_nb.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m sponge_amd.build   (or __graft_entry__.build()).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsponge_nb.so")
DRIVER = os.path.join(HERE, "sponge_nb_driver")
FE_DRIVER = os.path.join(HERE, "sponge_fe_driver")
HOST_SOURCES = ["host/sponge_host.cpp", "host/nb_driver.cpp"]
FE_HOST_SOURCES = ["host/sponge_host.cpp", "host/fe_driver.cpp"]
SOURCES = ["abi.hip", "nl.hip", "pair.hip", "pme.hip", "batch.hip", "adjoint.hip", "md.hip", "fep.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
    # packed FP32 (v_pk_*_f32) stays off
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-warning-option",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(HERE, "..", "include"), "-c",
           os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and "unknown target feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{out}")
    if out.strip():
        print(out, file=sys.stderr)
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    _build_driver()
    if verbose:
        print(f"built {LIB}, {DRIVER} and {FE_DRIVER}")
    return LIB


def _build_driver() -> None:
    """Host-only C++ (the SPONGE class mirror, the step-loop driver and the
    FEP / TI frame driver), linked against libsponge_nb.so through an $ORIGIN
    rpath."""
    for exe, sources in ((DRIVER, HOST_SOURCES), (FE_DRIVER, FE_HOST_SOURCES)):
        tmp = exe + ".tmp"
        cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(HERE, "..", "include"),
               *[os.path.join(CSRC, s) for s in sources], "-o", tmp,
               "-L", HERE, "-lsponge_nb", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {os.path.basename(exe)}:\n{r.stdout}{r.stderr}")
        os.replace(tmp, exe)


if __name__ == "__main__":
    build()

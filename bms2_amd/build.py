"""Builds libbms2.so (HIP kernels for gfx950 + host C ABI) and the CPU oracle.

Run as ``python -m bms2_amd.build`` or through ``__graft_entry__.build()``.  Objects go to
``bms2_amd/_build``; the shared library lands next to this file so it travels with the tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libbms2.so"
ORACLE = ROOT / "oracle"

ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")
ARCH = "gfx950"

CXXFLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]
# Packed FP32 instruction forms are disabled for all device code.
HIPFLAGS = CXXFLAGS + [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}):\n{' '.join(cmd)}\n{proc.stdout}")
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if out.strip():
        print(out, file=sys.stderr)


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *HIPFLAGS, "-c", str(src), "-o", str(obj)]
    else:
        cmd = ["g++", *CXXFLAGS, "-I", str(ROCM / "include"), "-D__HIP_PLATFORM_AMD__", "-c",
               str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build_library(jobs: int | None = None) -> Path:
    BUILD.mkdir(exist_ok=True)
    sources = sorted([*CSRC.glob("*.cpp"), *CSRC.glob("*.hip")])
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_driver() -> Path:
    """bms2_main and bms2_batch: the reference's BMS2 and BMS2_batch drivers written against
    include/sigmo.hpp."""
    for name in ("bms2_main", "bms2_batch"):
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-pthread", "-I", str(ROOT / "include"),
              str(PKG / "examples" / f"{name}.cpp"), "-o", str(PKG / name), "-L", str(PKG),
              "-lbms2", "-Wl,-rpath,$ORIGIN"])
    return PKG / "bms2_main"


def build_dropin_check() -> Path:
    """tests/dropin/bms2_callshapes: a driver making the reference BMS2 driver's library calls in
    the same shapes against include/sigmo.hpp (the drop-in check the tests compile and run)."""
    src = ROOT / "tests" / "dropin" / "bms2_callshapes.cpp"
    out = src.with_suffix("")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-pthread", "-I", str(ROOT / "include"),
          str(src), "-o", str(out), "-L", str(PKG), "-lbms2", f"-Wl,-rpath,{PKG}",
          "-Wl,-rpath,$ORIGIN/../../bms2_amd"])
    return out


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ORACLE), "-B", "liboracle.so"])
    return ORACLE / "liboracle.so"


def build_all() -> None:
    build_library()
    build_driver()
    build_dropin_check()
    build_oracle()


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build_all()
    print(f"built {LIB}")

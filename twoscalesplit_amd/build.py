"""Build recipe of libtss.so (HIP, gfx950) and its C++ facade example.

Run as ``python -m twoscalesplit_amd.build`` or through
``__graft_entry__.build()``. Everything is built in-tree:

* ``twoscalesplit_amd/lib/libtss.so``  -- the product (hipcc, gfx950)
* ``twoscalesplit_amd/lib/facade_demo`` -- a program written against the
  reference's class API (WCSPHConfig / WCSPHSolver / ParticleSystem)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
LIB_DIR = os.path.join(ROOT, "lib")
LIB = os.path.join(LIB_DIR, "libtss.so")
DEMO = os.path.join(LIB_DIR, "facade_demo")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: index arithmetic rounds every operation (the method's
# bit-exact definition); packed FP32 instructions are disabled for all
# device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Wall",
    "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

SOURCES = ["tss_kernels.hip", "tss_host.cpp"]


def _run(cmd: list[str], cwd: str | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE,
                          stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): "
                           f"{' '.join(cmd)}\n{out}")


def build_lib() -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(LIB_DIR, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC, *HIP_FLAGS, *lang, "-c", os.path.join(CSRC, src),
              "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    for obj in objs:
        os.remove(obj)
    return LIB


def build_demo() -> str:
    src = os.path.join(ROOT, "host", "facade_demo.cpp")
    cxx = shutil.which("g++") or "g++"
    _run([cxx, "-O2", "-std=c++17", "-Wall", src, "-o", DEMO,
          f"-L{LIB_DIR}", "-ltss", "-Wl,-rpath,$ORIGIN"])
    return DEMO


def build_all() -> None:
    build_lib()
    build_demo()


if __name__ == "__main__":
    build_all()
    print(LIB)
    sys.exit(0)

"""Builds the gfx950 HIP library (libtagore.so) and the pybind11 module (Tagore*.so) in place.

    python -m tagore_amd.build        # or tagore_amd.build.build()

Every .hip file under csrc/ is compiled with hipcc for gfx950 and linked into
tagore_amd/libtagore.so; csrc/pymodule.cpp is compiled with g++ against include/tagore.h and
linked to that library with an $ORIGIN rpath, giving tagore_amd/Tagore<EXT_SUFFIX>.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtagore.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # decisions are bit-exact: no contraction, IEEE division and square root
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    # never emit packed FP32 arithmetic
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n  " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _hip_sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def build_lib(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for f in _hip_sources():
        obj = os.path.join(BUILD, f[:-4] + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, f), "-o", obj])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def module_path() -> str:
    return os.path.join(HERE, "Tagore" + sysconfig.get_config_var("EXT_SUFFIX"))


def build_module() -> str:
    import pybind11

    out = module_path()
    tmp = out + ".tmp"
    _run([
        "g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall",
        "-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"],
        os.path.join(CSRC, "pymodule.cpp"), "-o", tmp,
        "-L" + HERE, "-ltagore", "-Wl,-rpath,$ORIGIN",
    ])
    os.replace(tmp, out)
    return out


def build() -> None:
    build_lib()
    build_module()


if __name__ == "__main__":
    build()
    print("built", LIB, "and", module_path())
    sys.exit(0)

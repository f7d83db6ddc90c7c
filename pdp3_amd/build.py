"""Build libpdp3_hip.so (gfx950), libpdp3_host.so and the pdp3 driver in-tree.

    python -m pdp3_amd.build            # compile + link
    python -m pdp3_amd.build --clean

Device code is compiled with hipcc for gfx950 only; the host side
(pdp3_amd/host: classes, parameters.xml reader, loaders, probes, the pdp3
driver) with g++ into libpdp3_host.so and bin/pdp3.  The files
whose results must round exactly like the reference's IEEE build (the field
updates and the FP64 particle kernels) are compiled with -ffp-contract=off;
the FP32 form keeps contraction (FMA) on.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libpdp3_hip.so")
HOST = os.path.join(HERE, "host")
BINDIR = os.path.join(HERE, "bin")
HOST_LIB = os.path.join(LIBDIR, "libpdp3_host.so")
DRIVER = os.path.join(BINDIR, "pdp3")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"
# host C++ (no device code): IEEE arithmetic without contraction, as the
# reference's loaders are compiled
HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra", "-Wno-unused-parameter",
              "-I", os.path.join(ROOT, "include")]
HOST_UNITS = ["xml.cpp", "parameters.cpp", "loaders.cpp", "pdp3.cpp", "probe_plain.cpp", "load_init_param.cpp",
              "host_capi.cpp"]

COMMON = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"),
]
# (source, extra flags)
UNITS = [
    ("capi.cpp", ["-ffp-contract=off", "-x", "hip"]),
    ("fields.hip", ["-ffp-contract=off"]),
    ("particles_fp64.hip", ["-ffp-contract=off"]),
    ("init.hip", ["-ffp-contract=off"]),
    ("particles_fp32.hip", ["-ffp-contract=fast", "-fno-hip-fp32-correctly-rounded-divide-sqrt"]),
    ("sort.hip", []),
    ("adjoint.hip", ["-ffp-contract=off"]),
    ("probes.hip", ["-ffp-contract=off"]),
]


def _compile(src: str, extra: list[str]) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def _run(cmd: list[str], what: str) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def _compile_host(src: str) -> str:
    obj = os.path.join(BUILD, "host_" + os.path.splitext(src)[0] + ".o")
    _run([CXX, *HOST_FLAGS, "-c", os.path.join(HOST, src), "-o", obj], "compile")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BINDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        dev = [ex.submit(_compile, *u) for u in UNITS]
        host = [ex.submit(_compile_host, s) for s in HOST_UNITS]
        objs = [f.result() for f in dev]
        hobjs = [f.result() for f in host]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], "link")
    os.replace(tmp, LIB)
    # host side: the classes, XML reader, loaders, probes and the driver,
    # linked against the device library found next to it
    tmp = HOST_LIB + ".tmp"
    _run([CXX, "-shared", "-fPIC", "-o", tmp, *hobjs, "-L", LIBDIR, "-lpdp3_hip", "-Wl,-rpath,$ORIGIN"], "link")
    os.replace(tmp, HOST_LIB)
    main_obj = _compile_host("main.cpp")
    tmp = DRIVER + ".tmp"
    _run([CXX, "-o", tmp, main_obj, "-L", LIBDIR, "-lpdp3_host", "-lpdp3_hip", "-Wl,-rpath,$ORIGIN/../lib"], "link")
    os.replace(tmp, DRIVER)
    if verbose:
        print(f"built {LIB}\nbuilt {HOST_LIB}\nbuilt {DRIVER}")
    return LIB


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)
    for f in (LIB, HOST_LIB, DRIVER):
        if os.path.exists(f):
            os.remove(f)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--clean", action="store_true")
    a = ap.parse_args()
    if a.clean:
        clean()
        sys.exit(0)
    build(verbose=True)

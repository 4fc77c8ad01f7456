"""In-tree build of the native library and the bench drivers.

Produces, next to this file:
  libdmf.so                 host formats/tools + HIP engine (C-ABI of include/dmf.h)
  bin/dyn_{data,pp,topo}_bench{,_large}   drivers with the reference's argv/stdout
  bin/dmf_input             seeded input tool (graph / update .bin writer)
  bin/dmf_multi_bench       many independent instances solved as one batched engine
  bin/dmf_serve_bench       streaming update batches against a resident graph (server)

Everything is compiled for gfx950 with hipcc; packed-fp32 instruction
selection is disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
BIN = os.path.join(HERE, "bin")
LIB = os.path.join(HERE, "libdmf.so")

ARCH = os.environ.get("DMF_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wno-unused-result",
    "-Wno-unused-value",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wextra"]

BENCH_VARIANTS = {
    "dyn_data_bench": ("data", 0),
    "dyn_data_bench_large": ("data", 1),
    "dyn_topo_bench": ("topo", 0),
    "dyn_topo_bench_large": ("topo", 1),
    "dyn_pp_bench": ("pp", 0),
    "dyn_pp_bench_large": ("pp", 1),
}


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    objs = []
    host_src = os.path.join(CSRC, "host_io.cpp")
    host_obj = os.path.join(BUILD, "host_io.o")
    _run(["g++", *HOST_FLAGS, "-I", INCLUDE, "-c", host_src, "-o", host_obj])
    objs.append(host_obj)
    srv_src = os.path.join(CSRC, "server.cpp")
    if os.path.exists(srv_src):
        srv_obj = os.path.join(BUILD, "server.o")
        _run(["g++", *HOST_FLAGS, "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", "-I", INCLUDE,
              "-c", srv_src, "-o", srv_obj])
        objs.append(srv_obj)
    eng_src = os.path.join(CSRC, "engine.hip")
    eng_obj = os.path.join(BUILD, "engine.o")
    _run([HIPCC, *DEVICE_FLAGS, "-I", INCLUDE, "-c", eng_src, "-o", eng_obj])
    objs.append(eng_obj)
    for extra in ("sched.hip", "multi.hip"):
        p = os.path.join(CSRC, extra)
        if os.path.exists(p):
            o = os.path.join(BUILD, extra.replace(".hip", ".o"))
            _run([HIPCC, *DEVICE_FLAGS, "-I", INCLUDE, "-c", p, "-o", o])
            objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    # drivers (host C++ linking libdmf.so through its C-ABI only)
    drv = os.path.join(CSRC, "bench_main.cpp")
    if os.path.exists(drv):
        for name, (variant, large) in BENCH_VARIANTS.items():
            _run(["g++", *HOST_FLAGS, "-I", INCLUDE, f"-DDMF_BENCH_VARIANT=\"{variant}\"",
                  f"-DDMF_BENCH_LARGE={large}", drv, "-o", os.path.join(BIN, name),
                  "-L", HERE, "-ldmf", f"-Wl,-rpath,$ORIGIN/.."])
    for src_name, exe in (("multi_bench.cpp", "dmf_multi_bench"), ("serve_bench.cpp", "dmf_serve_bench")):
        src = os.path.join(CSRC, src_name)
        if os.path.exists(src):
            _run(["g++", *HOST_FLAGS, "-I", INCLUDE, src, "-o", os.path.join(BIN, exe),
                  "-L", HERE, "-ldmf", "-pthread", f"-Wl,-rpath,$ORIGIN/.."])
    tool = os.path.join(CSRC, "input_tool.cpp")
    if os.path.exists(tool):
        _run(["g++", *HOST_FLAGS, "-I", INCLUDE, tool, "-o", os.path.join(BIN, "dmf_input"),
              "-L", HERE, "-ldmf", f"-Wl,-rpath,$ORIGIN/.."])
    if verbose:
        print(f"built {LIB}")
    return LIB


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)

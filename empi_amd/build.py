"""In-tree build of the HIP engine (gfx950) and the C++ host tools.

    python -m empi_amd.build          # libempi_hip.so + empi-hip CLI
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 (v_pk_*_f32) stays off on this part
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "csrc"),
]

# the column kernels are split by class range so that the translation units
# compile in parallel
ENGINE_SOURCES = ["csrc/columns_large.hip", "csrc/columns_mid.hip", "csrc/columns_small.hip", "csrc/engine.hip",
                  "csrc/mlp.hip", "csrc/dictionary.cpp"]
OBJ_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libempi_hip.so")

HOST_SOURCES = ["host/empi_main.cpp", "host/workspace.cpp", "host/config.cpp", "host/book.cpp",
                "host/signal_reader.cpp"]
CLI = os.path.join(HERE, "empi-hip")
TOOL_SOURCES = ["host/host_tool.cpp", "host/config.cpp", "host/book.cpp", "host/signal_reader.cpp"]
TOOL = os.path.join(HERE, "empi-host-tool")


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_engine(jobs=None):
    """Compiles every engine source to an object (in parallel) and links
    libempi_hip.so."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = [os.path.join(OBJ_DIR, os.path.basename(s) + ".o") for s in ENGINE_SOURCES]
    cmds = [[HIPCC] + HIP_FLAGS + ["-c", "-o", o, os.path.join(HERE, s)] for s, o in zip(ENGINE_SOURCES, objs)]
    jobs = jobs or min(len(cmds), max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        for f in [pool.submit(_run, c) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-shared", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_host():
    srcs = [os.path.join(HERE, s) for s in HOST_SOURCES if os.path.exists(os.path.join(HERE, s))]
    if len(srcs) != len(HOST_SOURCES):
        return None
    tmp = CLI + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-pthread", "-I" + os.path.join(ROOT, "include"),
          "-I" + os.path.join(HERE, "host"), "-o", tmp] + srcs +
         ["-L" + HERE, "-lempi_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, CLI)
    return CLI


def build_tool():
    tmp = TOOL + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-I" + os.path.join(ROOT, "include"),
          "-I" + os.path.join(HERE, "host"), "-o", tmp] + [os.path.join(HERE, s) for s in TOOL_SOURCES])
    os.replace(tmp, TOOL)
    return TOOL


def build_all():
    build_engine()
    build_host()
    build_tool()


if __name__ == "__main__":
    build_all()
    sys.exit(0)

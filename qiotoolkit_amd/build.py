"""Build the native libraries in-tree.

libqmc.so   -- HIP kernels for gfx950 + the thin C ABI (include/qmc.h)
libqiotk.so -- C++ host solvers / JSON runner over libqmc (include/qiotk.h)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    # packed FP32 instructions are kept out of generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]

DEVICE_SOURCES = ["qmc_kernels.hip", "qmc_api.cpp"]
HOST_SOURCES = ["host/json.cc", "host/model.cc", "host/schedule.cc", "host/solver.cc", "host/pa.cc",
                "host/capi.cc", "host/server.cc", "host/proto.cc", "host/observer.cc",
                "host/param_free.cc", "host/model_stream.cc"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_device():
    out = os.path.join(HERE, "libqmc.so")
    objs = []
    for src in DEVICE_SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"]
        _run([HIPCC] + HIP_FLAGS + lang + ["-I", os.path.join(ROOT, "include"), "-c",
                                           os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out] + objs)
    return out


def _commit():
    try:
        r = subprocess.run(["git", "-C", ROOT, "rev-parse", "--short", "HEAD"], capture_output=True, text=True,
                           timeout=10)
        return r.stdout.strip() or "unknown"
    except (OSError, subprocess.SubprocessError):
        return "unknown"


def build_host():
    if not HOST_SOURCES:
        return None
    out = os.path.join(HERE, "libqiotk.so")
    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES]
    # -fno-gnu-unique: unique symbols would make the library impossible to
    # unload (the Python binding unloads it at exit, before the C runtime's
    # exit handlers)
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-pthread", "-fno-gnu-unique",
          '-DQIOTK_COMMIT="%s"' % _commit(),
          "-I", os.path.join(ROOT, "include"),
          "-o", out] + srcs + [os.path.join(HERE, "libqmc.so"), "-Wl,-rpath,$ORIGIN"])
    return out


def build_all():
    build_device()
    build_host()


if __name__ == "__main__":
    build_all()
    sys.exit(0)

"""Build recipe of the native libraries (hipcc for gfx950, in-tree outputs).

  python -m puresoft3d_amd.build        # libpsgpu.so, libpuresoft3d.so, host/psrender, host/psdemo
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Exact FP32: no contraction, no fast math, IEEE division and sqrt (HIP's
# default), denormals kept; packed FP32 instructions disabled.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
             "-fno-gpu-flush-denormals-to-zero", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-value", "-Wno-unused-result", "-fPIC"]

SOURCES = ["ps_kernels.hip", "psgpu_api.hip", "ps_adjoint.hip"]
BUILD_DIR = os.path.join(HERE, "_build")


def run(cmd, cwd=None):
    print("+ " + " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s" % (r.returncode, " ".join(cmd)))


def build_psgpu():
    """libpsgpu.so; it links ROCm's HIP runtime by soname (libamdhip64.so.7),
    so in a process that imported torch first it binds to the runtime torch
    already loaded, and the two share device memory and streams."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        o = os.path.join(BUILD_DIR, src.replace(".hip", ".o"))
        run([HIPCC] + HIP_FLAGS + ["-c", "-o", o, os.path.join(CSRC, src)])
        objs.append(o)
    out = os.path.join(HERE, "libpsgpu.so")
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp"] + objs)
    os.replace(out + ".tmp", out)
    return out


HOST = os.path.join(HERE, "host")
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]


def build_facade():
    """C++ host library over libpsgpu (the PuresoftPipeline facade and the
    serving loop), the psrender program that drives the facade, and psdemo,
    the reference demos' resource idioms through it."""
    lib = os.path.join(HERE, "libpuresoft3d.so")
    run(["g++"] + CXX_FLAGS + ["-shared", "-o", lib + ".tmp", os.path.join(HOST, "puresoft3d.cpp"),
                               os.path.join(HOST, "psserve.cpp"), "-L" + HERE, "-lpsgpu", "-Wl,-rpath,$ORIGIN"])
    os.replace(lib + ".tmp", lib)
    exes = []
    for prog in ("psrender", "psdemo"):
        exe = os.path.join(HOST, prog)
        run(["g++"] + CXX_FLAGS + ["-o", exe + ".tmp", os.path.join(HOST, prog + ".cpp"),
                                   "-L" + HERE, "-lpuresoft3d", "-lpsgpu", "-Wl,-rpath,$ORIGIN/.."])
        os.replace(exe + ".tmp", exe)
        exes.append(exe)
    return lib, exes


def main():
    build_psgpu()
    build_facade()


if __name__ == "__main__":
    sys.exit(main())

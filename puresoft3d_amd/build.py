"""Build recipe of the native libraries (hipcc for gfx950, in-tree outputs).

  python -m puresoft3d_amd.build        # libpsgpu.so, libpuresoft3d.so, host/psrender
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

SOURCES = ["ps_kernels.hip", "psgpu_api.hip"]


def run(cmd, cwd=None):
    print("+ " + " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s" % (r.returncode, " ".join(cmd)))


def build_psgpu():
    out = os.path.join(HERE, "libpsgpu.so")
    tmp = out + ".tmp"
    run([HIPCC] + HIP_FLAGS + ["-shared", "-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES])
    os.replace(tmp, out)
    return out


HOST = os.path.join(HERE, "host")
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]


def build_facade():
    """C++ PuresoftPipeline facade (host code only) over libpsgpu, and the
    psrender program that drives it."""
    lib = os.path.join(HERE, "libpuresoft3d.so")
    run(["g++"] + CXX_FLAGS + ["-shared", "-o", lib + ".tmp", os.path.join(HOST, "puresoft3d.cpp"),
                               "-L" + HERE, "-lpsgpu", "-Wl,-rpath,$ORIGIN"])
    os.replace(lib + ".tmp", lib)
    exe = os.path.join(HOST, "psrender")
    run(["g++"] + CXX_FLAGS + ["-o", exe + ".tmp", os.path.join(HOST, "psrender.cpp"),
                               "-L" + HERE, "-lpuresoft3d", "-lpsgpu", "-Wl,-rpath,$ORIGIN/.."])
    os.replace(exe + ".tmp", exe)
    return lib, exe


def main():
    build_psgpu()
    build_facade()


if __name__ == "__main__":
    sys.exit(main())

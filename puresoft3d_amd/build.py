"""Build recipe of the native libraries (hipcc for gfx950, in-tree outputs).

  python -m puresoft3d_amd.build        # builds libpsgpu.so
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


def main():
    build_psgpu()


if __name__ == "__main__":
    sys.exit(main())

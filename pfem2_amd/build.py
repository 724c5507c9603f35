"""Build recipes for the HIP library (gfx950) and the CPU oracle.

`python -m pfem2_amd.build` builds both in-tree:
  pfem2_amd/libpfem2_amd.so   the particle stage (hand-written HIP, C ABI)
  oracle/build/libpf2_oracle.so   the CPU restatement used by tests only
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "pfem2_amd", "libpfem2_amd.so")
ORACLE_LIB = os.path.join(ROOT, "oracle", "build", "libpf2_oracle.so")

HIP_SOURCES = [os.path.join(ROOT, "pfem2_amd", "csrc", "pf2_lib.hip"),
               os.path.join(ROOT, "pfem2_amd", "csrc", "pf2_server.hip")]
HIP_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    # IEEE evaluation order: no contraction into FMA, so FP64 results
    # match a host computation of the same expressions bit for bit
    "-ffp-contract=off", "-Wno-unused-result", "-Wno-unused-value",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_lib(hipcc="/opt/rocm/bin/hipcc"):
    tmp = LIB + ".tmp"
    _run([hipcc] + HIP_FLAGS + ["-o", tmp] + HIP_SOURCES)
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-s"])
    return ORACLE_LIB


def main():
    what = sys.argv[1:] or ["lib", "oracle"]
    if "oracle" in what:
        build_oracle()
    if "lib" in what:
        build_lib()


if __name__ == "__main__":
    main()

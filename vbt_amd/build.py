"""Build the in-tree native libraries.

    python -m vbt_amd.build          # or __graft_entry__.build()

  vbt_amd/libvbt_hip.so    HIP decode library (gfx950), C-ABI in include/vbt_hip.h
  vbt_amd/libvbt_synth.so  synthetic fields + asset encoder (host C++, OpenMP)
  vbt_amd/libvbt_hip_cxx.so  the reference's C++ decode API over the C-ABI (host C++)
  vbt_amd/libvbt_loadgen.so  open-loop load generator over libvbt_hip.so (host C++)
  tests/native/shim_check  consumer of the C++ drop-in headers (test program)
  oracle/liboracle.so      CPU oracle (test infrastructure)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

HIP_SOURCES = ["format.hip", "decode.hip", "server.hip", "batch.hip", "march.hip", "cycles.hip", "tuning.hip"]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # bit-exact FP32: no FMA contraction (the reference computes v0 + a*(v1-v0)
    # and the trilinear weights as separate multiplies and adds)
    "-ffp-contract=off",
    "-fno-fast-math",
    # packed FP32 instructions stay off (compiler-generated v_pk_* FP32 ops)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def hipcc() -> str:
    p = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found (ROCm is required to build libvbt_hip.so)")
    return p


def build_hip(jobs: int = 4):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [hipcc()] + HIPCC_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            for p in procs:
                if p.wait() != 0:
                    raise RuntimeError("hipcc failed")
            procs = []
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    out = os.path.join(HERE, "libvbt_hip.so")
    tmp = out + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs +
         [f"-Wl,-rpath,{ROCM}/lib", f"-L{ROCM}/lib", "-lamdhip64"])
    os.replace(tmp, out)
    return out


def build_synth():
    out = os.path.join(HERE, "libvbt_synth.so")
    tmp = out + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-fopenmp", "-fPIC", "-Wall", "-Wextra", "-shared", "-o", tmp,
          os.path.join(CSRC, "tools", "vbt_synth.cpp")])
    os.replace(tmp, out)
    return out


def build_loadgen():
    out = os.path.join(HERE, "libvbt_loadgen.so")
    tmp = out + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-Wall", "-Wextra", "-shared", "-D__HIP_PLATFORM_AMD__",
          f"-I{ROCM}/include", "-o", tmp, os.path.join(CSRC, "tools", "vbt_loadgen.cpp"),
          f"-L{HERE}", "-lvbt_hip", f"-L{ROCM}/lib", "-lamdhip64", "-lpthread",
          "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, out)
    return out


def build_cxx():
    """vbt_amd/libvbt_hip_cxx.so: the reference's C++ decode API as exported
    functions over the C-ABI (links libvbt_hip.so)."""
    out = os.path.join(HERE, "libvbt_hip_cxx.so")
    tmp = out + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-Wall", "-Wextra", "-shared", "-D__HIP_PLATFORM_AMD__",
          f"-I{ROCM}/include", "-o", tmp, os.path.join(CSRC, "vbt_cxx.cpp"), f"-L{HERE}", "-lvbt_hip",
          f"-L{ROCM}/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, out)
    return out


def build_shim_check():
    """tests/native/shim_check: a consumer of the C++ drop-in headers
    (include/vbt/*.h) linked against libvbt_hip_cxx.so, used by
    tests/test_cpp_shim.py."""
    out = os.path.join(ROOT, "tests", "native", "shim_check")
    tmp = out + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", "-o", tmp,
          os.path.join(ROOT, "tests", "native", "shim_check.cpp"), f"-L{HERE}", "-lvbt_hip_cxx", "-lvbt_hip",
          f"-L{ROCM}/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN/../../vbt_amd", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, out)
    return out


def build_oracle():
    _run(["make", "-C", os.path.join(ROOT, "oracle")])
    if os.path.isdir("/root/reference"):
        # reference decode + encoders (checker only; needs the reference sources)
        _run(["make", "-C", os.path.join(ROOT, "oracle"), "ref", "-j4"])


def build_dropin():
    """The reference's own unit tests and a render tool linked against
    libvbt_hip_cxx.so (oracle/_ref/ref_*; needs the reference sources)."""
    if os.path.isdir("/root/reference"):
        _run(["make", "-C", os.path.join(ROOT, "oracle"), "dropin", "-j4"])


def build_all():
    build_synth()
    build_oracle()
    build_hip()
    build_cxx()
    build_loadgen()
    build_shim_check()
    build_dropin()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "synth": build_synth, "oracle": build_oracle,
         "loadgen": build_loadgen, "cxx": build_cxx, "shim": build_shim_check,
         "dropin": build_dropin}[w]()

"""Build the native library noor_amd/libnoor_hip.so with hipcc for gfx950.

Every HIP source is compiled for gfx950 only, with FMA contraction off (the
kernels match the CPU restatement operation by operation), IEEE division and
square root, and packed-FP32 instructions disabled.  The host scene builder is
compiled into the same library.
"""
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libnoor_hip.so")
BUILD = os.path.join(ROOT, "build", "noor_amd")

HIP_SOURCES = ["capi.hip", "kernels_mega.hip", "kernels_env.hip", "kernels_wave.hip", "kernels_adjoint.hip",
               "bvh_build.hip"]
CXX_SOURCES = ["scene_build.cpp", "wide_bvh.cpp", "texture_store.cpp", "loaders.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable", "-I", os.path.join(ROOT, "include")]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-fhip-fp32-correctly-rounded-divide-sqrt"]


def _run(cmd):
    t = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")
    return time.time() - t


def build(verbose=True, jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    objs, cmds = [], []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        cmds.append([HIPCC, *COMMON, *HIP_FLAGS, "-c", "-o", o, os.path.join(CSRC, s)])
        objs.append(o)
    for s in CXX_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        cmds.append(["g++", *COMMON, "-pthread", "-c", "-o", o, os.path.join(CSRC, s)])
        objs.append(o)
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for cmd, dt in zip(cmds, ex.map(_run, cmds)):
            if verbose:
                print(f"  compiled {os.path.basename(cmd[-1])} in {dt:.1f}s", flush=True)
    tmp = OUT + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-pthread", "-lz"])
    os.replace(tmp, OUT)
    if verbose:
        print(f"  linked {OUT}", flush=True)
    return OUT


if __name__ == "__main__":
    build()

"""Build libnbvh_hip.so in-tree: host C++ with g++, HIP kernels with hipcc for gfx950.

Usage: python -m nbvh_amd.build   (or nbvh_amd.build.build())
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnbvh_hip.so")
DRIVER = os.path.join(HERE, "nbvh_driver")
BENCH_LIB = os.path.join(HERE, "libnbvh_bench.so")
ROOT = os.path.dirname(HERE)
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fvisibility=hidden", "-Wall",
              "-Wno-unused-function", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include")]
# -ffp-contract=off: multiply-adds are fused only where fmaf() is written (bit-exact FP32 path).
# -packed-fp32-ops: no v_pk_*_f32 generation.
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH, "-ffp-contract=off", "-fvisibility=hidden",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def _sources():
    host = sorted(os.path.join(CSRC, "host", f) for f in os.listdir(os.path.join(CSRC, "host")) if f.endswith(".cpp"))
    hip = sorted(os.path.join(CSRC, "kernels", f) for f in os.listdir(os.path.join(CSRC, "kernels"))
                 if f.endswith(".hip"))
    return host, hip


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose=False, jobs=None):
    """Compile every source and link nbvh_amd/libnbvh_hip.so (always rebuilds)."""
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    os.makedirs(OUT_DIR, exist_ok=True)
    host, hip = _sources()
    cmds, objs = [], []
    for src in host:
        obj = os.path.join(OUT_DIR, os.path.basename(src) + ".o")
        cmds.append(["g++"] + HOST_FLAGS + ["-c", src, "-o", obj])
        objs.append(obj)
    for src in hip:
        obj = os.path.join(OUT_DIR, os.path.basename(src) + ".o")
        cmds.append([hipcc] + HIP_FLAGS + ["-c", src, "-o", obj])
        objs.append(obj)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # Measurement helper of the benchmark (random-gather line rate), a library of its own.
    bench_src = os.path.join(CSRC, "bench", "gather_ceiling.hip")
    tmp = BENCH_LIB + ".tmp"
    _run([hipcc] + HIP_FLAGS + ["-shared", bench_src, "-o", tmp])
    os.replace(tmp, BENCH_LIB)
    # C++ host driver over include/nbvh.hpp, linked against the library just built.
    tmp = DRIVER + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROOT, "include"),
          "-I" + os.path.join(ROCM, "include"), os.path.join(CSRC, "driver", "nbvh_driver.cpp"), "-o", tmp,
          "-L" + HERE, "-lnbvh_hip", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath,$ORIGIN",
          "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, DRIVER)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

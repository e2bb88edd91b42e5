"""Builds libhcm.so in-tree (hypercraft_amd/_build/) for gfx950.

The kernels are compiled by hipcc for gfx950 only; host-only C++ (registry, workload
generators) is compiled by g++. Packed FP32 instructions are disabled for every HIP
translation unit.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libhcm.so")
WRAPPER_CHECK = os.path.join(OUT, "hcm_wrapper_check")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall"]


def _run(cmds, verbose):
    procs = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 target feature; that warning is expected
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if p.returncode != 0:
            failed.append((cmd, "\n".join(lines)))
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        msg = "\n\n".join(" ".join(c) + "\n" + t for c, t in failed)
        raise RuntimeError("build failed:\n" + msg)


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    src = lambda n: os.path.join(CSRC, n)
    obj = lambda n: os.path.join(OUT, n + ".o")
    hip_units = ["hcm_kernels.hip", "hcm_light.hip", "hcm_draw.hip", "hcm_api.cpp", "hcm_serve.cpp"]
    cxx_units = ["hcm_registry.cpp", "hcm_workload.cpp"]
    cmds = [[HIPCC, *HIP_FLAGS, "-c", src(u), "-o", obj(u)] for u in hip_units]
    cmds += [["g++", *CXX_FLAGS, "-c", src(u), "-o", obj(u)] for u in cxx_units]
    _run(cmds, verbose)
    tmp = LIB + ".tmp"
    _run([[HIPCC, "--offload-arch=" + ARCH, "-shared", "-o", tmp] + [obj(u) for u in hip_units + cxx_units]],
         verbose)
    os.replace(tmp, LIB)
    # the C++ front end (include/hcm.hpp) built into a small check program the tests run
    root = os.path.dirname(HERE)
    _run([["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(root, "include"),
           os.path.join(root, "tests", "cpp", "wrapper_check.cpp"), "-o", WRAPPER_CHECK,
           "-L", OUT, "-lhcm", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]], verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

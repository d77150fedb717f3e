"""Build the gfx950 shared library sc_amd/lib/libscamd.so with hipcc.

One hipcc invocation compiles the HIP kernels + C ABI (csrc/sc_gpu.hip) and the
host-side input formats (csrc/sc_host.cpp).  Packed FP32 instructions are
disabled (-target-feature -packed-fp32-ops) and no fast-math flag is used, so
IEEE comparisons in the pair code behave as in the reference.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib", "libscamd.so")
SOURCES = [os.path.join(HERE, "csrc", "sc_gpu.hip"), os.path.join(HERE, "csrc", "sc_host.cpp")]
HEADERS = [os.path.join(HERE, "csrc", h) for h in ("sc_pair.h", "sc_rng.h", "sc_host.h", "sc_mlp.h")] + [
    os.path.join(ROOT, "include", "sc_gpu.h")]

# FP32 division and square root as the hardware's reciprocal and square root
# (about 1 ulp) instead of the correctly rounded sequences: the pair terms are
# FP32 anyway, and the sweep's chain of trials waits on ~35 of them per patch
# pair.  FP32 denormals flushed to zero: without it every division and square
# root carries a range-scaling guard (frexp/ldexp around the reciprocal, a
# scaled square root below FLT_MIN), 5-7 more instructions on those chains;
# for operands and results in the normal range the values are the same.
# FP64 keeps IEEE denormals.
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
               "-Wno-unused-result", "-Wno-unused-value", "-fno-hip-fp32-correctly-rounded-divide-sqrt",
               "-fgpu-flush-denormals-to-zero",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    return "hipcc"


def build(verbose=False):
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    srcs = [s for s in SOURCES if os.path.exists(s)]
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I" + os.path.join(ROOT, "include")] + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("hipcc failed building %s" % LIB)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))

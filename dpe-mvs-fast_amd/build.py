"""Build libdpe_hip.so (gfx950) in-tree with hipcc.

Flags: -ffp-contract=off keeps the numerics contract (fused multiply-adds only
where the source writes fmaf); -fhip-fp32-correctly-rounded-divide-sqrt keeps
IEEE divide/sqrt; packed fp32 ops are disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libdpe_hip.so")
SOURCES = [os.path.join(HERE, "csrc", "dpe_kernels.hip"), os.path.join(HERE, "csrc", "dpe_fuse.hip"),
           os.path.join(HERE, "csrc", "dpe_capi.hip")]
HEADERS = [os.path.join(HERE, "csrc", "dpe_device.h"), os.path.join(HERE, "csrc", "dpe_state.h"),
           os.path.join(HERE, "csrc", "dpe_fuse.h"),
           os.path.join(ROOT, "include", "dpe_hip.h")]

HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
               "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang",
               "-packed-fp32-ops", "-Wno-unused-result"]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c


def build(verbose=True, jobs=2):
    """Compile every HIP source for gfx950 and link libdpe_hip.so (always rebuilds)."""
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(HERE, "csrc", os.path.basename(src) + ".o")
        cmd = [hipcc()] + [f for f in HIPCC_FLAGS if f != "-shared"] + ["-c", src, "-o", obj,
                                                                      "-I", os.path.join(ROOT, "include")]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        objs.append(obj)
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)

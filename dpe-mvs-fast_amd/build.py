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
SOURCES = [os.path.join(HERE, "csrc", "dpe_kernels.hip"), os.path.join(HERE, "csrc", "dpe_adjoint.hip"),
           os.path.join(HERE, "csrc", "dpe_fuse.hip"), os.path.join(HERE, "csrc", "dpe_capi.hip")]
HEADERS = [os.path.join(HERE, "csrc", "dpe_device.h"), os.path.join(HERE, "csrc", "dpe_state.h"),
           os.path.join(HERE, "csrc", "dpe_ncc.h"),
           os.path.join(HERE, "csrc", "dpe_fuse.h"),
           os.path.join(ROOT, "include", "dpe_hip.h")]

HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
               "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang",
               "-packed-fp32-ops", "-Wno-unused-result"]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c


# exhaustive check of the projection's reciprocal (a GPU test runs it)
RCP_CHECK = os.path.join(HERE, "dpe_rcp_check")
RCP_CHECK_SOURCE = os.path.join(HERE, "csrc", "dpe_rcp_check.hip")

HOST_LIB = os.path.join(HERE, "libdpe_host.so")
DRIVER = os.path.join(HERE, "dpe_driver")
HOST_SOURCES = [os.path.join(HERE, "host", "dpe_hostio.cpp"), os.path.join(HERE, "host", "dpe_edges.cpp"),
                os.path.join(HERE, "host", "dpe_jpeg.cpp")]
DRIVER_SOURCES = [os.path.join(HERE, "host", "dpe_driver.cpp")]
GXX_FLAGS = ["-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include")]


def build_host(verbose=True):
    """libdpe_host.so (formats, pyramid, support maps; CPU only) and the
    dpe_driver binary over it and libdpe_hip.so (rpath $ORIGIN)."""
    cmd = ["g++"] + GXX_FLAGS + ["-fPIC", "-shared", "-o", HOST_LIB + ".tmp"] + HOST_SOURCES
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(HOST_LIB + ".tmp", HOST_LIB)
    cmd = (["g++"] + GXX_FLAGS + ["-o", DRIVER + ".tmp"] + DRIVER_SOURCES +
           ["-L", HERE, "-ldpe_hip", "-ldpe_host", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"])
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(DRIVER + ".tmp", DRIVER)
    return HOST_LIB, DRIVER


def build(verbose=True, jobs=2):
    """Compile every HIP source for gfx950 and link libdpe_hip.so, then the host
    library and the driver (always rebuilds)."""
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
    cmd = [hipcc()] + [f for f in HIPCC_FLAGS if f not in ("-shared", "-fPIC")] + [
        RCP_CHECK_SOURCE, "-o", RCP_CHECK + ".tmp", "-I", os.path.join(ROOT, "include")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(RCP_CHECK + ".tmp", RCP_CHECK)
    build_host(verbose)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)

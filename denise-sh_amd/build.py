"""Build libdenise_sh.so (HIP kernels + C-ABI + host setup) for gfx950.

Usage: python denise-sh_amd/build.py [--jobs N]

The library is built in-tree next to this file so that it travels with the
repository snapshot to the GPU host.  Every translation unit is compiled with
-ffp-contract=off: the kernels reproduce the reference's unfused FP32
arithmetic, and the host setup reproduces its coefficient tables bit for bit.
"""
import argparse
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdenise_sh.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-result",
          "-I" + CSRC, "-I" + os.path.join(os.path.dirname(HERE), "include")]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-fno-gpu-rdc"]

SOURCES = [
    ("dn_kernels.hip", True),
    ("dn_core.cpp", False),
    ("dn_setup.cpp", False),
    ("dn_dropin.cpp", False),
    ("dn_fwi.cpp", False),
    ("dn_io.cpp", False),
    ("dn_unet.hip", True),
    ("dn_unet.cpp", False),
    ("dn_gradproc.hip", True),
    ("dn_stf.hip", True),
    ("dn_misfit.hip", True),
    ("dn_misfit_noalloc.hip", True),
]


def compile_one(src, device, objdir):
    path = os.path.join(CSRC, src)
    obj = os.path.join(objdir, src + ".o")
    # host translation units include the HIP runtime and are compiled as HIP
    # as well (for gfx950 only) so no default offload target sneaks in
    cmd = [HIPCC] + COMMON + ["-x", "hip"] + DEVICE
    cmd += ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    if r.stderr.strip():
        # the host pass warns about the device-only target feature; keep the rest
        lines = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l]
        if any(l.strip() for l in lines):
            sys.stderr.write("\n".join(lines) + "\n")
    return obj


def build(jobs=8, verbose=True):
    if not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found at %s" % HIPCC)
    srcs = [(s, d) for s, d in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objdir = tempfile.mkdtemp(prefix="dnbuild_")
    try:
        with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
            objs = list(ex.map(lambda sd: compile_one(sd[0], sd[1], objdir), srcs))
        tmp = LIB + ".tmp"
        cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
        os.replace(tmp, LIB)
    finally:
        shutil.rmtree(objdir, ignore_errors=True)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    args = ap.parse_args()
    build(args.jobs)

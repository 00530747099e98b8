"""Build libcuev_amd.so (HIP kernels + C ABI) for gfx950 with hipcc.

Usage: python -m cuev_amd.build [-j N] [--debug]

Every source under cuev_amd/csrc is compiled to an object in cuev_amd/_build and
linked into cuev_amd/libcuev_amd.so, which the Python package loads with ctypes.
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcuev_amd.so")
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CUEV_OFFLOAD_ARCH", "gfx950")

# Packed FP32 VALU ops are disabled for every device compilation (hipcc warns
# that the host pass does not know the feature; that is expected).
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(CSRC, f))
    return out


def compile_one(src, debug):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, "-c", src, "-o", obj, "-std=c++17", "-fPIC",
           "-O1" if debug else "-O3", "-g0",
           f"--offload-arch={ARCH}", "-I", os.path.join(ROOT, "include"),
           "-Wall", "-Wno-unused-result", "-Wno-unused-function", "-Wno-unknown-pragmas"]
    if src.endswith(".hip"):
        cmd = cmd[:2] + ["-x", "hip"] + cmd[2:] + DEVICE_FLAGS
    else:
        cmd += ["-D__HIP_PLATFORM_AMD__"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    # drop the expected host-pass warning about the device-only feature flag
    err = "\n".join(l for l in p.stderr.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    return src, obj, p.returncode, err


def build(jobs=None, debug=False, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(16, os.cpu_count() or 4)
    srcs = sources()
    objs = []
    failed = []
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, obj, rc, err in ex.map(lambda s: compile_one(s, debug), srcs):
            if rc != 0:
                failed.append((src, err))
            elif err and verbose:
                print(f"[{os.path.basename(src)}]\n{err}", file=sys.stderr)
            objs.append(obj)
    if failed:
        for src, err in failed:
            print(f"==== {src}\n{err}", file=sys.stderr)
        raise RuntimeError(f"hipcc failed for {len(failed)} source(s)")
    tmp = LIB + ".tmp"
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + [
        f"-L{rocm_lib}", f"-Wl,-rpath,{rocm_lib}", "-ldl", "-lm", "-lpthread"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        print(p.stderr, file=sys.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--debug", action="store_true")
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.debug, a.v))


if __name__ == "__main__":
    main()

"""The C++ drop-in header (include/cuev.h) as a reference caller uses it.

tests/cpp/drop_in_caller.cpp calls cuev::symm_eig_solve<T>(dA, n, d_eval, d_evec,
stream, &timer) in the shape of the reference's bench/bench_solver.cpp:200-214.
On the CPU it is compiled and linked against libcuev_amd.so with hipcc (the header
and the exported C ABI must agree); on the GPU it runs and checks its own result.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "drop_in_caller.cpp")
LIBDIR = os.path.join(ROOT, "cuev_amd")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build(out_dir):
    exe = os.path.join(out_dir, "drop_in_caller")
    cmd = [HIPCC, "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
           SRC, "-o", exe, f"-L{LIBDIR}", "-lcuev_amd", f"-Wl,-rpath,{LIBDIR}"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    return exe


def test_drop_in_header_compiles_and_links(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libcuev_amd.so")):
        pytest.fail("libcuev_amd.so is missing: run `python -m cuev_amd.build`")
    exe = build(str(tmp_path))
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_drop_in_caller_runs(tmp_path, torch_dev):
    exe = build(str(tmp_path))
    p = subprocess.run([exe, "300"], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr)
    assert p.returncode == 0, (p.stdout, p.stderr)
    assert p.stdout.count("drop_in_caller") == 2

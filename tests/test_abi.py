"""CPU checks of the C-ABI library: it loads and exports every entry point of include/*.h.

No compute call is made here (no GPU in this container); the pure host helpers
(status strings, option defaults, the LAPACK probe) are exercised.
"""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(cuev_[a-z0-9_]+)\s*\(", text):
            syms.add(m.group(1))
    return syms


@pytest.fixture(scope="module")
def L():
    import cuev_amd
    if not os.path.exists(cuev_amd.lib_path):
        pytest.fail("libcuev_amd.so is missing: run `python -m cuev_amd.build`")
    return cuev_amd.lib()


def test_all_declared_symbols_exported(L):
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in sorted(syms) if not hasattr(L, s)]
    assert not missing, f"not exported: {missing}"


def test_status_strings(L):
    for code in range(10):
        assert L.cuev_status_string(code)
    assert b"watchdog" in L.cuev_status_string(6)


def test_options_defaults(L):
    import cuev_amd
    o = cuev_amd.Options(7, 7, 7, 7, 7, 7, 7, 7.0)
    L.cuev_options_init(ctypes.byref(o))
    assert (o.bandwidth, o.block, o.precision, o.refine_sweeps, o.bt_mode, o.dc_mode, o.dc_leaf,
            o.refine_tol) == (0, 0, 0, 0, 0, 0, 0, 0.0)
    # 7 ints, then the double at offset 32
    assert ctypes.sizeof(cuev_amd.Options) == 40
    assert cuev_amd.Options.refine_tol.offset == 32


def test_timer_layout_matches_header():
    import cuev_amd
    # 11 floats + 2 ints + 2 floats, all 4 bytes
    assert ctypes.sizeof(cuev_amd.SolveTimer) == 15 * 4


def test_host_lapack_found(L):
    import cuev_amd
    name = cuev_amd.lapack_name()
    assert name, "no host LAPACK with dstedc found"


def test_product_does_not_reference_oracle():
    """The shipped package must not import or load anything under oracle/."""
    pkg = os.path.join(ROOT, "cuev_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".hpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower(), f"{f} mentions the oracle"

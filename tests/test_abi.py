"""CPU-side checks of the C ABI: the library loads and exports every declared entry point.

No compute call is made here (this container has no GPU).
"""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in ("include/qsf_hip.h",):
        src = open(os.path.join(ROOT, h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:int|const char\*)\s+(qsf_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("qsf_fourier", "qsf_evolve", "qsf_average", "qsf_flux", "qsf_mask_region",
                     "qsf_set_post_compute", "qsf_prop_step", "qsf_prop_run_steps"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from qsf_amd import hip
    if not os.path.exists(hip.LIB_PATH):
        pytest.skip("libqsf_hip.so not built")
    lib = ctypes.CDLL(hip.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared_functions()) <= set(hip.SIGNATURES)


def test_types_mirror_header_sizes():
    from qsf_amd import qsf_types as T
    # offsets of the C structs (x86-64 / LP64 layout)
    assert ctypes.sizeof(T.GridDesc) == 120
    assert ctypes.sizeof(T.FieldDesc) == 64
    assert ctypes.sizeof(T.PropDesc) == 176


def test_enum_values_mirror_header():
    # every `QSF_NAME = value` enumerator of qsf_types.h has the same value in qsf_types.py
    from qsf_amd import qsf_types as T
    src = open(os.path.join(ROOT, "include/qsf_types.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    found = 0
    for body in re.findall(r"enum\s*\{(.*?)\}", src, flags=re.S):
        for name, value, shift in re.findall(r"QSF_(\w+)\s*=\s*(-?\d+)(?:\s*<<\s*(\d+))?", body):
            if hasattr(T, name):
                assert getattr(T, name) == int(value) << int(shift or 0), name
                found += 1
    assert found >= 3 and T.FP32_STATE == 2

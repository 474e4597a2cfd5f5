"""The C-ABI library loads and exports every entry point its header declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from gpumetropolis_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            names |= set(re.findall(r"\b(gpum_hip_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared()
    for want in ("gpum_hip_run_chains", "gpum_hip_available",
                 "gpum_hip_last_error", "gpum_hip_result_free",
                 "gpum_hip_loglik_batch", "gpum_hip_loglik_pointwise",
                 "gpum_hip_grad_batch"):
        assert want in names


def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(_native.LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_header():
    # field order and sizes of the ctypes mirrors against the header text
    text = open(os.path.join(ROOT, "include", "gpum_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    body = text[text.index("typedef struct gpum_run_ext"):]
    body = body[:body.index("} gpum_run_ext;")]
    fields = re.findall(r"^\s+[\w\s\*]+?\b(\w+);", body, re.M)
    assert fields == [f for f, _ in _native.RunExt._fields_]
    body = text[text.index("typedef struct gpum_chain_result"):]
    body = body[:body.index("} gpum_chain_result;")]
    fields = re.findall(r"^\s+[\w\s\*]+?\b(\w+);", body, re.M)
    assert fields == [f for f, _ in _native.ChainResult._fields_]

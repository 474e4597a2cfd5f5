"""The C-ABI library loads and exports every entry point its header declares,
the ctypes mirrors match the C compiler's layout of every header struct, and
a C program links against the library and calls it (no compute calls: this
runs without a GPU)."""
import ctypes
import json
import os
import re
import subprocess
import tempfile

import pytest

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


# header struct -> ctypes mirror
STRUCTS = {
    "gpum_chain_result": _native.ChainResult,
    "gpum_run_ext": _native.RunExt,
    "gpum_model_info": _native.ModelInfo,
    "gpum_train_opts": _native.TrainOpts,
    "gpum_da_ext": _native.DaExt,
    "gpum_batch_model": _native.BatchModel,
    "gpum_batch_opts": _native.BatchOpts,
    "gpum_warmup_opts": _native.WarmupOpts,
}


def _header_fields(name):
    text = open(os.path.join(ROOT, "include", "gpum_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    body = text[text.index(f"typedef struct {name} {{"):]
    body = body[:body.index(f"}} {name};")]
    return re.findall(r"^\s+[\w\s\*]+?\b(\w+);", body, re.M)


def _gcc(src, out, extra=()):
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I",
                    os.path.join(ROOT, "include"), src, "-o", out, *extra],
                   check=True, capture_output=True, text=True)


def test_every_header_struct_has_a_mirror():
    text = open(os.path.join(ROOT, "include", "gpum_hip.h")).read()
    named = set(re.findall(r"typedef struct (\w+) \{", text))
    assert named == set(STRUCTS)


def test_struct_layouts_match_the_c_compiler():
    """sizeof and every field's offsetof, as gcc lays the header out, equal
    the ctypes mirrors' (what a Rust #[repr(C)] mirror must also match)."""
    lines = ["#include <stdio.h>", "#include <stddef.h>", "#include \"gpum_hip.h\"",
             "int main(void) {", '  printf("{");']
    first = True
    for name in STRUCTS:
        fields = _header_fields(name)
        sep = "" if first else ","
        first = False
        lines.append(f'  printf("{sep}\\"{name}\\": {{\\"sizeof\\": %zu", sizeof({name}));')
        for f in fields:
            lines.append(f'  printf(", \\"{f}\\": %zu", offsetof({name}, {f}));')
        lines.append('  printf("}");')
    lines += ['  printf("}\\n");', "  return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "probe.c"), os.path.join(d, "probe")
        open(src, "w").write("\n".join(lines) + "\n")
        _gcc(src, exe)
        got = json.loads(subprocess.run([exe], check=True, capture_output=True,
                                        text=True).stdout)
    for name, mirror in STRUCTS.items():
        c = got[name]
        assert c["sizeof"] == ctypes.sizeof(mirror), name
        names = [f for f, _ in mirror._fields_]
        assert names == _header_fields(name), name
        for f in names:
            assert c[f] == getattr(mirror, f).offset, (name, f)


def test_c_program_links_and_calls_the_abi():
    """A C caller (as a Rust -sys crate would be) links the library and calls
    its entry points; without a GPU the run call fails with a message."""
    src_text = r"""
#include <stdio.h>
#include <string.h>
#include "gpum_hip.h"
int main(void) {
  printf("abi %d\n", GPUM_HIP_ABI_VERSION);
  int ok = gpum_hip_available();
  printf("available %d\n", ok);
  if (!ok) {
    const uint32_t code[2] = {1, 0};
    const double y = 0.0, init = 0.0, sd = 1.0, temp = 1.0;
    gpum_chain_result r;
    memset(&r, 0, sizeof r);
    int rc = gpum_hip_run_chains(code, 2, NULL, 0, NULL, 0, NULL, 0, 1, &y, 1,
                                 1, &init, 1, &sd, &temp, 4, 1, 0, 0.0, 0.0,
                                 NULL, &r);
    printf("rc %d\n", rc);
    printf("msg %s\n", gpum_hip_last_error());
    gpum_hip_result_free(&r);
  }
  return 0;
}
"""
    libdir = os.path.dirname(_native.LIB)
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "call.c"), os.path.join(d, "call")
        open(src, "w").write(src_text)
        _gcc(src, exe, ["-L", libdir, "-lgpum_hip", f"-Wl,-rpath,{libdir}"])
        out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = dict(l.split(" ", 1) for l in out.stdout.strip().splitlines())
    assert lines["abi"] == "1"
    if lines["available"] == "0":
        assert lines["rc"] != "0"
        assert lines["msg"].strip()


# The Rust binding written for the reference crate (integration/rust,
# uncompiled here): every #[repr(C)] struct mirrors its header struct field
# for field, with the C types mapped to the Rust ones.
RUST_STRUCTS = {
    "GpumChainResult": "gpum_chain_result", "GpumRunExt": "gpum_run_ext",
    "GpumModelInfo": "gpum_model_info", "GpumTrainOpts": "gpum_train_opts",
    "GpumDaExt": "gpum_da_ext", "GpumBatchModel": "gpum_batch_model",
    "GpumBatchOpts": "gpum_batch_opts", "GpumWarmupOpts": "gpum_warmup_opts",
}
C_TO_RUST = {"uint32_t": "u32", "int32_t": "i32", "double": "f64",
             "float": "f32", "uint8_t": "u8"}


def _header_typed_fields(name):
    text = open(os.path.join(ROOT, "include", "gpum_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    body = text[text.index(f"typedef struct {name} {{"):]
    body = body[body.index("{") + 1:body.index(f"}} {name};")]
    out = []
    for decl in body.split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        m = re.match(r"(const )?(\w+)\s*(\*?)\s*(\w+)$", decl)
        assert m, decl
        const, ctype, ptr, field = m.groups()
        if ptr:
            rt = C_TO_RUST.get(ctype, "Gpum")
            out.append((field, ("*const " if const else "*mut ") + rt))
        else:
            out.append((field, C_TO_RUST[ctype]))
    return out


def test_rust_binding_mirrors_the_header():
    src = open(os.path.join(ROOT, "integration", "rust", "hip_sys.rs")).read()
    for rname, cname in RUST_STRUCTS.items():
        body = src[src.index(f"pub struct {rname} {{"):]
        body = body[body.index("{") + 1:body.index("}")]
        rfields = re.findall(r"pub (\w+): ([^,]+),", body)
        want = _header_typed_fields(cname)
        assert [f for f, _ in rfields] == [f for f, _ in want], rname
        for (f, rt), (_, ct) in zip(rfields, want):
            if ct.endswith("Gpum"):            # pointer to an opaque handle
                assert rt.startswith(ct[:-4] + "Gpum"), (rname, f, rt)
            else:
                assert rt == ct, (rname, f, rt, ct)
    for name in declared():
        if name.startswith("gpum_hip_") and name in (
                "gpum_hip_run_chains", "gpum_hip_run_chains_ex",
                "gpum_hip_result_free", "gpum_hip_available",
                "gpum_hip_last_error", "gpum_hip_loglik_pointwise",
                "gpum_hip_grad_batch", "gpum_hip_run_de_sync",
                "gpum_hip_batch_run", "gpum_hip_model_run_da",
                "gpum_hip_model_warmup"):
            assert f"pub fn {name}(" in src, name

"""Per-model compiled log-likelihood for the CPU baseline.

TEST AND CPU-BASELINE INFRASTRUCTURE ONLY.  Mirrors the reference's Cranelift JIT
(src/rust/src/jit.rs:441-631): the bytecode becomes straight-line f64 code
inside the observation loop, accumulated in observation order, calling libm
exp/log/pow and the Lanczos log-gamma.  The C source is compiled with the
host gcc at -O2 without -ffast-math (Cranelift neither reassociates nor
vectorises), and its function pointer is handed to the oracle's sampler.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
import subprocess
import tempfile

from . import oracle

_BIN = {3: "+", 4: "-", 5: "*", 6: "/"}
_cache = {}


def c_source(code, consts) -> str:
    lines = []
    st = []
    tmp = 0
    for pc in range(len(code) // 2):
        op, arg = int(code[2 * pc]), int(code[2 * pc + 1])
        if op == 0:
            st.append(repr(float(consts[arg])))
            continue
        if op == 1:
            st.append(f"p[{arg}]")
            continue
        if op == 2:
            st.append(f"row[{arg}]")
            continue
        if op in (7, 8, 9, 10, 12):
            a = st.pop()
            expr = {7: f"-({a})", 8: f"exp({a})", 9: f"log({a})",
                    10: f"sqrt({a})", 12: f"orc_lgamma({a})"}[op]
        else:
            b = st.pop()
            a = st.pop()
            expr = f"({a}) {_BIN[op]} ({b})" if op in _BIN else f"pow({a}, {b})"
        name = f"t{tmp}"
        tmp += 1
        lines.append(f"    const double {name} = {expr};")
        st.append(name)
    term = st[-1] if st else "0.0"
    body = "\n".join(lines)
    return f"""#include <math.h>
#include <stddef.h>
#include <stdint.h>
double orc_lgamma(double);
double gpum_loglik(const double* p, const double* data, uint32_t n_obs,
                   uint32_t n_cols) {{
  double acc = 0.0;
  for (uint32_t i = 0; i < n_obs; ++i) {{
    const double* row = data + (size_t)i * n_cols;
    (void)row;
{body}
    acc += {term};
  }}
  return acc;
}}
"""


def compile_loglik(code, consts):
    """Compile and load; returns (ctypes function, address)."""
    src = c_source(code, consts)
    key = hashlib.sha1(src.encode()).hexdigest()[:16]
    if key in _cache:
        return _cache[key]
    oracle.lib()
    d = os.path.join(tempfile.gettempdir(), "gpum_codegen")
    os.makedirs(d, exist_ok=True)
    cfile = os.path.join(d, f"ll_{key}.c")
    so = os.path.join(d, f"ll_{key}.so")
    if not os.path.exists(so):
        with open(cfile, "w") as f:
            f.write(src)
        subprocess.run(["gcc", "-O2", "-fPIC", "-shared", cfile, "-o", so,
                        f"-L{oracle.HERE}", "-loracle", f"-Wl,-rpath,{oracle.HERE}",
                        "-lm"], check=True)
    lib = C.CDLL(so)
    fn = lib.gpum_loglik
    fn.restype = C.c_double
    fn.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_uint32,
                   C.c_uint32]
    addr = C.cast(fn, C.c_void_p).value
    _cache[key] = (lib, fn, addr)
    return _cache[key]

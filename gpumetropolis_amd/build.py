"""Build the gfx950 backend library in-tree.

``python -m gpumetropolis_amd.build`` compiles ``csrc/`` with hipcc into
``gpumetropolis_amd/lib/libgpum_hip.so`` (device code for gfx950 only):
one object per source, compiled in parallel under ``_build/``, then linked.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libgpum_hip.so")
OBJDIR = os.path.join(os.path.dirname(HERE), "_build")

# Slowest first: the interpreter's kernel instantiations dominate.
SOURCES = ["chain_kernel_inst0.hip", "chain_kernel_inst1.hip",
           "chain_kernel_inst2.hip", "chain_kernel_inst3.hip",
           "da_kernels.hip", "batch_kernels.hip", "mlp.hip", "chain_kernel.hip",
           "eval64.hip", "noalloc.hip", "warmup_kernels.hip", "gpum_hip.cpp", "da.cpp", "batch.cpp", "warmup.cpp", "jit.cpp",
           "lower.cpp", "eval64.cpp"]
HEADERS = ["vm.h", "lower.h", "chain_kernel.h", "chain_kernels.h",
           "device_vm.h", "math64.h", "philox.h", "runtime.h", "step_common.h",
           "bf16.h", "mlp.h", "da.h", "batch.h", "chain_body.h", "jit.h", "eval64_device.h", "eval64_kernels.h",
           "eval64.h", "warmup.h"]
# Device headers embedded into the library for the per-model kernels
# compiled at run time (csrc/jit.cpp).
JIT_HEADERS = ["chain_body.h", "chain_kernel.h", "device_vm.h", "math64.h",
               "philox.h", "step_common.h", "vm.h", "eval64.h",
               "eval64_device.h", "eval64_kernels.h"]
JIT_INC = os.path.join(CSRC, "jit_sources.inc")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Wno-unused-result", "-Wno-unused-value",
    # Contract a*b+c only within one expression, never across statements:
    # the per-model kernels (straight-line code) then round exactly as the
    # interpreter, which executes one VM instruction at a time.
    "-ffp-contract=on",
    # Packed FP32 arithmetic stays off (ordinary float code is emitted as
    # scalar v_*_f32).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    return "hipcc"


def write_jit_sources() -> None:
    """Embed the device headers as raw string literals (jit_sources.inc)."""
    names, bodies = [], []
    for h in JIT_HEADERS:
        with open(os.path.join(CSRC, h)) as f:
            text = f.read()
        if ")GPUMJIT" in text:
            raise RuntimeError(f"{h} contains the raw-string delimiter")
        names.append(f'"{h}"')
        bodies.append(f'R"GPUMJIT({text})GPUMJIT"')
    out = ("// Generated by gpumetropolis_amd/build.py from the device headers;"
           " do not edit.\n"
           f"static const int kJitHeaderCount = {len(names)};\n"
           "static const char* const kJitHeaderNames[] = {" + ", ".join(names)
           + "};\n"
           "static const char* const kJitHeaderSources[] = {\n"
           + ",\n".join(bodies) + "};\n")
    old = open(JIT_INC).read() if os.path.exists(JIT_INC) else None
    if old != out:
        with open(JIT_INC, "w") as f:
            f.write(out)


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    return r.returncode, r.stdout + r.stderr


def build(verbose: bool = False, jobs: int = 0) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    write_jit_sources()
    # Objects named after the whole source name (x.hip and x.cpp may pair).
    objs = [os.path.join(OBJDIR, s + ".o") for s in SOURCES]
    cmds = [[hipcc(), *HIPCC_FLAGS, "-c", os.path.join(CSRC, s), "-o", o]
            for s, o in zip(SOURCES, objs)]
    jobs = jobs or min(len(cmds), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        results = list(pool.map(lambda c: _run(c, verbose), cmds))
    failed = [(c, out) for c, (rc, out) in zip(cmds, results) if rc != 0]
    if failed:
        for c, out in failed:
            sys.stderr.write(" ".join(c) + "\n" + out)
        raise RuntimeError("hipcc failed building libgpum_hip.so")
    tmp = LIB + ".tmp"
    rc, out = _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs,
                    "-Wl,--no-undefined", "-L/opt/rocm/lib", "-lhiprtc",
                    "-lamdhip64", "-o", tmp], verbose)
    if rc != 0:
        sys.stderr.write(out)
        raise RuntimeError("linking libgpum_hip.so failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))

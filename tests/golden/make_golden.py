"""Regenerate the committed golden fixtures of this directory.

  philox_rocrand.json      Philox4x32-10 blocks from the open-source rocRAND
                           engine (host path of rocrand_philox4x32_10.h),
                           built from philox_rocrand.cpp with hipcc.
  reference_constants.json The triple32 shift/multiplier sequence and the
                           Lanczos coefficients as written in the reference
                           (src/rust/src/cpu_native.rs:13-23, :36-46), read
                           from the reference source as text, plus triple32
                           outputs computed from those constants.

Run from the repository root:  python tests/golden/make_golden.py
(the second fixture needs the reference checkout at /root/reference).
"""
import json
import os
import re
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/src/rust/src/cpu_native.rs"


def philox_rocrand():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "philox_rocrand")
        subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O1",
                        os.path.join(HERE, "philox_rocrand.cpp"), "-o", exe],
                       check=True, capture_output=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True)
    data = json.loads(out.stdout)
    with open(os.path.join(HERE, "philox_rocrand.json"), "w") as f:
        json.dump(data, f, indent=1)


def reference_constants():
    src = open(REF).read()
    body = src[src.index("fn triple32"):]
    body = body[:body.index("\n}\n")]
    steps = []
    for line in body.splitlines():
        m = re.search(r"h \^= h >> (\d+);", line)
        if m:
            steps.append(["xorshift", int(m.group(1))])
        m = re.search(r"wrapping_mul\((\d+)\)", line)
        if m:
            steps.append(["mul", int(m.group(1))])
    lz = src[src.index("const C: [f64; 9]"):]
    lz = lz[lz.index("[", lz.index("=")) + 1:lz.index("];")]
    coeffs = [float(t.replace("_", "")) for t in re.findall(r"-?[\d_]+\.[\d_]+(?:e-?\d+)?", lz)]

    def t32(h):
        for kind, v in steps:
            h = (h ^ (h >> v)) if kind == "xorshift" else (h * v) & 0xFFFFFFFF
        return h

    seeds = [0, 1, 7, 0xFFFFFFFF]
    outs = {str(s): [t32((s + k) & 0xFFFFFFFF) for k in range(16)] for s in seeds}
    data = {"source": "src/rust/src/cpu_native.rs",
            "triple32_steps": steps, "triple32_outputs": outs,
            "lanczos_g": 7.0, "lanczos_coeffs": coeffs}
    with open(os.path.join(HERE, "reference_constants.json"), "w") as f:
        json.dump(data, f, indent=1)


if __name__ == "__main__":
    philox_rocrand()
    if os.path.exists(REF):
        reference_constants()

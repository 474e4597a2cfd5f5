"""The FP32 log-gamma of the per-observation VM (csrc/device_vm.h,
vm_lgamma), restated step for step in numpy float32 and measured against a
50-digit log-gamma (mpmath), plus a check that the header holds the fitted
coefficients.  The device itself is compared with the oracle's FP64 Lanczos
log-gamma in tests/test_gpu_parity.py.
"""
import math
import os
import re

import numpy as np
import pytest

from gpumetropolis_amd.csrc import lgamma_fit

mp = pytest.importorskip("mpmath")
f = np.float32
HEADER = os.path.join(os.path.dirname(lgamma_fit.__file__), "device_vm.h")


def _header_coeffs():
    src = open(HEADER).read()
    body = src[src.index("vm_lgamma_pos"):src.index("const float poly")]
    vals = re.findall(r"([-+]?\d\.\d+e[-+]\d+)f", body)
    return np.array([float(v) for v in vals], np.float32)[::-1]  # c0 first


def test_header_holds_the_fitted_coefficients():
    np.testing.assert_array_equal(_header_coeffs(), lgamma_fit.fp32_coeffs())


def _pos(x, c):
    big = x >= f(8)
    k = f(np.floor(x) - f(2))
    z = f(x - k)
    pd = f(1)
    for j in range(1, 6):
        if f(j) <= k:
            pd = f(pd * f(x - f(j)))
    pu = f(x * f(x + f(1))) if x < f(1) else x
    p = pu if k < f(0) else pd
    lg = f(np.log(x if big else p))
    t = f(z - f(2.5))
    q = c[-1]
    for ci in c[-2::-1]:
        q = f(f(q * t) + ci)
    poly = f(f(z - f(2)) * q)
    small = f(poly - lg) if k < f(0) else f(poly + lg)
    inv = f(f(1) / x)
    inv2 = f(inv * inv)
    series = f(inv * f(f(1 / 12) - f(inv2 * f(f(1 / 360) - f(inv2 * f(1 / 1260))))))
    stir = f(f(f(f(x - f(0.5)) * lg) - x) + f(f(0.91893853320467274) + series))
    return stir if big else small


def _lgamma32(x, c):
    x = f(x)
    if x > f(0):
        return _pos(x, c)
    t = f(x * f(0.5))
    t = f(t - f(np.rint(t)))
    s = f(math.sin(2 * math.pi * float(t)))
    return f(f(np.log(abs(f(f(math.pi) * f(f(1) / s))))) - _pos(f(f(1) - x), c))


def test_fp32_algorithm_accuracy():
    mp.mp.dps = 30
    c = _header_coeffs()
    xs = np.concatenate([np.geomspace(1e-6, 0.5, 60), np.linspace(0.5, 8.5, 400),
                         np.geomspace(8.5, 1e6, 60)]).astype(np.float32)
    worst = 0.0
    for x in xs:
        ref = float(mp.loggamma(mp.mpf(float(x))))
        worst = max(worst, abs(float(_lgamma32(x, c)) - ref) / max(abs(ref), 1.0))
    assert worst < 3e-7, worst
    # negative arguments away from the poles: the conditioning of the
    # argument bounds the error
    for x in np.linspace(-9.7, -0.05, 97).astype(np.float32):
        ref = float(mp.loggamma(mp.mpf(float(x))).real)
        assert abs(float(_lgamma32(x, c)) - ref) <= 2e-6 * max(abs(ref), 1.0), x
    # far from the origin the reduced sine stays correct (no wrap); at the
    # poles the sine of a whole half-revolution is 0 or a rounding residue,
    # so the result is +inf or very large
    x = f(-1000.25)
    ref = float(mp.loggamma(mp.mpf(float(x))).real)
    assert abs(float(_lgamma32(x, c)) - ref) <= 1e-6 * abs(ref)
    with np.errstate(divide="ignore"):
        for x in (-3.0, -2.0, 0.0, -4096.0):
            assert _lgamma32(f(x), c) > 30.0, x

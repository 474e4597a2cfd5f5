"""Derivation of the FP32 log-gamma polynomial used by device_vm.h.

On [2, 3] log-gamma vanishes at z = 2, so it is written lgamma(z) =
(z - 2) q(z) with q smooth (q(2) = 1 - Euler's gamma, q(3) = ln 2).  q is
fitted in t = z - 2.5 by least squares on Chebyshev nodes against mpmath's
loggamma at 50 digits, the coefficients are rounded to FP32, and the
rounded polynomial is re-measured in FP32 Horner arithmetic.

    python -m gpumetropolis_amd.csrc.lgamma_fit      # prints the table

The printed table is what `vm_lgamma` holds.  `poly_coeffs()` is also used
by the CPU test that emulates the device routine in FP32.
"""
from __future__ import annotations

import numpy as np

DEGREE = 7


def poly_coeffs(degree: int = DEGREE) -> np.ndarray:
    """FP64 least-squares coefficients c[0..degree] of q(2.5 + t), t in
    [-0.5, 0.5], lowest order first."""
    import mpmath as mp
    mp.mp.dps = 50
    n = 4 * (degree + 1)
    k = np.arange(n)
    t = 0.5 * np.cos(np.pi * (k + 0.5) / n)
    q = []
    for tk in t:
        z = mp.mpf(2.5) + mp.mpf(float(tk))
        q.append(float(mp.loggamma(z) / (z - 2)))
    V = np.vander(t, degree + 1, increasing=True)
    c, *_ = np.linalg.lstsq(V, np.array(q), rcond=None)
    return c


def fp32_coeffs(degree: int = DEGREE) -> np.ndarray:
    return poly_coeffs(degree).astype(np.float32)


def _check(c32: np.ndarray) -> float:
    import mpmath as mp
    mp.mp.dps = 30
    f = np.float32
    worst = 0.0
    for z in np.linspace(2.0, 3.0, 2001, dtype=np.float32):
        t = f(z - f(2.5))
        acc = c32[-1]
        for ci in c32[-2::-1]:
            acc = f(f(acc * t) + ci)
        v = f(f(z - f(2.0)) * acc)
        ref = float(mp.loggamma(mp.mpf(float(z))))
        worst = max(worst, abs(float(v) - ref) / max(abs(ref), 1e-30))
    return worst


if __name__ == "__main__":
    c32 = fp32_coeffs()
    for i, ci in enumerate(c32):
        print(f"  c{i} = {float(ci):.9e}f")
    print(f"max relative error of (z-2) q(z) in FP32 on [2, 3]: {_check(c32):.2e}")

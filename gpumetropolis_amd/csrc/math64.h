// math64.h -- FP64 special functions shared by the host lowering and the
// device step code.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace gpum {

// Lanczos series for log-gamma, valid for x >= 0.5: g = 7 and the nine
// published coefficients (Lanczos 1964; the coefficient set of Numerical
// Recipes / Godfrey).  The terms are accumulated in index order.
__host__ __device__ inline double lgamma64_main(double x) {
  const double two_pi = 6.283185307179586;
  x = x - 1.0;
  double t = (x + 7.0) + 0.5;
  double a = 0.99999999999980993;
  a += 676.5203681218851 / (x + 1.0);
  a += -1259.1392167224028 / (x + 2.0);
  a += 771.32342877765313 / (x + 3.0);
  a += -176.61502916214059 / (x + 4.0);
  a += 12.507343278686905 / (x + 5.0);
  a += -0.13857109526572012 / (x + 6.0);
  a += 9.9843695780195716e-6 / (x + 7.0);
  a += 1.5056327351493116e-7 / (x + 8.0);
  return 0.5 * log(two_pi) + (x + 0.5) * log(t) - t + log(a);
}

// Log-gamma with the reflection formula below 0.5.  The reference's CPU path
// uses the same approximation (src/rust/src/cpu_native.rs:34-60), so the FP64
// values of hoisted parameter subtrees match the oracle's.
__host__ __device__ inline double lgamma64(double x) {
  if (x < 0.5) {
    const double pi = 3.14159265358979323846;
    return log(fabs(pi / sin(pi * x))) - lgamma64_main(1.0 - x);
  }
  return lgamma64_main(x);
}

// Digamma: the recurrence psi(x) = psi(x + 1) - 1/x moves the argument to
// x >= 6, then the asymptotic series ln x - 1/(2x) - 1/(12 x^2)
// + 1/(120 x^4) - 1/(252 x^6) (Abramowitz & Stegun 6.3.5, 6.3.18).  The
// reference's CPU path takes the same steps (src/rust/src/cpu_native.rs:66-77)
// with no bound on the recurrence; here arguments below -64 first use the
// reflection psi(x) = psi(1 - x) - pi / tan(pi x), so every call ends within
// 70 steps, and -inf gives NaN.
__host__ __device__ inline double digamma64(double x) {
  // Rounded operation by operation, as the reference's (no fused a*b+c).
#pragma clang fp contract(off)
  double result = 0.0;
  if (x < -64.0) {
    const double pi = 3.14159265358979323846;
    if (x == -__builtin_inf()) return __builtin_nan("");
    result = -pi / tan(pi * x);
    x = 1.0 - x;
  }
  while (x < 6.0) {
    result -= 1.0 / x;
    x += 1.0;
  }
  const double f = 1.0 / (x * x);
  const double inner = 1.0 / 120.0 - f * (1.0 / 252.0);
  const double series = 1.0 / 12.0 - f * inner;
  return result + log(x) - 1.0 / (2.0 * x) - f * series;
}

}  // namespace gpum

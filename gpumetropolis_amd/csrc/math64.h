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

}  // namespace gpum

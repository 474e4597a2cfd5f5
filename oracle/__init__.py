"""CPU oracle of the HIP backend: a restatement of the reference's f64 CPU
sampler.  TEST INFRASTRUCTURE ONLY -- imported by tests/, smoke() and the
CPU-baseline leg of bench.py, never by the product package."""

"""MI355X-native (gfx950) sampling backend for gpumetropolis.

The product path: the DSL compiler (``dsl``), the C-ABI binding of the HIP
library (``_native``), the host-side mirror of ``gpu_metropolis`` (``api``),
and the convergence diagnostics used to report ESS/s (``diagnostics``).
"""
from .dsl import DSLError, Model, Program, compile_expr, gpum_model  # noqa: F401

__all__ = ["DSLError", "Model", "Program", "compile_expr", "gpum_model"]

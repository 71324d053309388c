"""MI355X-native nonlocal-density fill and hard-sphere Schur operator for
Tramonto's Newton-Krylov loop (HIP kernels for gfx950 behind a C ABI)."""
from .problem import ProblemSpec, SplitMix64  # noqa: F401

__all__ = ["ProblemSpec", "SplitMix64"]

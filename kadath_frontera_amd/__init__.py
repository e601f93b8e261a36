"""MI355X-native Newton-Raphson path of Kadath/FUKA (spectral transforms,
Jacobian fill, dense LU with mixed-precision refinement) behind a C-ABI.

The shared library must be built (see build.py); nothing here falls back to a
CPU implementation.
"""
from ._lib import lib, LIB_PATH, declared_symbols, KadError  # noqa: F401

__all__ = ["lib", "LIB_PATH", "declared_symbols", "KadError"]

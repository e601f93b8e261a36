"""The C-ABI library loads and exports every symbol include/kadath_amd.h declares (no GPU calls)."""
import ctypes

from kadath_frontera_amd import lib, declared_symbols


def test_header_symbols_exported():
    names = declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_null_context_rejected():
    # argument checks run before any device call
    assert lib.kad_sync(None) < 0
    assert lib.kad_lu_factor(None, 0, 10, None, 10, None, None, 0) < 0
    assert lib.kad_gemm(None, 0, 0, 4, 4, 4, None, 4, None, 4, None, 4) < 0

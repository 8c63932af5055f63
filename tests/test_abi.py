"""The C ABI library loads and exports every entry point include/fastmat_amd.h
declares (CPU-only: no compute call is made)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fastmat_amd.h")
LIB = os.path.join(ROOT, "fastmat_amd", "libfastmat_amd.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "fm_hadamard_apply" in syms and "fm_fourier_apply" in syms
    assert "fm_conv_apply" in syms and "fm_dense_apply" in syms
    assert len(syms) >= 20


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    import torch  # noqa: F401  (bind torch's HIP runtime first, as the package does)
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_python_binding_covers_header():
    from fastmat_amd import _lib
    assert set(declared_symbols()) == set(_lib.EXPORTED)
    lib = _lib.load()
    assert lib.fm_abi_version() == 1


def test_errors_map_to_reference_exceptions():
    import fastmat_amd as fm
    import numpy as np
    with pytest.raises(ValueError):
        fm.Fourier(0)
    with pytest.raises(ValueError):
        fm.Hadamard(0)
    with pytest.raises(ValueError):
        fm.Toeplitz(np.ones(4))  # even length without split
    with pytest.raises(ValueError):
        fm.Product()
    with pytest.raises(TypeError):
        fm.Sum(fm.Eye(3), 3)
    with pytest.raises(ValueError):
        fm.Kron(fm.Eye(3))
    with pytest.raises(ValueError):
        fm.Permutation(np.array([0, 0, 1]))
    with pytest.raises(NotImplementedError):
        fm.Eye(3) // 2
    with pytest.raises(ValueError):
        fm.Partial(fm.Eye(3), rows=np.array([5]))


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_batch_backward_entry_points_reject_a_null_plan():
    """The redirected apply and the least-squares backward validate their
    arguments on the host before any launch (no device is touched)."""
    from fastmat_amd import _lib
    lib = _lib.load()
    assert lib.fm_batch_apply_redirected(None, 1, 0, None, None) != 0
    assert b"NULL" in lib.fm_get_last_error()
    assert lib.fm_batch_lsq_backward(None, 0, None, 0, None) != 0
    assert b"NULL" in lib.fm_get_last_error()

"""Fourier: unnormalised DFT (fastmat/Fourier.pyx:26-285), GPU-backed.

Power-of-two orders run radix FFT kernels; any other order runs Bluestein's
chirp-z algorithm with a power-of-two convolution length L >= 2N - 1 and
FP64-generated chirps (Fourier.pyx:109-156 picks between the two for the
CPU; here the choice is structural, the result y = F x is the same).

Precision follows the input: the matrix dtype is complex64 unless given as
complex128 (``dtype=np.complex128``), and results are promote(input, dtype):
complex64 for int8/int16/float32/complex64 data (FP32 butterflies, FP64
twiddles), complex128 otherwise.  The reference always reports complex128
(Fourier.pyx:159-161); its radix path nevertheless returns complex64 for
complex64 input under numpy 2.
"""
import ctypes

import numpy as np

from . import _device as D
from . import _lib
from .matrix import Matrix


class _PlanCache:
    """Per-precision native plans, created on first use and destroyed with
    the owning matrix."""

    def __init__(self, create, destroy):
        self._create = create
        self._destroy = destroy
        self._plans = {}

    def get(self, prec):
        h = self._plans.get(prec)
        if h is None:
            h = self._plans[prec] = self._create(prec)
        return h

    def drop(self):
        lib = _lib._lib
        for h in self._plans.values():
            if lib is not None and h:
                getattr(lib, self._destroy)(h)
        self._plans = {}


def precision_of(dtype):
    dt = D.np_dtype(dtype)
    if dt == np.complex64:
        return _lib.FM_COMPLEX64
    if dt == np.complex128:
        return _lib.FM_COMPLEX128
    raise TypeError("Fourier-type operators compute in complex64 or complex128, not %s" % dt)


def complex_of(dtype):
    """The complex dtype an operator computes in for a promoted dtype."""
    dt = D.np_dtype(dtype)
    return np.promote_types(dt, np.complex64)


class Fourier(Matrix):
    def __init__(self, order, **options):
        order = int(order)
        if order < 1:
            raise ValueError("Fourier order cannot be smaller than 1.")
        self._order = order
        self._optimize = options.pop("optimize", True)
        self._maxStage = options.pop("maxStage", 4)
        dtype = D.np_dtype(options.pop("dtype", np.complex64))
        if dtype not in (np.dtype(np.complex64), np.dtype(np.complex128)):
            raise TypeError("Fourier dtype must be complex64 or complex128")
        self._content = ()
        self._plans = _PlanCache(self._make_plan, "fm_fourier_plan_destroy")
        self._initProperties(order, order, dtype, **options)

    def _getLargestSingularValue(self):
        # F^H F = N I (Fourier.pyx:188-189): every singular value is sqrt(N)
        return float(np.sqrt(self.numRows))

    def _getColNorms(self):
        # Fourier.pyx:176-177: every column has norm sqrt(N)
        return np.full((self.numCols,), np.sqrt(self.numCols))

    def _getRowNorms(self):
        # Fourier.pyx:179-180
        return np.full((self.numRows,), np.sqrt(self.numRows))

    @property
    def order(self):
        return self._order

    def __del__(self):
        try:
            self._plans.drop()
        except Exception:
            pass

    def _make_plan(self, prec):
        h = ctypes.c_void_p()
        _lib.call("fm_fourier_plan_create", self._order, prec, D.stream(), ctypes.byref(h))
        return h

    def plan_info(self, dtype=np.complex64):
        """(fft_length, bluestein, passes) of the native plan for `dtype`."""
        h = self._plans.get(precision_of(dtype))
        L = ctypes.c_int64()
        b = ctypes.c_int()
        p = ctypes.c_int()
        _lib.call("fm_fourier_plan_info", h, ctypes.byref(L), ctypes.byref(b), ctypes.byref(p))
        return L.value, bool(b.value), p.value

    def compute_dtype(self, in_dtype):
        return complex_of(D.promote(in_dtype, self.dtype))

    def _run(self, X, adjoint):
        cdt = self.compute_dtype(X.dtype)
        if D.np_dtype(X.dtype) != cdt:
            X = X.to(D.torch_dtype(cdt))
        h = self._plans.get(precision_of(cdt))
        k, n = X.shape
        Y = D.empty(k, n, cdt)
        need = ctypes.c_int64()
        _lib.call("fm_fourier_workspace_size", h, k, ctypes.byref(need))
        ws = D.workspace(need.value)
        _lib.call("fm_fourier_apply", h, int(adjoint), _lib.ptr(X), n, _lib.ptr(Y), n, k,
                  None if ws is None else _lib.ptr(ws), need.value, D.stream())
        return Y

    def _apply_views(self, X, Y, adjoint):
        cdt = self.compute_dtype(X.dtype)
        if D.np_dtype(X.dtype) != cdt or D.np_dtype(Y.dtype) != cdt:
            return False
        h = self._plans.get(precision_of(cdt))
        k, n = X.shape
        need = ctypes.c_int64()
        _lib.call("fm_fourier_workspace_size", h, k, ctypes.byref(need))
        ws = D.workspace(need.value)
        _lib.call("fm_fourier_apply", h, int(adjoint), _lib.ptr(X), D.ld(X), _lib.ptr(Y),
                  D.ld(Y), k, None if ws is None else _lib.ptr(ws), need.value, D.stream())
        return True

    def _fwd(self, X):
        return self._run(X, False)

    def _bwd(self, X):
        return self._run(X, True)

    def _apply_strided(self, X, adjoint, outer, inner):
        """Mode-wise DFT for Kron (single-pass power-of-two plans)."""
        h = self._plans.get(precision_of(X.dtype))
        Y = X.new_empty(X.shape)
        _lib.call("fm_fourier_apply_strided", h, int(adjoint), _lib.ptr(X), _lib.ptr(Y), outer,
                  inner, D.stream())
        return Y

    def _strided_ok(self, dtype):
        dt = D.np_dtype(dtype)
        if dt not in (np.dtype(np.complex64), np.dtype(np.complex128)):
            return False
        limit = 8192 if dt == np.complex64 else 4096
        return (self._order & (self._order - 1)) == 0 and self._order <= limit

    # -------------------------------------------------------------- dense
    def _reference(self):
        n = self._order
        idx = np.arange(n, dtype=np.int64)
        return np.exp(np.mod(np.outer(idx, idx), n) * (-2j * np.pi / n))

    def _getArray(self):
        return self._reference().astype(self.dtype)

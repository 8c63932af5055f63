"""Diag: diagonal matrix (fastmat/Diag.pyx), GPU-backed element-wise kernel.
Backward multiplies by conj(d) (Diag.pyx:159-167)."""
import numpy as np

from . import _device as D
from . import _lib
from .matrix import Matrix


class Diag(Matrix):
    def __init__(self, vecD, **options):
        vecD = np.asarray(vecD)
        d = D.squeeze1d(vecD)
        if vecD.ndim != 1 or d.ndim != 1:
            raise ValueError("Diag: Definition vector must have exactly one dimension.")
        if d.dtype not in D.SUPPORTED:
            raise TypeError("Unsupported data type %s" % d.dtype)
        self._vecD = d
        self._dev = {}
        self._content = ()
        self._initProperties(d.size, d.size, d.dtype, **options)
        self._forceContiguousInput = True

    def _getLargestSingularValue(self):
        return float(np.max(np.abs(self._vecD))) if self.numRows else 0.0

    def _getColNorms(self):
        # Diag.pyx:126-127
        return np.abs(self._vecD)

    def _getRowNorms(self):
        # Diag.pyx:129-130
        return np.abs(self._vecD)

    @property
    def vecD(self):
        return self._vecD

    def _d(self, dt):
        t = self._dev.get(dt)
        if t is None:
            t = self._dev[dt] = D.as_tensor(self._vecD, dt)
        return t

    def _run(self, X, conj_d):
        dt = D.promote(X.dtype, self.dtype)
        if D.np_dtype(X.dtype) != dt:
            X = X.to(D.torch_dtype(dt))
        k, n = X.shape
        Y = D.empty(k, n, dt)
        d = self._d(dt)
        _lib.call("fm_diag_apply", D.fm_type(dt), int(conj_d), n, _lib.ptr(d), _lib.ptr(X), n,
                  _lib.ptr(Y), n, k, D.stream())
        return Y

    def _apply_views(self, X, Y, adjoint):
        dt = D.promote(X.dtype, self.dtype)
        if D.np_dtype(X.dtype) != dt or D.np_dtype(Y.dtype) != dt:
            return False
        k, n = X.shape
        _lib.call("fm_diag_apply", D.fm_type(dt), int(adjoint), n, _lib.ptr(self._d(dt)),
                  _lib.ptr(X), D.ld(X), _lib.ptr(Y), D.ld(Y), k, D.stream())
        return True

    def _fwd(self, X):
        return self._run(X, False)

    def _bwd(self, X):
        return self._run(X, True)

    def _getT(self):
        return self

    def _reference(self):
        return np.diag(self._vecD)

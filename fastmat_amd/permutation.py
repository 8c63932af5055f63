"""Permutation: y = x[sigma], backward y = x[argsort(sigma)]
(fastmat/Permutation.pyx:62-125); exact row gather on the device."""
import numpy as np

from . import _device as D
from . import _lib
from .matrix import Matrix


class Permutation(Matrix):
    def __init__(self, sigma, **options):
        sigma = np.asarray(sigma)
        s = D.squeeze1d(sigma)
        if sigma.ndim != 1 or s.ndim != 1:
            raise ValueError("Diag: Definition vector must have exactly one dimension.")
        n = sigma.shape[0]
        if not np.allclose(np.sort(s), np.arange(n)):
            raise ValueError("Not a permutation.")
        self._sigma = s.astype(np.int64)
        self._tau = np.argsort(self._sigma).astype(np.int64)
        self._dev = {}
        self._content = ()
        self._initProperties(n, n, np.int8, **options)

    def _getLargestSingularValue(self):
        return 1.0

    def _getColNorms(self):
        # Permutation.pyx:104-105
        return np.ones((self.numCols,), dtype=self.dtype)

    def _getRowNorms(self):
        # Permutation.pyx:107-108
        return np.ones((self.numRows,), dtype=self.dtype)

    @property
    def sigma(self):
        return self._sigma

    @property
    def tau(self):
        return self._tau

    def _idx(self, which):
        t = self._dev.get(which)
        if t is None:
            t = self._dev[which] = D.as_tensor(self._sigma if which == "s" else self._tau,
                                               np.int64)
        return t

    def _run(self, X, which):
        dt = D.promote(X.dtype, self.dtype)
        if D.np_dtype(X.dtype) != dt:
            X = X.to(D.torch_dtype(dt))
        k, n = X.shape
        Y = D.empty(k, n, dt)
        _lib.call("fm_gather_rows", D.fm_type(dt), n, _lib.ptr(self._idx(which)), _lib.ptr(X), n,
                  _lib.ptr(Y), n, k, D.stream())
        return Y

    def _fwd(self, X):
        return self._run(X, "s")

    def _bwd(self, X):
        return self._run(X, "t")

    def _reference(self):
        return np.eye(self.numRows, dtype=self.dtype)[self._sigma, :]

"""LFSRCirculant: the circulant matrix of a +-1 shift-register sequence
(fastmat/LFSRCirculant.pyx).

The apply follows the reference's algorithm (LFSRCirculant.pyx:316-396):
because every register state is a linear function of the start state over
GF(2), entry (i, j) of the circulant matrix is a Walsh-Hadamard entry
H[tap_i, state_j] of the order-n Hadamard matrix.  So the product is a
scatter of the input rows to the register states (zeros elsewhere), one
Hadamard transform of length 2^n in the data's own type (integer data wraps
like the reference), and a gather of the tap-side states.  The state
sequences come from the library's host routine (fm_lfsr_sequence); the
scatter, transform and gather run in one library call (fm_lfsr_apply) that
moves whole rows of all columns at once through a row-major work array.
"""
import ctypes
import warnings

import numpy as np

from . import _device as D
from . import _lib
from .hadamard import Hadamard
from .matrix import Matrix


class LFSRCirculant(Matrix):
    """LFSRCirculant(polynomial, start) (LFSRCirculant.pyx:159-237)."""

    def __init__(self, polynomial, start, **options):
        self.polynomial = int(polynomial)
        order, mask = 0, 1
        while (~mask & self.polynomial) > mask:
            mask <<= 1
            order += 1
        mask = 1 << order
        self.order = order
        self.start = int(start) & (mask - 1)
        if self.order > 31 or self.order < 1:
            raise ValueError("Only polynomials of order 1 to 31 are supported.")
        if self.start == 0:
            raise ValueError("Initial state must be non-zero.")
        period = ctypes.c_int64()
        _lib.call("fm_lfsr_sequence", self.polynomial & 0xFFFFFFFF, self.start, self.order,
                  ctypes.byref(period), None, None, 0)
        self.period = int(period.value)
        self._content = (Hadamard(self.order),)
        self._seq = None
        self._vecC = None
        self._states = None
        self._dev = {}
        self._initProperties(self.period, self.period, np.int8, **options)

    # ---------------------------------------------------------- sequences
    def _sequences(self):
        """(states, taps): generator states from `start` and tap-side states
        from 1, one period each (int64)."""
        if self._seq is None:
            n = self.period
            st = np.empty(n, dtype=np.int64)
            tp = np.empty(n, dtype=np.int64)
            per = ctypes.c_int64()
            _lib.call("fm_lfsr_sequence", self.polynomial & 0xFFFFFFFF, self.start, self.order,
                      ctypes.byref(per), st.ctypes.data_as(ctypes.c_void_p),
                      tp.ctypes.data_as(ctypes.c_void_p), n)
            self._seq = (st, tp)
        return self._seq

    @property
    def vecC(self):
        """The +-1 sequence (first column), int8 (LFSRCirculant.pyx:297-314)."""
        if self._vecC is None:
            st, _ = self._sequences()
            self._vecC = (1 - 2 * (st & 1)).astype(np.int8)
        return self._vecC

    @property
    def states(self):
        """Register states over one period, uint32 (LFSRCirculant.pyx:278-295)."""
        if self._states is None:
            self._states = self._sequences()[0].astype(np.uint32)
        return self._states

    @property
    def size(self):
        warnings.warn('size is deprecated. WIll be removed in furure releases.', FutureWarning)
        return self.order

    @property
    def taps(self):
        warnings.warn('taps is deprecated. Use polynomial.', FutureWarning)
        return self.polynomial

    def _index(self, name):
        t = self._dev.get(name)
        if t is None:
            st, tp = self._sequences()
            n = self.period
            neg = (-np.arange(n)) % n
            idx = {"scatter_fwd": st[neg], "scatter_bwd": st,
                   "gather_fwd": tp, "gather_bwd": tp[neg]}[name]
            t = self._dev[name] = D.as_tensor(idx, np.int64)
        return t

    # ---------------------------------------------------------- transform
    def _run(self, X, adjoint):
        """forward: y[i] = H[taps[i], :] . z with z[states[t]] = x[-t mod N];
        backward: z[states[t]] = x[t] and y[-i mod N] = (H z)[taps[i]]
        (LFSRCirculant.pyx:316-407), one fm_lfsr_apply call."""
        k, n = X.shape
        dt = D.np_dtype(X.dtype)
        Y = D.empty(k, n, dt)
        if k == 0:
            return Y
        sc = self._index("scatter_bwd" if adjoint else "scatter_fwd")
        ga = self._index("gather_bwd" if adjoint else "gather_fwd")
        need = 2 * (1 << self.order) * k * dt.itemsize
        ws = D.workspace(need)
        _lib.call("fm_lfsr_apply", D.fm_type(dt), self.order, n, _lib.ptr(sc), _lib.ptr(ga),
                  _lib.ptr(X), n, _lib.ptr(Y), n, k, _lib.ptr(ws), need, D.stream())
        return Y

    def _fwd(self, X):
        return self._run(X, False)

    def _bwd(self, X):
        return self._run(X, True)

    # ---------------------------------------------------------- properties
    def _getColNorms(self):
        # LFSRCirculant.pyx:250-251
        return np.full((self.numCols,), np.sqrt(self.numCols))

    def _getRowNorms(self):
        # LFSRCirculant.pyx:253-254
        return np.full((self.numRows,), np.sqrt(self.numRows))

    def _getLargestSingularValue(self):
        # a circulant matrix is normal: sigma_max = max |DFT(first column)|
        return float(np.max(np.abs(np.fft.fft(self.vecC.astype(np.float64)))))

    def _getArray(self):
        return self._reference()

    def _reference(self):
        """Columns are the sequence rolled by the column index
        (LFSRCirculant.pyx:409-438)."""
        seq = self.vecC
        n = self.period
        i = np.arange(n)
        return seq[(i[:, None] - i[None, :]) % n].astype(self.dtype)

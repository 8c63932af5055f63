"""Sum: A1 + A2 + ... (fastmat/Sum.pyx): sum of the terms' transforms
(Sum.pyx:151-175), nested Sums flatten (:73-84), input widened to the sum's
dtype (:120-121)."""
import numpy as np

from . import _device as D
from .matrix import Matrix, accumulate_block


class Sum(Matrix):
    def __init__(self, *matrices, **options):
        terms = []

        def add(items):
            for m in items:
                if not isinstance(m, Matrix):
                    raise TypeError("Sum: Term is not a fastmat Matrix.")
                if isinstance(m, Sum):
                    add(m.content)
                else:
                    terms.append(m)

        add(matrices)
        if len(terms) < 1:
            raise ValueError("Sum: No terms given.")
        numRows, numCols = terms[0].numRows, terms[0].numCols
        for m in terms[1:]:
            if m.numRows != numRows or m.numCols != numCols:
                raise ValueError("Sum: Term dimension mismatch: " + repr(m))
        dt = np.dtype(np.int8)
        for m in terms:
            dt = np.promote_types(dt, m.dtype)
        self._content = tuple(terms)
        self._initProperties(numRows, numCols, dt, **options)
        self._widenInputDatatype = True

    def _run(self, X, adjoint):
        out_dt = D.promote(X.dtype, self.dtype)
        Y = None
        for m in self._content:
            T = m._apply(X, adjoint)
            if D.np_dtype(T.dtype) != out_dt:
                T = T.to(D.torch_dtype(out_dt))
            if Y is None:
                Y = T.clone() if T is X else T
            else:
                accumulate_block(Y, T)
        return Y

    def _fwd(self, X):
        return self._run(X, False)

    def _bwd(self, X):
        return self._run(X, True)

    def _reference(self):
        res = np.zeros((self.numRows, self.numCols), dtype=self.dtype)
        for m in self._content:
            res = res + m.reference()
        return res

"""Product: s * A1 A2 ... An (fastmat/Product.pyx).

Forward applies the last factor first, backward the first factor's adjoint
first with conj(s) (Product.pyx:205-240); nested Products flatten and their
scalars multiply (:101-105); dtype = promote(factors, scalars) expanded by
``typeExpansion`` (default safeTypeExpansion, :120-122)."""
import numpy as np

from . import _device as D
from .matrix import Matrix, scale_block


class Product(Matrix):
    def __init__(self, *matrices, **options):
        factors = []
        scalar = [np.int8(1)]
        dtype = [np.dtype(np.int8)]

        def add(items):
            for f in items:
                if np.isscalar(f):
                    dtype[0] = np.promote_types(dtype[0], np.array(f).dtype)
                    scalar[0] = np.inner(scalar[0], f)
                    continue
                if not isinstance(f, Matrix):
                    raise TypeError("Product: Term not scalar nor Matrix.")
                if isinstance(f, Product):
                    if f._scalar != 1:
                        scalar[0] = scalar[0] * f._scalar
                    add(f.content)
                else:
                    factors.append(f)
                dtype[0] = np.promote_types(dtype[0], f.dtype)

        add(matrices)
        dt = dtype[0]
        typeExpansion = options.pop("typeExpansion", D.safe_type_expansion(dt))
        dt = dt if typeExpansion is None else np.promote_types(dt, typeExpansion)
        options.pop("debug", None)
        if len(factors) < 1:
            raise ValueError("Product has no terms.")
        numRows, numCols = factors[0].numRows, factors[0].numCols
        for ii in range(1, len(factors)):
            if factors[ii].numRows != numCols:
                raise ValueError("Product: Dimension mismatch for term %d [%dx%d]" % (
                    ii, numRows, numCols))
            numCols = factors[ii].numCols
        self._scalar = np.asarray(scalar[0]).astype(dt)
        self._content = tuple(factors)
        self._initProperties(numRows, numCols, dt, **options)

    @property
    def scalar(self):
        return self._scalar

    def _prep(self, X, conj):
        dt = D.promote(X.dtype, self.dtype)
        if D.np_dtype(X.dtype) != dt:
            X = X.to(D.torch_dtype(dt))
        if self._scalar != 1:
            s = complex(self._scalar)
            if conj:
                s = s.conjugate()
            X = scale_block(X, s)
        return X

    def _fwd(self, X):
        X = self._prep(X, False)
        for f in reversed(self._content):
            X = f._apply(X, False)
        return X

    def _bwd(self, X):
        X = self._prep(X, np.iscomplexobj(self._scalar))
        for f in self._content:
            X = f._apply(X, True)
        return X

    def _reference(self):
        dt = np.promote_types(np.float64, self.dtype)
        res = np.inner(self._content[-1].reference(), self._scalar.astype(dt))
        for f in self._content[-2::-1]:
            res = f.reference().dot(res)
        return res

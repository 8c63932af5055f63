"""Kron: Kronecker product of square factors, applied mode-wise
(fastmat/Kron.pyx:79-353).

np.kron convention: factor 1 is the slowest index.  Factor m acts on the axis
of length d_m of x viewed as (k, d_1, ..., d_M) (Kron.pyx:267-303); here each
factor runs its own strided device kernel along that axis (Fourier, Circulant,
Hadamard, dense float32 on MFMA) with no materialised transposes, and any
other factor falls back to moving the axis to the end, applying the factor
and moving it back.  dtype = promote(factors) with safe type expansion; the
input is widened to it (Kron.pyx:104-129)."""
import numpy as np
import torch

from . import _device as D
from . import _lib
from .matrix import Matrix


class Kron(Matrix):
    def __init__(self, *matrices, **options):
        if len(matrices) < 2:
            raise ValueError("Kronecker: Product must have at least two terms")
        dt = np.dtype(np.int8)
        numRows = 1
        for f in matrices:
            if not isinstance(f, Matrix):
                raise TypeError("Kronecker: terms must be fastmat matrices")
            if f.numRows != f.numCols:
                raise ValueError("Kronecker: Product terms must be symmetric")
            numRows *= f.numRows
            dt = np.promote_types(dt, f.dtype)
        self._content = tuple(matrices)
        self._dims = tuple(f.numRows for f in matrices)
        typeExpansion = options.pop("typeExpansion", D.safe_type_expansion(dt))
        dt = dt if typeExpansion is None else np.promote_types(dt, typeExpansion)
        self._initProperties(numRows, numRows, dt, **options)
        self._widenInputDatatype = True

    def _getLargestSingularValue(self):
        # singular values of a Kronecker product are the products of the
        # factors' singular values
        out = 1.0
        for f in self._content:
            out *= f.largestSingularValue
        return float(out)

    def _getColNorms(self):
        # Kron.pyx:165-172: the Kronecker product of the factors' norms
        out = self._content[0].colNorms
        for f in self._content[1:]:
            out = np.kron(out, f.colNorms)
        return out

    def _getRowNorms(self):
        # Kron.pyx:174-182
        out = self._content[0].rowNorms
        for f in self._content[1:]:
            out = np.kron(out, f.rowNorms)
        return out

    def _run(self, X, adjoint):
        k = X.shape[0]
        dims = self._dims
        fused = _fused_tail(self._content, X.dtype)
        last = len(self._content) - (2 if fused else 0)
        for m, f in enumerate(self._content[:last]):
            outer = k * int(np.prod(dims[:m], dtype=np.int64))
            inner = int(np.prod(dims[m + 1:], dtype=np.int64))
            X = _apply_mode(f, X, adjoint, outer, dims[m], inner)
        if fused:
            # the last two factors (Circulant(128), dense 128 x 128 float32)
            # in one pass over the data (fm_kron_conv_dense_apply)
            C, A = self._content[-2:]
            Xc = X.contiguous()
            Y = torch.empty_like(Xc)
            outer = k * int(np.prod(dims[:-2], dtype=np.int64))
            _lib.call("fm_kron_conv_dense_apply", C.plan_handle(np.complex64), int(adjoint),
                      _lib.ptr(A._device_f32(transposed=adjoint)), 128, _lib.ptr(Xc),
                      _lib.ptr(Y), outer, D.stream())
            X = Y
        return X

    def _fwd(self, X):
        return self._run(X, False)

    def _bwd(self, X):
        return self._run(X, True)

    def _reference(self):
        res = self._content[0].reference().astype(np.promote_types(self.dtype, np.float64))
        for f in self._content[1:]:
            res = np.kron(res, f.reference())
        return res


def _fused_tail(content, dtype):
    """True when the product ends in a one-level Circulant of 128 points and a
    dense 128 x 128 float32 Matrix and the data is complex64: those two modes
    then run as one pass (the fused kernel's only shape)."""
    from .circulant import Circulant
    if len(content) < 2 or D.np_dtype(dtype) != np.dtype(np.complex64):
        return False
    C, A = content[-2:]
    return (type(C) is Circulant and C._multi is None and C.numRows == 128 and
            C.compute_dtype(dtype) == np.dtype(np.complex64) and
            type(A) is Matrix and A.numRows == 128 and A.numCols == 128 and
            A._array.dtype == np.float32)


def _apply_mode(f, X, adjoint, outer, n, inner):
    """Apply factor f along the middle axis of X viewed as (outer, n, inner)."""
    k, N = X.shape
    dt = f.compute_dtype(X.dtype) if hasattr(f, "compute_dtype") else f._out_dtype(X.dtype)
    strided = getattr(f, "_strided_ok", None)
    if strided is not None and strided(dt):
        Xc = X if D.np_dtype(X.dtype) == dt else X.to(D.torch_dtype(dt))
        Y = f._apply_strided(Xc.contiguous(), adjoint, outer, inner)
        return Y.reshape(k, N)
    # generic path: move the mode axis last, apply as a block, move it back
    T = X.reshape(outer, n, inner).permute(0, 2, 1).reshape(outer * inner, n).contiguous()
    R = f._apply(T, adjoint)
    return R.reshape(outer, inner, n).permute(0, 2, 1).reshape(k, N).contiguous()

"""Matrix base class and its transposition wrappers (GPU-backed).

Mirrors fastmat.Matrix (fastmat/Matrix.pyx):
  forward / backward           Matrix.pyx:1844-1910, 1937-2007
  input preparation, dtypes    Matrix.pyx:1737-1817
  .H / .T / .conj              Matrix.pyx:1209-1269, Hermitian 2224-2300,
                               Conjugate 2326-2412, Transpose 2417-2482
  operator overloads           Matrix.pyx:1663-1734
  reference()                  Matrix.pyx:2010-2057

Operators act on internal device blocks (k, n) (see _device).  A subclass
implements ``_fwd(X)`` / ``_bwd(X)`` on blocks whose dtype has already been
promoted by ``_apply`` following the reference's rules:
  fInternal = promote(input, minType)  ; fOutput = promote(fInternal, dtype)
  fInternal = fOutput if widenInputDatatype.
"""
import numpy as np
import torch

from . import _device as D
from . import _lib


class FastmatFlags:
    """Global flags (Matrix.pyx:43-48).  The dense-bypass strategy needs
    calibration data and is off in the reference by default
    (Matrix.pyx:1406-1409); it is not used here."""

    def __init__(self):
        self.bypassAllow = False
        self.bypassAutoArray = False


flags = FastmatFlags()


class Matrix:
    """Dense matrix; base of all structured operators.

    ``Matrix(arr)`` stores a 2-D array on the device and applies it with the
    library's FP32-MFMA GEMM for float32 matrices (float32 / complex64 data),
    with the real 2m x 2n form for complex64 matrices, and with the generic
    typed GEMM kernel otherwise.
    """

    __array_priority__ = 20.

    def __init__(self, arrMatrix=None, **options):
        if arrMatrix is None:
            raise ValueError("Matrix: data array required")
        arr = np.array(arrMatrix)
        if arr.ndim != 2:
            raise NotImplementedError("Matrix data array must be 2D")
        if arr.dtype not in D.SUPPORTED:
            raise TypeError("Unsupported data type %s" % arr.dtype)
        self._array = arr
        self._content = ()
        self._dev = {}
        self._initProperties(arr.shape[0], arr.shape[1], arr.dtype, **options)

    # ------------------------------------------------------------ set-up
    def _initProperties(self, numRows, numCols, dataType, **options):
        """Matrix.pyx:1572-1610."""
        self.numRows = int(numRows)
        self.numCols = int(numCols)
        self.dtype = D.np_dtype(dataType)
        self._forceContiguousInput = options.get("forceContiguousInput", False)
        self._widenInputDatatype = options.get("widenInputDatatype", False)
        self._fortranStyle = options.get("fortranStyle", True)
        self._minType = D.np_dtype(options.get("minType", np.int8))
        self.bypassAllow = options.get("bypassAllow", flags.bypassAutoArray)
        self.bypassAutoArray = options.get("bypassAutoArray", False)
        self._H = self._T = self._conj = None
        self._arrayCache = None
        if not hasattr(self, "_content"):
            self._content = ()

    def _getProperties(self):
        return {
            "forceContiguousInput": self._forceContiguousInput,
            "widenInputDatatype": self._widenInputDatatype,
            "fortranStyle": self._fortranStyle,
            "minType": self._minType,
            "bypassAllow": self.bypassAllow,
            "bypassAutoArray": self.bypassAutoArray,
        }

    # ------------------------------------------------------------ spectrum
    @property
    def largestSingularValue(self):
        """sigma_max(A) (Matrix.pyx:824-834), computed once and kept."""
        if getattr(self, "_largestSV", None) is None:
            return self.getLargestSingularValue()
        return self._largestSV

    def getLargestSingularValue(self):
        """Matrix.pyx:836-899: the reference runs scipy's svds (ARPACK) on the
        operator.  Here classes with a closed form override
        `_getLargestSingularValue`; the generic form is a power iteration of
        A^H A on the device (one forward and one adjoint apply per step)."""
        result = float(self._getLargestSingularValue())
        if not np.isnan(result):
            self._largestSV = result
        return result

    def _getLargestSingularValue(self):
        if getattr(self, "_array", None) is not None and max(self.shape) <= 4096:
            return float(np.linalg.svd(np.atleast_2d(self._array), compute_uv=False)[0])
        return power_iteration_sv(self)

    @property
    def colNorms(self):
        """Euclidean norm of every column, float64 (Matrix.pyx:1031-1046),
        computed once and kept."""
        if getattr(self, "_colNorms", None) is None:
            return self.getColNorms()
        return self._colNorms

    def getColNorms(self):
        self._colNorms = self._getColNorms()
        return self._colNorms

    def _getColNorms(self):
        """Matrix.pyx:1048-1087: the operator applied to slices of the
        float64 identity, one column norm per output column.  The slices
        are applied on the device, at most 1024 columns and about 2^24
        output elements at a time; classes with a closed form override
        this."""
        n, m = self.numCols, self.numRows
        out = np.empty(n, dtype=np.float64)
        if n == 0:
            return out
        if m == 0:
            out[:] = 0.0
            return out
        s = max(1, min(1024, (1 << 24) // m))
        dev = D.device()
        for i in range(0, n, s):
            j = min(n, i + s)
            E = torch.zeros(j - i, n, dtype=torch.float64, device=dev)
            E[torch.arange(j - i, device=dev), torch.arange(i, j, device=dev)] = 1.0
            F = self._apply(E, False)
            out[i:j] = torch.linalg.vector_norm(F.to(torch.complex128)
                                                if F.is_complex() else F.double(),
                                                dim=1).cpu().numpy()
        return out

    @property
    def rowNorms(self):
        """Euclidean norm of every row, float64 (Matrix.pyx:1091-1148)."""
        if getattr(self, "_rowNorms", None) is None:
            return self.getRowNorms()
        return self._rowNorms

    def getRowNorms(self):
        self._rowNorms = self._getRowNorms()
        return self._rowNorms

    def _getRowNorms(self):
        """The adjoint applied to slices of the float64 identity (row i of A
        has the norm of column i of A^H), on the device; classes with a
        closed form override this."""
        m, n = self.numRows, self.numCols
        out = np.empty(m, dtype=np.float64)
        if m == 0:
            return out
        if n == 0:
            out[:] = 0.0
            return out
        s = max(1, min(1024, (1 << 24) // n))
        dev = D.device()
        for i in range(0, m, s):
            j = min(m, i + s)
            E = torch.zeros(j - i, m, dtype=torch.float64, device=dev)
            E[torch.arange(j - i, device=dev), torch.arange(i, j, device=dev)] = 1.0
            F = self._apply(E, True)
            out[i:j] = torch.linalg.vector_norm(F.to(torch.complex128)
                                                if F.is_complex() else F.double(),
                                                dim=1).cpu().numpy()
        return out

    @property
    def colNormalized(self):
        """A Diag(1 / colNorms) (Matrix.pyx:1151-1178)."""
        if getattr(self, "_colNormalized", None) is None:
            return self.getColNormalized()
        return self._colNormalized

    def getColNormalized(self):
        self._colNormalized = self._getColNormalized()
        return self._colNormalized

    def _getColNormalized(self):
        from .diag import Diag
        norms = np.asarray(self.colNorms)
        if np.any(norms == 0):
            raise ValueError("Normalization: Matrix has zero-norm column.")
        return self * Diag(1. / norms)

    @property
    def rowNormalized(self):
        """Diag(1 / rowNorms) A (Matrix.pyx:1180-1207)."""
        if getattr(self, "_rowNormalized", None) is None:
            return self.getRowNormalized()
        return self._rowNormalized

    def getRowNormalized(self):
        self._rowNormalized = self._getRowNormalized()
        return self._rowNormalized

    def _getRowNormalized(self):
        from .diag import Diag
        norms = np.asarray(self.rowNorms)
        if np.any(norms == 0):
            raise ValueError("Normalization: Matrix has zero-norm row.")
        return Diag(1. / norms) * self

    @property
    def gram(self):
        """A^H A as a fastmat product (Matrix.pyx:1009-1028)."""
        if getattr(self, "_gram", None) is None:
            return self.getGram()
        return self._gram

    def getGram(self):
        self._gram = self._getGram()
        return self._gram

    def _getGram(self):
        from .product import Product
        return Product(self.H, self)

    @property
    def largestSV(self):
        import warnings
        warnings.warn('largestSV is deprecated. Use largestSingularValue.', FutureWarning)
        return self.largestSingularValue

    @property
    def numN(self):
        import warnings
        warnings.warn('numN is deprecated. Use numRows.', FutureWarning)
        return self.numRows

    @property
    def numM(self):
        import warnings
        warnings.warn('numM is deprecated. Use numCols.', FutureWarning)
        return self.numCols

    @property
    def normalized(self):
        import warnings
        warnings.warn('normalized is deprecated. Use colNormalized.', FutureWarning)
        return self.colNormalized

    @property
    def nbytes(self):
        """Bytes held by this instance and its terms: host arrays and device
        buffers (the reference's getMemoryFootprint, Matrix.pyx:317-322)."""
        return _footprint(self, set())

    @property
    def scipyLinearOperator(self):
        """scipy.sparse.linalg.LinearOperator over forward / backward
        (Matrix.pyx:977-1006)."""
        if getattr(self, "_scipyLinearOperator", None) is None:
            return self.getScipyLinearOperator()
        return self._scipyLinearOperator

    def getScipyLinearOperator(self):
        self._scipyLinearOperator = self._getScipyLinearOperator()
        return self._scipyLinearOperator

    def _getScipyLinearOperator(self):
        from scipy.sparse.linalg import LinearOperator
        return LinearOperator(shape=self.shape, matvec=self.forward, rmatvec=self.backward,
                              matmat=self.forward, dtype=self.dtype)

    # ------------------------------------------------------------ basics
    @property
    def shape(self):
        return (self.numRows, self.numCols)

    @property
    def content(self):
        return self._content

    def __len__(self):
        return len(self._content)

    def __iter__(self):
        return iter(self._content)

    def __repr__(self):
        return "<%s[%dx%d]:0x%12x>" % (self.__class__.__name__, self.numRows,
                                       self.numCols, id(self))

    def __copy__(self):
        return self  # matrices are immutable (Matrix.pyx:341-358)

    def __deepcopy__(self, memo):
        return self

    # --------------------------------------------------------- transforms
    def forward(self, arrX):
        """A @ x for 1-D (n,) or 2-D (n, k) input (Matrix.forward)."""
        X, b = D.to_block(arrX, self.numCols)
        return b.finish(self._apply(X, False))

    def backward(self, arrX):
        """A^H @ x (Matrix.backward)."""
        X, b = D.to_block(arrX, self.numRows)
        return b.finish(self._apply(X, True))

    def _out_dtype(self, in_dtype):
        fin = D.promote(in_dtype, self._minType)
        return D.promote(fin, self.dtype)

    def _apply(self, X, adjoint):
        """Type promotion (Matrix._prepareInputArray) then the transform."""
        fin = D.promote(X.dtype, self._minType)
        fout = D.promote(fin, self.dtype)
        if self._widenInputDatatype:
            fin = fout
        if D.np_dtype(X.dtype) != fin:
            X = X.to(D.torch_dtype(fin))
        return self._bwd(X) if adjoint else self._fwd(X)

    def _apply_into(self, X, Y, adjoint):
        """Write the transform of the (k, n) block X into the (k, m) block Y.
        X and Y may be column slices of wider blocks (unit element stride,
        row stride >= width), as the block compositions hand them over; Y has
        the composition's output dtype.  Classes whose kernels take leading
        dimensions override this to read and write the slices in place; the
        generic form applies to a packed copy and copies the result in."""
        if _views_direct(X, Y) and D.np_dtype(Y.dtype) == self._out_dtype(X.dtype) and \
                self._apply_views(X, Y, adjoint):
            return
        T = self._apply(X if X.is_contiguous() else X.contiguous(), adjoint)
        Y.copy_(T)

    def _apply_views(self, X, Y, adjoint):
        """Leading-dimension apply of X into Y (same dtypes as _apply would
        produce); returns False when the class has none."""
        return False

    # dense implementation (subclasses override)
    def _fwd(self, X):
        return _dense_apply(self, X, False)

    def _bwd(self, X):
        return _dense_apply(self, X, True)

    def _strided_ok(self, dtype):
        """True when the mode-wise (Kron) kernel can apply this dense matrix
        to data of `dtype` directly: float32 A on float32 / complex64 data."""
        return (type(self) is Matrix and self._array.dtype == np.float32 and
                D.np_dtype(dtype) in (np.dtype(np.float32), np.dtype(np.complex64)))

    def _device_f32(self, transposed=False):
        """The float32 array on the device, contiguous row-major (or its
        transpose), cached."""
        key = "f32T" if transposed else "f32"
        A_d = self._dev.get(key)
        if A_d is None:
            A = np.asarray(self._array, dtype=np.float32)
            A_d = self._dev[key] = D.as_tensor(np.ascontiguousarray(A.T if transposed else A),
                                               np.float32)
        return A_d

    def _apply_strided(self, X, adjoint, outer, inner):
        A_d = self._device_f32()
        Y = X.new_empty(X.shape)
        n = self.numRows
        _lib.call("fm_dense_apply_strided", D.fm_type(np.float32), D.fm_type(X.dtype),
                  int(adjoint), n, _lib.ptr(A_d), n, _lib.ptr(X), _lib.ptr(Y), outer, inner,
                  D.stream())
        return Y

    def __mul__(self, factor):
        if isinstance(factor, (np.ndarray, torch.Tensor)):
            return self.forward(factor)
        from .product import Product
        return Product(self, factor, typeExpansion=np.int8)

    def __rmul__(self, factor):
        from .product import Product
        if np.isscalar(factor) or isinstance(factor, Matrix):
            return Product(factor, self, typeExpansion=np.int8)
        raise TypeError("Invalid product term for fastmat Matrix.")

    def __add__(self, element):
        from .sum import Sum
        if isinstance(element, Matrix):
            return Sum(self, element)
        raise TypeError("Not an addition of fastmat matrices.")

    def __radd__(self, element):
        from .sum import Sum
        if isinstance(element, Matrix):
            return Sum(self, element)
        raise TypeError("Not an addition of fastmat matrices.")

    def __sub__(self, element):
        from .product import Product
        from .sum import Sum
        if isinstance(element, Matrix):
            return Sum(self, Product(element, np.int8(-1), typeExpansion=np.int8))
        raise TypeError("Not a subtraction of fastmat matrices.")

    def __rsub__(self, element):
        from .product import Product
        from .sum import Sum
        if isinstance(element, Matrix):
            return Sum(element, Product(self, np.int8(-1), typeExpansion=np.int8))
        raise TypeError("Not a subtraction of fastmat matrices.")

    def __truediv__(self, divisor):
        from .product import Product
        if np.isscalar(divisor):
            if divisor != 0:
                return Product(self, 1. / divisor)
            raise ZeroDivisionError("Division by Zero.")
        raise TypeError("Only scalars allowed as divisors.")

    def __floordiv__(self, divisor):
        raise NotImplementedError("Matrix Floor division not allowed.")

    # ------------------------------------------------------ transposition
    @property
    def H(self):
        if self._H is None:
            self._H = self._getH()
        return self._H

    @property
    def T(self):
        if self._T is None:
            self._T = self._getT()
        return self._T

    @property
    def conj(self):
        if self._conj is None:
            self._conj = self._getConj()
        return self._conj

    def getH(self):
        return self.H

    def getT(self):
        return self.T

    def getConj(self):
        return self.conj

    def _getH(self):
        return Hermitian(self)

    def _getT(self):
        return Transpose(self)

    def _getConj(self):
        return getConjugate(self)

    # --------------------------------------------------------- dense views
    @property
    def array(self):
        return self.getArray()

    def getArray(self):
        if self._arrayCache is None:
            self._arrayCache = self._getArray()
        return self._arrayCache

    def _getArray(self):
        if getattr(self, "_array", None) is not None:
            return self._array
        return self.forward(np.eye(self.numCols, dtype=self.dtype))

    def getCol(self, idx):
        e = np.zeros(self.numCols, dtype=self.dtype)
        e[idx] = 1
        return self.forward(e)

    def getCols(self, indices):
        """Columns by index: an int gives getCol, a slice or 1-D index array
        a 2-D array (Matrix.pyx:389-427), all columns in one batched apply."""
        if np.isscalar(indices):
            return self.getCol(indices)
        if getattr(self, "_array", None) is not None:
            return self._array[:, indices]
        idx = (np.arange(*indices.indices(self.numCols)) if isinstance(indices, slice)
               else np.array(indices))
        if idx.ndim > 1:
            raise IndexError("Index array must have at most one dimension.")
        E = np.zeros((self.numCols, idx.size), dtype=self.dtype)
        E[idx, np.arange(idx.size)] = 1
        return self.forward(E)

    def getRows(self, indices):
        """Rows by index (Matrix.pyx:460-499), all rows in one batched
        adjoint apply."""
        if np.isscalar(indices):
            return self.getRow(indices)
        if getattr(self, "_array", None) is not None:
            return self._array[indices, :]
        idx = (np.arange(*indices.indices(self.numRows)) if isinstance(indices, slice)
               else np.array(indices))
        if idx.ndim > 1:
            raise IndexError("Index array must have at most one dimension.")
        E = np.zeros((self.numRows, idx.size), dtype=self.dtype)
        E[idx, np.arange(idx.size)] = 1
        return np.conj(self.backward(E)).T

    def getRow(self, idx):
        e = np.zeros(self.numRows, dtype=self.dtype)
        e[idx] = 1
        return np.conj(self.backward(e))

    def __getitem__(self, key):
        return self.getArray()[key]

    # ----------------------------------------------------------- reference
    def reference(self):
        """Dense matrix built without the fast algorithms (Matrix.pyx:2010-2057)."""
        return self._reference()

    def _reference(self):
        return self._array


def _footprint(obj, seen):
    """Bytes of the numpy arrays and torch tensors an instance holds
    (attributes, dicts and the terms of compositions), each counted once."""
    if id(obj) in seen:
        return 0
    seen.add(id(obj))
    if isinstance(obj, np.ndarray):
        return int(obj.nbytes)
    if isinstance(obj, torch.Tensor):
        return int(obj.numel() * obj.element_size())
    if isinstance(obj, dict):
        return sum(_footprint(v, seen) for v in obj.values())
    if isinstance(obj, (list, tuple)):
        return sum(_footprint(v, seen) for v in obj)
    if isinstance(obj, Matrix):
        return sum(_footprint(v, seen) for v in vars(obj).values())
    return 0


def _views_direct(X, Y):
    """X, Y are row-major (k, w) views with unit element stride whose rows
    are at least their width apart (a column slice of a wider block)."""
    return (X.dim() == 2 and Y.dim() == 2 and X.shape[0] == Y.shape[0] and
            X.shape[0] > 0 and X.shape[1] > 0 and Y.shape[1] > 0 and
            X.stride(1) == 1 and Y.stride(1) == 1 and
            D.ld(X) >= X.shape[1] and D.ld(Y) >= Y.shape[1])


def _realified(A):
    """Real 2m x 2n form of a complex m x n matrix, interleaved to act on
    (re, im) storage: [[Re, -Im], [Im, Re]] per element."""
    m, n = A.shape
    R = np.empty((2 * m, 2 * n), dtype=np.float32)
    R[0::2, 0::2] = A.real
    R[0::2, 1::2] = -A.imag
    R[1::2, 0::2] = A.imag
    R[1::2, 1::2] = A.real
    return R


def _dense_apply(M, X, adjoint):
    """Y = A X or A^H X for a dense Matrix on the device, in the promoted
    type of the operands like ndarray.dot (Matrix.pyx:1842, 1935): FP32
    MFMA for float32 / complex64 results (a real matrix multiplies complex
    data as interleaved real columns, a complex one as its real 2m x 2n
    form), the library's typed GEMM otherwise (integers wrapping modulo
    2^bits like numpy's integer dot, float64 / complex128 in FP64)."""
    A = M._array
    m, n = A.shape
    fout = D.promote(X.dtype, M.dtype) if not M._widenInputDatatype else D.np_dtype(X.dtype)
    fout = np.dtype(fout)
    k = X.shape[0]
    rows = n if adjoint else m
    if D.np_dtype(X.dtype) != fout:
        X = X.to(D.torch_dtype(fout)).contiguous()
    Y = D.empty(k, rows, fout)
    if fout == np.float32 or (fout == np.complex64 and A.dtype.kind != "c"):
        A_d = M._dev.get("f32")
        if A_d is None:
            A_d = M._dev["f32"] = D.as_tensor(A.real if A.dtype.kind == "c" else A, np.float32)
        _lib.call("fm_dense_apply", D.fm_type(np.float32), D.fm_type(fout), int(adjoint), m, n,
                  _lib.ptr(A_d), n, _lib.ptr(X), X.shape[1], _lib.ptr(Y), rows, k, D.stream())
        return Y
    if fout == np.complex64:
        key = "re_h" if adjoint else "re"
        A_d = M._dev.get(key)
        if A_d is None:
            Ac = A.astype(np.complex64)
            A_d = M._dev[key] = D.as_tensor(_realified(Ac.conj().T if adjoint else Ac),
                                            np.float32)
        mm, nn = (n, m) if adjoint else (m, n)
        _lib.call("fm_dense_apply", D.fm_type(np.float32), D.fm_type(np.float32), 0, 2 * mm,
                  2 * nn, _lib.ptr(A_d), 2 * nn, _lib.ptr(X), 2 * X.shape[1], _lib.ptr(Y),
                  2 * rows, k, D.stream())
        return Y
    key = ("t", fout)
    A_d = M._dev.get(key)
    if A_d is None:
        A_d = M._dev[key] = D.as_tensor(A, fout)
    _lib.call("fm_dense_apply", D.fm_type(fout), D.fm_type(fout), int(adjoint), m, n,
              _lib.ptr(A_d), n, _lib.ptr(X), X.shape[1], _lib.ptr(Y), rows, k, D.stream())
    return Y


# ------------------------------------------------------------------ wrappers
class Hermitian(Matrix):
    """A^H (Matrix.pyx:2224-2300): forward = nested backward."""

    def __init__(self, matrix):
        if not isinstance(matrix, Matrix):
            raise TypeError("Hermitian: Not a fastmat Matrix")
        self._nested = matrix
        self._content = (matrix,)
        self._initProperties(matrix.numCols, matrix.numRows, matrix.dtype,
                             **matrix._getProperties())

    def __repr__(self):
        return "<%s.H>" % (self._nested.__repr__(),)

    def _apply(self, X, adjoint):
        return self._nested._apply(X, not adjoint)

    def _getH(self):
        return self._nested

    def _getT(self):
        return getConjugate(self._nested)

    def _getConj(self):
        return Transpose(self._nested)

    def _getArray(self):
        return self._nested.getArray().T.conj()

    def _reference(self):
        return self._nested.reference().T.conj()


def getConjugate(matrix):
    """Conjugate(matrix) for complex matrices, else the matrix itself."""
    return Conjugate(matrix) if np.iscomplexobj(np.empty(0, matrix.dtype)) else matrix


class Conjugate(Matrix):
    """conj(A) (Matrix.pyx:2326-2412): conj(A conj(x))."""

    def __init__(self, matrix):
        if not isinstance(matrix, Matrix):
            raise TypeError("Conjugate: Not a fastmat Matrix")
        self._nested = matrix
        self._content = (matrix,)
        self._initProperties(matrix.numRows, matrix.numCols, matrix.dtype,
                             **matrix._getProperties())

    def __repr__(self):
        return "<conj(%s)>" % (self._nested.__repr__(),)

    def _apply(self, X, adjoint):
        Xc = _conj_block(X)
        return _conj_block(self._nested._apply(Xc, adjoint))

    def _getH(self):
        return Transpose(self._nested)

    def _getT(self):
        return Hermitian(self._nested)

    def _getConj(self):
        return self._nested

    def _getArray(self):
        return self._nested.getArray().conj()

    def _reference(self):
        return self._nested.reference().conj()


class Transpose(Hermitian):
    """A^T = Hermitian(Conjugate(A)) (Matrix.pyx:2417-2482)."""

    def __init__(self, matrix):
        if not isinstance(matrix, Matrix):
            raise TypeError("Transpose: Not a fastmat Matrix")
        self._nestedConj = matrix
        nested = getConjugate(matrix)
        self._nested = nested
        self._content = (nested,)
        self._initProperties(nested.numCols, nested.numRows, nested.dtype,
                             **nested._getProperties())

    def __repr__(self):
        return "<%s.T>" % (self._nestedConj.__repr__(),)

    def _getT(self):
        return self._nestedConj

    def _getH(self):
        return getConjugate(self._nestedConj)

    def _getConj(self):
        return Hermitian(self._nestedConj)

    def _getArray(self):
        return self._nestedConj.getArray().T

    def _reference(self):
        return self._nestedConj.reference().T


def _conj_block(X):
    """Conjugated copy of a complex block through the library (axpby with
    conj); real blocks are returned as they are."""
    if not X.is_complex():
        return X
    Y = torch.empty_like(X)
    k, n = X.shape
    _lib.call("fm_axpby", D.fm_type(X.dtype), n, k, _lib.dvec(1.0), _lib.ptr(X), n, 1,
              _lib.dvec(0.0), _lib.ptr(Y), n, D.stream())
    return Y


def scale_block(X, s, conj=False, out=None):
    """out = s * op(X) through the library."""
    Y = torch.empty_like(X) if out is None else out
    k, n = X.shape
    _lib.call("fm_axpby", D.fm_type(X.dtype), n, k, _lib.dvec(s), _lib.ptr(X), n, int(conj),
              _lib.dvec(0.0), _lib.ptr(Y), n, D.stream())
    return Y


def accumulate_block(Y, X):
    """Y += X through the library."""
    k, n = X.shape
    _lib.call("fm_axpby", D.fm_type(X.dtype), n, k, _lib.dvec(1.0), _lib.ptr(X), n, 0,
              _lib.dvec(1.0), _lib.ptr(Y), n, D.stream())
    return Y


def power_iteration_sv(M, tol=1e-13, max_steps=5000, seed=20261202):
    """sigma_max of M by power iteration of M^H M on the device: x <- M^H M x
    / ||M^H M x||, estimate = the Rayleigh quotient x^H M^H M x, until it
    changes by < tol (relative).  Works in complex128 (or float64 for real
    operators) so the estimate reaches double precision: the solvers' step
    size 1 / sigma^2 then matches an SVD-based one to ~1e-13."""
    wdt = np.complex128 if np.iscomplexobj(np.zeros(1, M.dtype)) else np.float64
    g = torch.Generator(device="cpu").manual_seed(seed)
    x = torch.randn(1, M.numCols, generator=g, dtype=torch.float64)
    if wdt == np.complex128:
        x = torch.complex(x, torch.randn(1, M.numCols, generator=g, dtype=torch.float64))
    x = x.to(D.device())
    x = x / torch.linalg.vector_norm(x)
    est = 0.0
    for _ in range(max_steps):
        z = M._apply(M._apply(x, False), True)
        if D.np_dtype(z.dtype) != np.dtype(wdt):
            z = z.to(D.torch_dtype(wdt))
        rq = float(torch.sum(x.conj() * z).real)  # x normalised
        nz = float(torch.linalg.vector_norm(z))
        if nz == 0.0:
            return 0.0
        x = z / nz
        if abs(rq - est) <= tol * abs(rq):
            est = rq
            break
        est = rq
    return float(np.sqrt(max(est, 0.0)))

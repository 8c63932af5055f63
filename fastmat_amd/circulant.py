"""Circulant and Toeplitz matrices via circulant embedding, GPU-backed.

Reference: fastmat/Circulant.pyx (1-D branch 104-137, reference 414-425) and
fastmat/Toeplitz.pyx (constructor 146-322, embedding 357-366, reference
719-732).  The reference builds ``Partial(Product(F.H, Diag(fft(t')/L), F))``;
here one native plan holds the spectrum fft(t')/L (computed on the device in
FP64) and applies pad -> FFT -> spectrum -> inverse FFT -> crop in one fused
kernel (L up to the LDS tile) or three passes (larger L).  Embedding lengths
are powers of two: any L >= 2N-1 (Circulant) or L >= m+n-1 (Toeplitz) embeds
the matrix exactly, so y = A x is unchanged.

dtype: promote(generator dtype, complex64) unless ``dtype=`` is given; the
computation precision follows promote(input, dtype) like ``Fourier``.  The
reference reports complex128 always (Circulant.pyx:223-227,
Toeplitz.pyx:318-322).
"""
import ctypes

import numpy as np

from . import _device as D
from . import _lib
from .fourier import _PlanCache, complex_of, precision_of
from .matrix import Matrix


class _ConvBase(Matrix):
    """Shared native-plan handling of Circulant and Toeplitz."""

    def _setup(self, generator, dtype, options):
        gen = np.asarray(generator)
        if dtype is None:
            dtype = np.promote_types(gen.dtype if gen.dtype.kind in "biufc" else np.float64,
                                     np.complex64)
        dtype = D.np_dtype(dtype)
        if dtype not in (np.dtype(np.complex64), np.dtype(np.complex128)):
            raise TypeError("dtype must be complex64 or complex128")
        self._gen128 = gen.astype(np.complex128)
        self._plans = _PlanCache(self._make_plan, "fm_conv_plan_destroy")
        self._content_cache = None
        self._multi = None
        self._multi_chains = {}
        return dtype

    def __del__(self):
        try:
            if getattr(self, "_gplans", None) is not None:
                self._gplans.drop()
            self._plans.drop()
        except Exception:
            pass

    def plan_info(self, dtype=np.complex64):
        """(rows, cols, fft_length, passes) of the native plan."""
        h = self._plans.get(precision_of(dtype))
        r, c, L = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        p = ctypes.c_int()
        _lib.call("fm_conv_plan_info", h, ctypes.byref(r), ctypes.byref(c), ctypes.byref(L),
                  ctypes.byref(p))
        return r.value, c.value, L.value, p.value

    def plan_handle(self, dtype):
        return self._plans.get(precision_of(dtype))

    def compute_dtype(self, in_dtype):
        return complex_of(D.promote(in_dtype, self.dtype))

    def _run(self, X, adjoint):
        cdt = self.compute_dtype(X.dtype)
        if self._multi is not None:
            if D.np_dtype(X.dtype) != cdt:
                X = X.to(D.torch_dtype(cdt))
            return self._multilevel_chain(cdt)._apply(X, adjoint)
        if D.np_dtype(X.dtype) != cdt:
            X = X.to(D.torch_dtype(cdt))
        return self._run_plan(self._plans.get(precision_of(cdt)), X, adjoint)

    def _run_plan(self, h, X, adjoint):
        """Apply the native plan `h` to the (k, n) block X of its precision."""
        k, n = X.shape
        m = self.numCols if adjoint else self.numRows
        Y = D.empty(k, m, X.dtype)
        need = ctypes.c_int64()
        _lib.call("fm_conv_workspace_size", h, k, ctypes.byref(need))
        ws = D.workspace(need.value)
        _lib.call("fm_conv_apply", h, int(adjoint), _lib.ptr(X), n, _lib.ptr(Y), m, k,
                  None if ws is None else _lib.ptr(ws), need.value, D.stream())
        return Y

    def _apply_views(self, X, Y, adjoint):
        cdt = self.compute_dtype(X.dtype)
        if self._multi is not None or D.np_dtype(X.dtype) != cdt or D.np_dtype(Y.dtype) != cdt:
            return False
        h = self._plans.get(precision_of(cdt))
        k = X.shape[0]
        need = ctypes.c_int64()
        _lib.call("fm_conv_workspace_size", h, k, ctypes.byref(need))
        ws = D.workspace(need.value)
        _lib.call("fm_conv_apply", h, int(adjoint), _lib.ptr(X), D.ld(X), _lib.ptr(Y),
                  D.ld(Y), k, None if ws is None else _lib.ptr(ws), need.value, D.stream())
        return True

    def _getArray(self):
        # the dense array is the closed form (Circulant.pyx:241-242,
        # Toeplitz.pyx:325-326)
        return self._reference()

    def _fwd(self, X):
        return self._run(X, False)

    def _bwd(self, X):
        return self._run(X, True)

    def _generator_grad_1d(self, X, G):
        """Gradient of a real loss w.r.t. the generator (c, or [vecC, vecR]),
        given the forward input block X (k, cols) and dLoss/dY block G
        (k, rows), PyTorch convention.  Returns a device tensor."""
        cdt = self.compute_dtype(X.dtype)
        X = X.to(D.torch_dtype(cdt)).contiguous()
        G = G.to(D.torch_dtype(cdt)).contiguous()
        h = self._plans.get(precision_of(cdt))
        k = X.shape[0]
        need = ctypes.c_int64()
        _lib.call("fm_conv_generator_grad_workspace_size", h, k, ctypes.byref(need))
        ws = D.workspace(need.value)
        out = D.empty(1, self._gen128.size, cdt)[0]
        _lib.call("fm_conv_generator_grad", h, _lib.ptr(X), X.shape[1], _lib.ptr(G), G.shape[1],
                  k, _lib.ptr(out), _lib.ptr(ws), need.value, D.stream())
        return out

    def _multilevel_chain(self, cdt):
        """Partial(Product(K.H, Diag(fftn(t')/prod L), K)), K = Kron(Fourier(L_i)),
        for a multi-level generator (Circulant.pyx:138-215,
        Toeplitz.pyx:261-316), with the spectrum rounded once to the compute
        precision.  The per-level lengths L_i are powers of two."""
        cdt = np.dtype(cdt)
        chain = self._multi_chains.get(cdt)
        if chain is None:
            from .diag import Diag
            from .fourier import Fourier
            from .kron import Kron
            from .partial import Partial
            from .product import Product
            Ls, t, rows, cols = self._multi
            Lt = int(np.prod(Ls))
            spec = (np.fft.fftn(t) / Lt).reshape(-1)
            if _two_level_ok(Ls, cdt):
                P = _TwoLevelConv(Ls, spec, cdt)
            else:
                K = Kron(*[Fourier(L, dtype=cdt) for L in Ls])
                P = Product(K.H, Diag(spec.astype(cdt)), K)
            chain = Partial(P, rows=None if len(rows) == Lt else rows,
                            cols=None if len(cols) == Lt else cols)
            self._multi_chains[cdt] = chain
        return chain

    def generator_grad(self, X, G):
        if self._multi is not None:
            raise NotImplementedError("generator gradient of a multi-level matrix")
        return self._generator_grad_1d(X, G)

    # ------------------------------------------- spectrum from a device generator
    def _gen_plan(self, cdt):
        """Private copy of the native plan of precision `cdt` whose spectrum is
        set from device generators (fm_conv_set_generator); the matrix's own
        plan keeps the spectrum of the constructor's generator."""
        if getattr(self, "_gplans", None) is None:
            def clone(prec):
                h = ctypes.c_void_p()
                _lib.call("fm_conv_plan_clone", self._plans.get(prec), D.stream(),
                          ctypes.byref(h))
                return h
            self._gplans = _PlanCache(clone, "fm_conv_plan_destroy")
            self._gtokens = {}
        return self._gplans.get(precision_of(cdt))

    def load_generator(self, gen, cdt):
        """Make the private plan of precision `cdt` the operator of the device
        generator `gen` (c, or [vecC, vecR], any real or complex float type):
        its spectrum becomes fft(t')/L of gen's current values (the operator
        *is* its generator's spectrum, Circulant.pyx:130, Toeplitz.pyx:253-255).
        Returns (plan handle, token); the token changes with every load."""
        if self._multi is not None:
            raise NotImplementedError("device generators of a multi-level matrix")
        if gen.dim() != 1 or gen.numel() != self._gen128.size:
            raise ValueError("generator must be a vector of %d elements" % self._gen128.size)
        if D.np_dtype(gen.dtype) not in (np.dtype(np.float32), np.dtype(np.float64),
                                         np.dtype(np.complex64), np.dtype(np.complex128)):
            gen = gen.to(D.torch_dtype(cdt))
        gen = gen.detach().contiguous()
        h = self._gen_plan(cdt)
        need = ctypes.c_int64()
        _lib.call("fm_conv_set_generator_workspace_size", h, ctypes.byref(need))
        ws = D.workspace(need.value)
        _lib.call("fm_conv_set_generator", h, D.fm_type(gen.dtype), _lib.ptr(gen),
                  _lib.ptr(ws), need.value, D.stream())
        prec = precision_of(cdt)
        tok = self._gtokens.get(prec, 0) + 1
        self._gtokens[prec] = tok
        return h, tok

    def generator_token(self, cdt):
        """Token of the generator last loaded into the private plan of `cdt`."""
        return getattr(self, "_gtokens", {}).get(precision_of(cdt), 0)

    def _strided_ok(self, dtype):
        if self._multi is not None:
            return False
        dt = D.np_dtype(dtype)
        n = self.numRows
        limit = 8192 if dt == np.complex64 else 4096
        return (dt in (np.dtype(np.complex64), np.dtype(np.complex128)) and
                self.numRows == self.numCols and (n & (n - 1)) == 0 and n <= limit)

    def _apply_strided(self, X, adjoint, outer, inner):
        h = self._plans.get(precision_of(X.dtype))
        Y = X.new_empty(X.shape)
        _lib.call("fm_conv_apply_strided", h, int(adjoint), _lib.ptr(X), _lib.ptr(Y), outer,
                  inner, D.stream())
        return Y


def _next_pow2(n):
    return 1 if n <= 1 else 1 << (int(n) - 1).bit_length()


class Circulant(_ConvBase):
    """C[i, j] = c[(i - j) mod N] (1-D generator)."""

    def __init__(self, tenC, **options):
        tenC = np.asarray(tenC)
        c = D.squeeze1d(tenC)
        if tenC.ndim < 1 or c.ndim < 1:
            raise ValueError("Column-definition tensor must be at least 1D.")
        self._optimize = options.pop("optimize", True)
        self._maxStage = options.pop("maxStage", 4)
        dtype = self._setup(c, options.pop("dtype", None), options)
        self._tenC = c
        n = c.size
        self._content = ()
        if c.ndim > 1:
            # per level: keep a power-of-two size, else embed into the next
            # power of two >= 2 d - 1 as [t, 0 ... 0, t[1:]] along that axis
            # (Circulant._preProcSlice, Circulant.pyx:243-280)
            dims = c.shape
            Ls = [d if (d & (d - 1)) == 0 else _next_pow2(2 * d - 1) for d in dims]
            t = c.astype(np.complex128)
            for ax, (d, L) in enumerate(zip(dims, Ls)):
                if L != d:
                    t = np.concatenate([t, np.zeros(t.shape[:ax] + (L - 2 * d + 1,) +
                                                    t.shape[ax + 1:]),
                                        np.take(t, np.arange(1, d), axis=ax)], axis=ax)
            sel = _grid_selection(dims, Ls)
            self._multi = (Ls, t, sel, sel)
        self._initProperties(n, n, dtype, **options)

    def _getLargestSingularValue(self):
        # a (multi-level) circulant matrix is normal and diagonalised by the
        # (multi-dimensional) DFT: sigma_max = max |fftn(c)|
        return float(np.max(np.abs(np.fft.fftn(self._tenC.astype(np.complex128)))))

    def _getColNorms(self):
        # Circulant.pyx:229-230: every column holds all of tenC
        return np.full((self._tenC.size,), np.linalg.norm(self._tenC))

    def _getRowNorms(self):
        # Circulant.pyx:232-233
        return np.full((self._tenC.size,), np.linalg.norm(self._tenC))

    @property
    def tenC(self):
        return self._tenC

    @property
    def fftLength(self):
        if self._multi is not None:
            return int(np.prod(self._multi[0]))
        n = self._tenC.size
        return n if (n & (n - 1)) == 0 else _next_pow2(2 * n - 1)

    def _make_plan(self, prec):
        gen = D.as_tensor(self._gen128, np.complex128)
        h = ctypes.c_void_p()
        _lib.call("fm_circulant_plan_create", prec, self._gen128.size, _lib.ptr(gen), D.stream(),
                  ctypes.byref(h))
        return h

    @property
    def content(self):
        """The equivalent Partial(Product(F.H, Diag(fft(c')/L), F)) chain
        (Circulant.pyx:127-137, 217-221), GPU-backed."""
        if self._multi is not None:
            return (self._multilevel_chain(self.dtype),)
        if self._content_cache is None:
            from .diag import Diag
            from .fourier import Fourier
            from .partial import Partial
            from .product import Product
            n, L = self._tenC.size, self.fftLength
            c = self._gen128
            t = c if L == n else np.concatenate([c, np.zeros(L - 2 * n + 1), c[1:]])
            F = Fourier(L, dtype=self.dtype)
            P = Product(F.H, Diag((np.fft.fft(t) / L).astype(self.dtype)), F)
            sel = np.arange(n) if L != n else None
            self._content_cache = (Partial(P, rows=sel, cols=sel),)
        return self._content_cache

    def _reference(self):
        c = self._tenC.astype(self.dtype)
        return _multilevel_dense(c, c.shape, c.shape)


class Toeplitz(_ConvBase):
    """T[i, j] = tenT[(i - j) mod (m + n - 1)], tenT = [vecC (m), vecR (n-1)]."""

    def __init__(self, *args, **options):
        split = options.pop("split", None)
        arrSplit = np.array([] if split is None else split)
        if len(args) == 1:
            # the reference stores it in a typed ndarray slot, which rejects
            # anything else with TypeError (Toeplitz.pyx:147-149)
            if not isinstance(args[0], np.ndarray):
                raise TypeError("Argument 'tenT' has incorrect type (expected "
                                "numpy.ndarray, got %s)" % type(args[0]).__name__)
            tenT = args[0]
        elif len(args) == 2:
            if not all(isinstance(a, np.ndarray) for a in args):
                raise ValueError(
                    "You must specify two 1D-ndarrays containing the " +
                    "column- and row-definition vectors or one ndarray tensor")
            if arrSplit.size != 0:
                raise ValueError(
                    "You must not define split points when supplying " +
                    "column- and row-definition vectors.")
            dt = np.promote_types(args[0].dtype, args[1].dtype)
            vecC = D.squeeze1d(args[0].astype(dt))
            vecR = D.squeeze1d(args[1].astype(dt)) if args[1].size else args[1].reshape(0)
            if vecC.ndim != 1 or vecR.ndim != 1:
                raise ValueError("Column- and row-definition vectors must be 1D.")
            arrSplit = np.array(vecC.size)
            tenT = np.hstack((vecC, vecR))
        else:
            raise ValueError(
                "Invalid number of arguments to Toeplitz: Expecting exactly " +
                "one or two fixed arguments")
        arrDim = np.array(tenT.shape)
        arrSplit = np.atleast_1d(arrSplit)
        if arrSplit.size == 0:
            if not all((ll + 1) % 2 == 0 for ll in arrDim):
                raise ValueError(
                    "Defining a tensor with non-square levels requires " +
                    "explicit split points.")
            arrSplit = (arrDim + 1) // 2
        if arrSplit.size != tenT.ndim:
            raise ValueError("The split point vector must have one entry for each " +
                             "dimension of the defining tensor")
        if any(ll < 1 or ll > arrDim[ii] for ii, ll in enumerate(arrSplit)):
            raise ValueError("Entry in split vector outside of defining tensor bounds")
        self._optimize = options.pop("optimize", True)
        self._maxStage = options.pop("maxStage", 4)
        self._tenT = tenT
        self._arrDimRows = np.array(arrSplit, dtype=np.int64)
        self._arrDimCols = arrDim - self._arrDimRows + 1
        m = int(np.prod(self._arrDimRows))
        n = int(np.prod(self._arrDimCols))
        dtype = self._setup(tenT, options.pop("dtype", None), options)
        self._content = ()
        if tenT.ndim > 1:
            # per level: zeros inserted at the split point up to the next
            # power of two (Toeplitz._preProcSlice, Toeplitz.pyx:331-369)
            dims = tuple(int(d) for d in arrDim)
            Ls = [_next_pow2(d) for d in dims]
            t = tenT.astype(np.complex128)
            for ax, (d, L, sp) in enumerate(zip(dims, Ls, self._arrDimRows)):
                if L != d:
                    t = np.concatenate([np.take(t, np.arange(sp), axis=ax),
                                        np.zeros(t.shape[:ax] + (L - d,) + t.shape[ax + 1:]),
                                        np.take(t, np.arange(sp, d), axis=ax)], axis=ax)
            rows = _grid_selection(tuple(self._arrDimRows), Ls)
            cols = _grid_selection(tuple(self._arrDimCols), Ls)
            self._multi = (Ls, t, rows, cols)
        self._initProperties(m, n, dtype, **options)

    @property
    def tenT(self):
        return self._tenT

    @property
    def vecC(self):
        """tenT up to the first level's split point; deprecated like
        Toeplitz.pyx:63-69."""
        import warnings
        warnings.warn('vecC is deprecated.', FutureWarning)
        return self._tenT[:self._arrDimRows[0]]

    @property
    def vecR(self):
        """The last cols-1 entries of tenT along the first level; deprecated
        like Toeplitz.pyx:71-77 (same slice, so a one-column level returns
        all of tenT there too)."""
        import warnings
        warnings.warn('vecR is deprecated.', FutureWarning)
        return self._tenT[-int(self._arrDimCols[0]) + 1:]

    @property
    def fftLength(self):
        if self._multi is not None:
            return int(np.prod(self._multi[0]))
        return _next_pow2(self.numRows + self.numCols - 1)

    def _make_plan(self, prec):
        gen = D.as_tensor(self._gen128, np.complex128)
        h = ctypes.c_void_p()
        _lib.call("fm_toeplitz_plan_create", prec, self.numRows, self.numCols, _lib.ptr(gen),
                  D.stream(), ctypes.byref(h))
        return h

    def _reference(self):
        return _multilevel_dense(self._tenT.astype(self.dtype), tuple(self._arrDimRows),
                                 tuple(self._arrDimCols))

    def _getColNorms(self):
        # the closed form of Toeplitz._normalizeColCore (Toeplitz.pyx:371-470)
        return _multilevel_col_norms(self._tenT, tuple(self._arrDimRows),
                                     tuple(self._arrDimCols))


def _two_level_ok(Ls, cdt):
    """Two levels that the three-pass two-level plan takes: each 2^5 ..
    2^13 (complex64) or 2^12 (complex128), together beyond one tile."""
    if len(Ls) != 2:
        return False
    single = 8192 if np.dtype(cdt) == np.complex64 else 4096
    top = 13 if np.dtype(cdt) == np.complex64 else 12
    logs = [int(L).bit_length() - 1 for L in Ls]
    return min(logs) >= 5 and max(logs) <= top and Ls[0] * Ls[1] > single


class _TwoLevelConv(Matrix):
    """The (L0 L1) x (L0 L1) two-level circulant with power-of-two levels
    and spectrum fftn(t)/(L0 L1) (C order), applied by the library's
    three-pass two-level plan (fm_circulant2_plan_create): level-0 DFT,
    level-1 DFT / spectrum / inverse, level-0 inverse DFT."""

    def __init__(self, Ls, spectrum, cdt):
        self._Ls = (int(Ls[0]), int(Ls[1]))
        self._content = ()
        n = self._Ls[0] * self._Ls[1]
        self._initProperties(n, n, cdt)
        spec = D.as_tensor(np.asarray(spectrum, dtype=np.complex128), np.complex128)
        h = ctypes.c_void_p()
        _lib.call("fm_circulant2_plan_create", precision_of(cdt), self._Ls[0], self._Ls[1],
                  _lib.ptr(spec), D.stream(), ctypes.byref(h))
        self._h = h

    def __del__(self):
        try:
            if getattr(self, "_h", None) and _lib._lib is not None:
                _lib._lib.fm_conv_plan_destroy(self._h)
        except Exception:
            pass

    def _run(self, X, adjoint):
        cdt = self.dtype
        if D.np_dtype(X.dtype) != cdt:
            X = X.to(D.torch_dtype(cdt))
        k, n = X.shape
        Y = D.empty(k, n, cdt)
        need = ctypes.c_int64()
        _lib.call("fm_conv_workspace_size", self._h, k, ctypes.byref(need))
        ws = D.workspace(need.value)
        _lib.call("fm_conv_apply", self._h, int(adjoint), _lib.ptr(X), n, _lib.ptr(Y), n, k,
                  None if ws is None else _lib.ptr(ws), need.value, D.stream())
        return Y

    def _fwd(self, X):
        return self._run(X, False)

    def _bwd(self, X):
        return self._run(X, True)


def _grid_selection(dims, Ls):
    """Flat indices (C order) of the grid prod(range(d_i)) inside the padded
    grid prod(range(L_i)) (Circulant._genArrS, Circulant.pyx:292-347)."""
    idx = np.zeros(1, dtype=np.int64)
    for d, L in zip(dims, Ls):
        idx = (idx[:, None] * L + np.arange(d)[None, :]).ravel()
    return idx


def _multilevel_dense(t, rows, cols):
    """Dense multi-level Toeplitz/Circulant: entry ((i_1..i_d), (j_1..j_d)),
    multi-indices in C order, is t[(i_1 - j_1) mod d_1, ..., (i_d - j_d) mod
    d_d] with d_k = t.shape[k] (Circulant._refRecursion, Circulant.pyx:
    344-412; Toeplitz._refRecursion, Toeplitz.pyx:628-719)."""
    t = np.asarray(t)
    rows = tuple(int(r) for r in np.atleast_1d(rows))
    cols = tuple(int(c) for c in np.atleast_1d(cols))
    if t.ndim == 1 and len(rows) == 1:
        d = t.size
        i = np.arange(rows[0])[:, None]
        j = np.arange(cols[0])[None, :]
        if d == rows[0] == cols[0]:
            return t[np.mod(i - j, d)]
        return t[np.mod(i - j, rows[0] + cols[0] - 1)]
    nd = len(rows)
    idx = np.meshgrid(*[np.arange(r) for r in rows], indexing="ij")
    jdx = np.meshgrid(*[np.arange(c) for c in cols], indexing="ij")
    I = [a.ravel()[:, None] for a in idx]
    J = [b.ravel()[None, :] for b in jdx]
    sub = tuple(np.mod(I[k] - J[k], t.shape[k]) for k in range(nd))
    return t[sub]


def _multilevel_col_norms(t, rows, cols):
    """Column norms of the matrix of `_multilevel_dense`: along level k,
    column j_k holds the cyclic window of rows_k entries of t starting at
    (-j_k) mod d_k, so the squared norm of column (j_1..j_d) is the sum of
    |t|^2 over the product of those windows.  The window sums are formed one
    level at a time from prefix sums (in extended precision)."""
    a = np.abs(np.asarray(t)).astype(np.longdouble) ** 2
    rows = tuple(int(r) for r in np.atleast_1d(rows))
    cols = tuple(int(c) for c in np.atleast_1d(cols))
    if a.ndim == 1 and len(rows) == 1:
        a = a.reshape(-1)
    for ax, r in enumerate(rows):
        d = a.shape[ax]
        ext = np.concatenate([a, np.take(a, np.arange(r), axis=ax)], axis=ax)
        c = np.cumsum(ext, axis=ax)
        c = np.concatenate([np.zeros_like(np.take(c, [0], axis=ax)), c], axis=ax)
        s = np.arange(d)
        a = np.take(c, s + r, axis=ax) - np.take(c, s, axis=ax)
    idx = [np.mod(-np.arange(c), a.shape[k]) for k, c in enumerate(cols)]
    return np.sqrt(np.maximum(a[np.ix_(*idx)].reshape(-1), 0)).astype(np.float64)

"""Partial: row/column sub-selection of a matrix (fastmat/Partial.pyx).

Forward scatters x into a zero-padded vector of the nested matrix's width at
``colSelection``, applies the matrix and gathers ``rowSelection``
(Partial.pyx:268-280); backward mirrors it (:282-294).  Scatter and gather run
as library kernels."""
import numpy as np

from . import _device as D
from . import _lib
from .matrix import Matrix


def _check_selection(sel, size, name):
    if sel is None:
        return None
    sel = np.array(sel)
    if sel.dtype == bool:
        sel = np.arange(size)[sel]
    elif np.issubdtype(sel.dtype, np.integer):
        pass
    else:
        raise TypeError("Partial: Type of %s indices must be int or bool." % (name,))
    bounded = ((len(sel) != size or np.sum(sel - np.arange(size)) != 0) and
               np.any((sel >= size) | (sel < 0)))
    if bounded:
        raise ValueError("Partial: A %s index exceeds matrix dimensions." % (name,))
    return sel.astype(np.int64)


def _device_index(sel, size):
    """Row indices for the scatter / gather kernels with numpy's indexing
    rules: an index in [-size, -1] counts from the end, anything outside
    [-size, size) raises IndexError when the selection is used (the
    reference's check above lets e.g. [-1, 1, 3] of size 3 through, and
    numpy then raises at apply time, Partial.pyx:272-279)."""
    bad = (sel < -size) | (sel >= size)
    if np.any(bad):
        raise IndexError("index %d is out of bounds for axis 0 with size %d"
                         % (int(sel[np.argmax(bad)]), size))
    return np.where(sel < 0, sel + size, sel).astype(np.int64)


class Partial(Matrix):
    def __init__(self, mat, **options):
        if not isinstance(mat, Matrix):
            raise TypeError("Partial: fastmat Matrix required.")
        if "N" in options:
            options["rows"] = options.pop("N")
        if "M" in options:
            options["cols"] = options.pop("M")
        self._content = (mat,)
        self._rowSelection = _check_selection(options.pop("rows", None), mat.numRows, "row")
        self._colSelection = _check_selection(options.pop("cols", None), mat.numCols, "col")
        self._dev = {}
        self._initProperties(
            len(self._rowSelection) if self._rowSelection is not None else mat.numRows,
            len(self._colSelection) if self._colSelection is not None else mat.numCols,
            mat.dtype, **options)

    @property
    def rowSelection(self):
        return self._rowSelection

    @property
    def colSelection(self):
        return self._colSelection

    def _getColNorms(self):
        # Partial.pyx:234-241: without a row selection the norms are the
        # parent's, restricted to the selected columns
        if self._rowSelection is not None:
            return super()._getColNorms()
        norms = self._content[0].colNorms
        return norms if self._colSelection is None else norms[self._colSelection]

    def _getRowNorms(self):
        # Partial.pyx:243-250
        if self._colSelection is not None:
            return super()._getRowNorms()
        norms = self._content[0].rowNorms
        return norms if self._rowSelection is None else norms[self._rowSelection]

    def _idx(self, name, sel):
        t = self._dev.get(name)
        if t is None:
            M = self._content[0]
            size = M.numRows if name == "r" else M.numCols
            t = self._dev[name] = D.as_tensor(_device_index(sel, size), np.int64)
        return t

    def _scatter(self, X, sel, name, n_full):
        k, n = X.shape
        Y = D.empty(k, n_full, X.dtype)
        _lib.call("fm_scatter_rows", D.fm_type(X.dtype), n, _lib.ptr(self._idx(name, sel)),
                  _lib.ptr(X), n, _lib.ptr(Y), n_full, n_full, k, D.stream())
        return Y

    def _gather(self, X, sel, name):
        k, n = X.shape
        m = len(sel)
        Y = D.empty(k, m, X.dtype)
        _lib.call("fm_gather_rows", D.fm_type(X.dtype), m, _lib.ptr(self._idx(name, sel)),
                  _lib.ptr(X), n, _lib.ptr(Y), m, k, D.stream())
        return Y

    def _fwd(self, X):
        M = self._content[0]
        if self._colSelection is not None:
            X = self._scatter(X, self._colSelection, "c", M.numCols)
        Y = M._apply(X, False)
        if self._rowSelection is not None:
            Y = self._gather(Y, self._rowSelection, "r")
        return Y

    def _bwd(self, X):
        M = self._content[0]
        if self._rowSelection is not None:
            X = self._scatter(X, self._rowSelection, "r", M.numRows)
        Y = M._apply(X, True)
        if self._colSelection is not None:
            Y = self._gather(Y, self._colSelection, "c")
        return Y

    def _reference(self):
        full = self._content[0].reference()
        if self._rowSelection is not None:
            full = full[self._rowSelection, :]
        if self._colSelection is not None:
            full = full[:, self._colSelection]
        return full

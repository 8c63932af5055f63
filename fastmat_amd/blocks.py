"""Block compositions: BlockDiag and Blocks (fastmat/BlockDiag.pyx,
fastmat/Blocks.pyx).

Both work on column slices of the caller's (k, n) block without packing
them: each term reads its slice of the input and writes its slice of the
output in place through the leading-dimension applies (Matrix._apply_into).
The 1-D Circulant / Toeplitz blocks of a BlockDiag are applied together by
one ConvBatch bound to those slices, so they run as a few persistent
launches over the batch descriptor table rather than one transform after
another.  Input is widened to the composition's dtype first, as the
reference does (BlockDiag.pyx:130, Blocks.pyx:195).
"""
import numpy as np
import torch

from . import _device as D
from . import _lib
from .matrix import Matrix


def _conv_1d(term):
    from .circulant import _ConvBase
    return isinstance(term, _ConvBase) and term._multi is None


class BlockDiag(Matrix):
    """diag(A_1, ..., A_n) (BlockDiag.pyx:95-189)."""

    def __init__(self, *matrices, **options):
        if len(matrices) < 1:
            raise ValueError("You must specify at least one matrix.")
        rows = cols = 0
        dt = np.dtype(np.int8)
        for term in matrices:
            if not isinstance(term, Matrix):
                raise ValueError("Only fastmat matrices supported, %s given." % (str(term)))
            rows += term.numRows
            cols += term.numCols
            dt = np.promote_types(dt, term.dtype)
        self._content = tuple(matrices)
        self._batches = {}
        self._initProperties(rows, cols, dt, **options)
        self._widenInputDatatype = True

    def _run(self, X, adjoint):
        k = X.shape[0]
        out_dt = D.np_dtype(X.dtype)
        Y = D.empty(k, self.numCols if adjoint else self.numRows, out_dt)
        if k == 0:
            return Y
        convs, i_in, i_out = [], 0, 0
        for idx, term in enumerate(self._content):
            nin = term.numRows if adjoint else term.numCols
            mout = term.numCols if adjoint else term.numRows
            xs, ys = X[:, i_in:i_in + nin], Y[:, i_out:i_out + mout]
            i_in += nin
            i_out += mout
            if mout == 0:
                continue
            if nin == 0:
                ys.zero_()
                continue
            if _conv_1d(term) and term.compute_dtype(out_dt) == out_dt:
                convs.append((idx, xs, ys))
            else:
                term._apply_into(xs, ys, adjoint)
        # binding a batch to new addresses allocates and synchronises, which a
        # stream under graph capture does not allow: captured applies take the
        # blocks one by one
        if len(convs) == 1 or (convs and torch.cuda.is_current_stream_capturing()):
            for idx, xs, ys in convs:
                self._content[idx]._apply_into(xs, ys, adjoint)
        elif convs:
            b = self._conv_batch(convs, k, out_dt)
            b.bind_views([c[1] for c in convs], [c[2] for c in convs], adjoint)
            b.launch(adjoint)
        return Y

    def _conv_batch(self, convs, k, dt):
        from .batch import ConvBatch
        key = (tuple(c[0] for c in convs), k, dt.name)
        b = self._batches.get(key)
        if b is None:
            b = self._batches[key] = ConvBatch([self._content[c[0]] for c in convs],
                                               [k] * len(convs), dtype=dt)
        return b

    def _fwd(self, X):
        return self._run(X, False)

    def _bwd(self, X):
        return self._run(X, True)

    def _getColNorms(self):
        return np.concatenate([np.asarray(t.colNorms, dtype=np.float64)
                               for t in self._content])

    def _getLargestSingularValue(self):
        # the singular values of a block diagonal matrix are its blocks'
        return float(max(t.largestSingularValue for t in self._content))

    def _reference(self):
        res = np.zeros((self.numRows, self.numCols), dtype=self.dtype)
        r = c = 0
        for t in self._content:
            res[r:r + t.numRows, c:c + t.numCols] = t.reference()
            r += t.numRows
            c += t.numCols
        return res


class Blocks(Matrix):
    """A grid of blocks [[A_11, A_12, ...], [A_21, ...], ...]
    (Blocks.pyx:65-312): block row r of the result is sum_c A_rc x_c."""

    def __init__(self, arrMatrices, **options):
        if not isinstance(arrMatrices, (list, tuple)):
            raise ValueError("Blocks: Not a nested list of fastmat matrices.")
        if len(arrMatrices) < 1:
            raise ValueError("Blocks: Contains no matrices.")
        for rr, rowMatrices in enumerate(arrMatrices):
            if isinstance(rowMatrices, Matrix):
                raise TypeError(
                    ("Blocks.row(%d) is of type %s. Instantiate Blocks " +
                     "with a 2D structure of iterables (e.g. nested " +
                     "lists) holding fastmat Matrix instances in their " +
                     "second level.") % (rr, repr(rowMatrices)))
            for cc, item in enumerate(rowMatrices):
                if not isinstance(item, Matrix):
                    raise TypeError(
                        ("Blocks[%d, %d] is of type %s and not a " +
                         "fastmat Matrix isntance.") % (rr, cc, repr(item)))
        self._rows = tuple(tuple(r) for r in arrMatrices)
        nrows, ncols = len(self._rows), len(self._rows[0])
        first = self._rows[0]
        self._rowSize = tuple(r[0].numRows for r in self._rows)
        self._colSize = tuple(t.numCols for t in first)
        dt = np.dtype(np.int8)
        for rr, row in enumerate(self._rows):
            if len(row) != ncols:
                raise ValueError("Blocks.row(%d) has incompatible number of entries" % (rr,))
            for cc, term in enumerate(row):
                if term.numRows != row[0].numRows:
                    raise ValueError(
                        ("Blocks[%d,%d] with shape %s is incompatible with " +
                         "row height (%d,:)") % (rr, cc, str(term.shape), row[0].numRows))
                if term.numCols != first[cc].numCols:
                    raise ValueError(
                        ("Blocks[%d,%d] with shape %s is incompatible with " +
                         "column width (:,%d)") % (rr, cc, str(term.shape), first[cc].numCols))
                dt = np.promote_types(dt, term.dtype)
        self._cols = tuple(tuple(self._rows[r][c] for r in range(nrows)) for c in range(ncols))
        self._content = tuple(t for row in self._rows for t in row)
        self._initProperties(sum(self._rowSize), sum(self._colSize), dt, **options)
        self._widenInputDatatype = True

    def _run(self, X, adjoint):
        k = X.shape[0]
        out_dt = D.np_dtype(X.dtype)
        # forward: output block r = sum_c A_rc x_c; adjoint: output block
        # c = sum_r A_rc^H x_r
        lines = self._cols if adjoint else self._rows
        in_sizes = self._rowSize if adjoint else self._colSize
        out_sizes = self._colSize if adjoint else self._rowSize
        Y = D.empty(k, sum(out_sizes), out_dt)
        if k == 0:
            return Y
        in_off = np.concatenate([[0], np.cumsum(in_sizes)]).astype(int)
        o = 0
        for line, mout in zip(lines, out_sizes):
            ys = Y[:, o:o + mout]
            o += mout
            if mout == 0:
                continue
            first = True
            for j, term in enumerate(line):
                xs = X[:, in_off[j]:in_off[j + 1]]
                if xs.shape[1] == 0:
                    continue
                if first:
                    term._apply_into(xs, ys, adjoint)
                    first = False
                    continue
                T = term._apply(xs if xs.is_contiguous() else xs.contiguous(), adjoint)
                if T.dtype != Y.dtype:
                    T = T.to(Y.dtype)
                _lib.call("fm_axpby", D.fm_type(Y.dtype), mout, k, _lib.dvec(1.0), _lib.ptr(T),
                          mout, 0, _lib.dvec(1.0), _lib.ptr(ys), D.ld(ys), D.stream())
            if first:
                ys.zero_()
        return Y

    def _fwd(self, X):
        return self._run(X, False)

    def _bwd(self, X):
        return self._run(X, True)

    def _reference(self):
        res = np.zeros((self.numRows, self.numCols), dtype=self.dtype)
        r = 0
        for row, h in zip(self._rows, self._rowSize):
            c = 0
            for term, w in zip(row, self._colSize):
                res[r:r + h, c:c + w] = term.reference()
                c += w
            r += h
        return res

"""Device-side array plumbing shared by all operators.

Internal block format: a contiguous torch tensor of shape (k, n) on the HIP
device, i.e. k right-hand-side vectors of length n stored one after another.
This is exactly the F-contiguous (n, k) array fastmat's C-style classes work
on (Matrix.pyx:1892-1901, core/cmath.pyx:343-386), so every block can be
handed to the C ABI as a column-major (n, k) array with leading dimension n.

User-facing inputs may be numpy arrays or torch tensors, 1-D (n,) or 2-D
(n, k) (Matrix.pyx:1770-1790).  Results are returned in the caller's kind:
numpy in -> numpy out (F-ordered, like fastmat), torch in -> torch out as an
(m, k) view whose columns are contiguous.
"""
import numpy as np
import torch

from . import _lib

_NP2T = {
    np.dtype(np.int8): torch.int8, np.dtype(np.int16): torch.int16,
    np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64,
    np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64,
    np.dtype(np.complex64): torch.complex64, np.dtype(np.complex128): torch.complex128,
}
_T2NP = {v: k for k, v in _NP2T.items()}

SUPPORTED = tuple(_NP2T)


def np_dtype(dt):
    """numpy dtype of a numpy dtype-like or torch dtype."""
    if isinstance(dt, torch.dtype):
        if dt not in _T2NP:
            raise TypeError("Unsupported data type %s" % dt)
        return _T2NP[dt]
    return np.dtype(dt)


def torch_dtype(dt):
    dt = np_dtype(dt)
    if dt not in _NP2T:
        raise TypeError("Unsupported data type %s" % dt)
    return _NP2T[dt]


def fm_type(dt):
    return _lib.NP_TO_FM[np_dtype(dt)]


def promote(a, b):
    """numpy promotion (the reference's promotion table equals
    np.promote_types, core/types.pyx:418-453)."""
    return np.promote_types(np_dtype(a), np_dtype(b))


def safe_type_expansion(dtype):
    """int8/int16 -> float32, int32/int64 -> float64 (core/types.pyx:378-394)."""
    dt = np_dtype(dtype)
    if dt in (np.dtype(np.int8), np.dtype(np.int16)):
        return np.dtype(np.float32)
    if dt in (np.dtype(np.int32), np.dtype(np.int64)):
        return np.dtype(np.float64)
    return dt


_device = None


def device():
    global _device
    if _device is None:
        if not torch.cuda.is_available():
            raise _lib.FastmatAmdError(
                "fastmat_amd needs a HIP device; none is visible")
        _device = torch.device("cuda", torch.cuda.current_device())
    return _device


class Boundary:
    """Records how the caller passed an array, to return results alike."""

    __slots__ = ("kind", "was1d", "dev")

    def __init__(self, kind, was1d, dev):
        self.kind = kind
        self.was1d = was1d
        self.dev = dev

    def finish(self, Y):
        """Y: internal (k, m) block -> caller's format."""
        out = Y.t()  # (m, k), columns contiguous
        if self.was1d:
            out = out[:, 0]
        if self.kind == "numpy":
            return out.cpu().numpy()
        if self.kind == "torch-cpu":
            return out.cpu()
        return out


def to_block(arr, required):
    """Validate and convert a caller array into an internal (k, n) block.

    Errors follow Matrix._prepareInputArray (Matrix.pyx:1770-1782).
    """
    if isinstance(arr, torch.Tensor):
        kind = "torch" if arr.is_cuda else "torch-cpu"
        t = arr
    else:
        a = np.asarray(arr)
        if a.dtype not in _NP2T:
            if a.dtype == np.bool_ or a.dtype.kind in "ui":
                a = a.astype(np.promote_types(a.dtype, np.int8)) if a.dtype.kind != "u" \
                    else a.astype(np.int64)
            else:
                raise TypeError("Unsupported data type %s" % a.dtype)
        kind = "numpy"
        t = torch.from_numpy(np.ascontiguousarray(a.T) if a.ndim == 2 else np.ascontiguousarray(a))
        t = t.t() if a.ndim == 2 else t
    if t.dim() < 1 or t.dim() > 2:
        raise ValueError("Input data array must be 1D or 2D")
    if t.shape[0] != required:
        raise ValueError("Mismatch of vector size %d to relevant matrix axis %d" % (
            t.shape[0], required))
    if t.dtype not in _T2NP:
        raise TypeError("Unsupported data type %s" % t.dtype)
    was1d = t.dim() == 1
    if was1d:
        t = t.reshape(required, 1)
    dev = device()
    block = t.t()
    if block.device != dev:
        block = block.to(dev, non_blocking=False)
    if not block.is_contiguous():
        block = block.contiguous()
    return block, Boundary(kind, was1d, dev)


def empty(k, n, dtype):
    return torch.empty((k, n), dtype=torch_dtype(dtype), device=device())


def stream():
    return _lib.stream_handle(device())


def ld(t):
    """Leading dimension (elements between rows) of a (k, w) view with unit
    element stride; a single row may report any row stride, so its width is
    used."""
    return int(t.shape[1]) if t.shape[0] == 1 else int(t.stride(0))


def workspace(nbytes):
    """Device scratch of at least nbytes (uint8 tensor; caller keeps it alive
    until the stream has consumed it, which torch's caching allocator
    guarantees for tensors freed after enqueueing on the current stream)."""
    if nbytes <= 0:
        return None
    return torch.empty(int(nbytes), dtype=torch.uint8, device=device())


def squeeze1d(a):
    """Squeezed, at least 1-D copy (core/cmath.pyx:389-441, _arrSqueezedCopy)."""
    return np.atleast_1d(np.squeeze(np.array(a))).copy()


def as_tensor(values, dtype):
    """Host array -> device tensor of `dtype` (1-D or 2-D)."""
    a = np.ascontiguousarray(np.asarray(values).astype(np_dtype(dtype)))
    return torch.from_numpy(a).to(device())

"""ctypes binding of oracle/fwht.c, the C restatement of the reference's
typed radix-2 Hadamard loop (Hadamard.pyx:36-61, 164-230).  Test
infrastructure: built by oracle/Makefile into oracle/_build/."""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libfwht_oracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        _lib = ctypes.CDLL(_SO)
    return _lib


_INT = {np.dtype(np.int8): "fwht_oracle_i8", np.dtype(np.int16): "fwht_oracle_i16",
        np.dtype(np.int32): "fwht_oracle_i32", np.dtype(np.int64): "fwht_oracle_i64"}


def hadamard(x):
    """Reference-order FWHT of x (n,) or (n, k); output dtype == input dtype
    (promote(dtype, int8) == dtype for all eight supported types)."""
    lib = _load()
    x = np.asarray(x)
    was1d = x.ndim == 1
    x2 = x[:, None] if was1d else x
    n, k = x2.shape
    order = n.bit_length() - 1
    if (1 << order) != n:
        raise ValueError("length must be a power of two")
    xf = np.asfortranarray(x2)
    y = np.empty_like(xf, order="F")
    p = ctypes.c_void_p
    i64 = ctypes.c_int64
    dt = xf.dtype
    if dt in _INT:
        getattr(lib, _INT[dt])(p(xf.ctypes.data), p(y.ctypes.data), ctypes.c_int(order),
                               i64(k), i64(n), i64(n))
    elif dt in (np.float32, np.complex64):
        lib.fwht_oracle_f32(p(xf.ctypes.data), p(y.ctypes.data), ctypes.c_int(order),
                            i64(k), i64(n), i64(n), ctypes.c_int(2 if dt == np.complex64 else 1))
    elif dt in (np.float64, np.complex128):
        lib.fwht_oracle_f64(p(xf.ctypes.data), p(y.ctypes.data), ctypes.c_int(order),
                            i64(k), i64(n), i64(n), ctypes.c_int(2 if dt == np.complex128 else 1))
    else:
        raise TypeError("unsupported dtype %s" % dt)
    return y[:, 0] if was1d else y

"""ctypes binding of ``libfastmat_amd.so`` (the C ABI in include/fastmat_amd.h).

The library is loaded after ``torch`` so that it binds to the HIP runtime
torch already mapped (same soname), i.e. one HIP runtime per process and
torch's streams are valid handles for it.  There is no fallback: if the
library is missing or fails to load, every GPU operator raises.
"""
import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfastmat_amd.so")

# element type ids (include/fastmat_amd.h; reference core/types.pxd:45-55)
FM_INT8, FM_INT16, FM_INT32, FM_INT64 = 0, 1, 2, 3
FM_FLOAT32, FM_FLOAT64, FM_COMPLEX64, FM_COMPLEX128 = 4, 5, 6, 7

NP_TO_FM = {
    np.dtype(np.int8): FM_INT8, np.dtype(np.int16): FM_INT16,
    np.dtype(np.int32): FM_INT32, np.dtype(np.int64): FM_INT64,
    np.dtype(np.float32): FM_FLOAT32, np.dtype(np.float64): FM_FLOAT64,
    np.dtype(np.complex64): FM_COMPLEX64, np.dtype(np.complex128): FM_COMPLEX128,
}

_c_i64 = ctypes.c_int64
_c_int = ctypes.c_int
_vp = ctypes.c_void_p
_pp = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes)
_PROTOTYPES = {
    "fm_abi_version": (_c_int, []),
    "fm_get_last_error": (ctypes.c_char_p, []),
    "fm_hadamard_apply": (_c_int, [_c_int, _c_int, _vp, _c_i64, _vp, _c_i64, _c_i64, _vp]),
    "fm_hadamard_apply_strided": (_c_int, [_c_int, _c_int, _vp, _vp, _c_i64, _c_i64, _vp]),
    "fm_wht_tile_check": (_c_int, [_c_int, _c_int, _c_int, _c_int, ctypes.POINTER(_c_i64)]),
    "fm_fourier_plan_create": (_c_int, [_c_i64, _c_int, _vp, _pp]),
    "fm_fourier_plan_info": (_c_int, [_vp, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_int),
                                      ctypes.POINTER(_c_int)]),
    "fm_fourier_workspace_size": (_c_int, [_vp, _c_i64, ctypes.POINTER(_c_i64)]),
    "fm_fourier_apply": (_c_int, [_vp, _c_int, _vp, _c_i64, _vp, _c_i64, _c_i64, _vp, _c_i64,
                                  _vp]),
    "fm_fourier_apply_strided": (_c_int, [_vp, _c_int, _vp, _vp, _c_i64, _c_i64, _vp]),
    "fm_fourier_plan_destroy": (_c_int, [_vp]),
    "fm_circulant_plan_create": (_c_int, [_c_int, _c_i64, _vp, _vp, _pp]),
    "fm_toeplitz_plan_create": (_c_int, [_c_int, _c_i64, _c_i64, _vp, _vp, _pp]),
    "fm_circulant2_plan_create": (_c_int, [_c_int, _c_i64, _c_i64, _vp, _vp, _pp]),
    "fm_conv_plan_info": (_c_int, [_vp, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i64),
                                   ctypes.POINTER(_c_i64), ctypes.POINTER(_c_int)]),
    "fm_conv_workspace_size": (_c_int, [_vp, _c_i64, ctypes.POINTER(_c_i64)]),
    "fm_conv_apply": (_c_int, [_vp, _c_int, _vp, _c_i64, _vp, _c_i64, _c_i64, _vp, _c_i64, _vp]),
    "fm_conv_apply_strided": (_c_int, [_vp, _c_int, _vp, _vp, _c_i64, _c_i64, _vp]),
    "fm_conv_apply_noalloc": (_c_int, [_vp, _c_int, _vp, _c_i64, _vp, _c_i64, _c_i64, _vp]),
    "fm_conv_generator_grad_workspace_size": (_c_int, [_vp, _c_i64, ctypes.POINTER(_c_i64)]),
    "fm_conv_generator_grad": (_c_int, [_vp, _vp, _c_i64, _vp, _c_i64, _c_i64, _vp, _vp, _c_i64,
                                        _vp]),
    "fm_conv_set_generator_workspace_size": (_c_int, [_vp, ctypes.POINTER(_c_i64)]),
    "fm_conv_set_generator": (_c_int, [_vp, _c_int, _vp, _vp, _c_i64, _vp]),
    "fm_conv_plan_clone": (_c_int, [_vp, _vp, _pp]),
    "fm_conv_plan_destroy": (_c_int, [_vp]),
    "fm_dense_apply": (_c_int, [_c_int, _c_int, _c_int, _c_i64, _c_i64, _vp, _c_i64, _vp, _c_i64,
                                _vp, _c_i64, _c_i64, _vp]),
    "fm_dense_apply_strided": (_c_int, [_c_int, _c_int, _c_int, _c_i64, _vp, _c_i64, _vp, _vp,
                                        _c_i64, _c_i64, _vp]),
    "fm_kron_conv_dense_apply": (_c_int, [_vp, _c_int, _vp, _c_i64, _vp, _vp, _c_i64, _vp]),
    "fm_diag_apply": (_c_int, [_c_int, _c_int, _c_i64, _vp, _vp, _c_i64, _vp, _c_i64, _c_i64,
                               _vp]),
    "fm_gather_rows": (_c_int, [_c_int, _c_i64, _vp, _vp, _c_i64, _vp, _c_i64, _c_i64, _vp]),
    "fm_scatter_rows": (_c_int, [_c_int, _c_i64, _vp, _vp, _c_i64, _vp, _c_i64, _c_i64, _c_i64,
                                 _vp]),
    "fm_axpby": (_c_int, [_c_int, _c_i64, _c_i64, ctypes.POINTER(ctypes.c_double), _vp, _c_i64,
                          _c_int, ctypes.POINTER(ctypes.c_double), _vp, _c_i64, _vp]),
    "fm_batch_plan_create": (_c_int, [_c_int, ctypes.POINTER(_vp), ctypes.POINTER(_c_i64), _pp]),
    "fm_batch_workspace_size": (_c_int, [_vp, ctypes.POINTER(_c_i64)]),
    "fm_batch_set_grid": (_c_int, [_vp, _c_int]),
    "fm_batch_set_schedule": (_c_int, [_vp, _c_int, _c_int, _c_int]),
    "fm_batch_bind": (_c_int, [_vp, _c_int, ctypes.POINTER(_vp), ctypes.POINTER(_c_i64),
                               ctypes.POINTER(_vp), ctypes.POINTER(_c_i64), _vp, _c_i64, _vp]),
    "fm_batch_apply": (_c_int, [_vp, _c_int, _vp]),
    "fm_batch_apply_redirected": (_c_int, [_vp, _c_int, _c_i64, _vp, _vp]),
    "fm_batch_lsq_backward": (_c_int, [_vp, _c_i64, _vp, _c_i64, _vp]),
    "fm_batch_bind_residual": (_c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_c_i64),
                                        ctypes.POINTER(_vp), ctypes.POINTER(_c_i64),
                                        ctypes.POINTER(_vp), ctypes.POINTER(_c_i64),
                                        ctypes.c_double, _vp, _c_i64, _vp]),
    "fm_batch_bind_lsq": (_c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_c_i64),
                                   ctypes.POINTER(_vp), ctypes.POINTER(_c_i64),
                                   ctypes.POINTER(_vp), ctypes.POINTER(_c_i64),
                                   ctypes.POINTER(_vp), ctypes.POINTER(_c_i64),
                                   ctypes.c_double, _vp, _c_i64, _vp]),
    "fm_batch_residual_loss": (_c_int, [_vp, _vp, _vp]),
    "fm_batch_grad_workspace_size": (_c_int, [_vp, ctypes.POINTER(_c_i64)]),
    "fm_batch_grad_bind": (_c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_c_i64),
                                    ctypes.POINTER(_vp), ctypes.POINTER(_c_i64), ctypes.POINTER(_vp),
                                    _vp, _c_i64, _vp]),
    "fm_batch_grad_apply": (_c_int, [_vp, _vp]),
    "fm_batch_grad_apply_reuse": (_c_int, [_vp, _vp]),
    "fm_batch_gen_workspace_size": (_c_int, [_vp, ctypes.POINTER(_c_i64)]),
    "fm_batch_gen_bind": (_c_int, [_vp, _c_int, ctypes.POINTER(_vp), _vp, _c_i64, _vp]),
    "fm_batch_gen_apply": (_c_int, [_vp, _vp]),
    "fm_batch_gen_forward_apply": (_c_int, [_vp, _vp]),
    "fm_batch_plan_destroy": (_c_int, [_vp]),
    "fm_shrink_step": (_c_int, [_c_int, _c_i64, _vp, _vp, ctypes.c_double, ctypes.c_double, _vp,
                                _vp, _vp, _vp, ctypes.c_double, _vp, _vp]),
    "fm_counter_advance": (_c_int, [_vp, _vp]),
    "fm_lfsr_sequence": (_c_int, [ctypes.c_uint32, ctypes.c_uint32, _c_int,
                                  ctypes.POINTER(_c_i64), _vp, _vp, _c_i64]),
    "fm_lfsr_apply": (_c_int, [_c_int, _c_int, _c_i64, _vp, _vp, _vp, _c_i64, _vp, _c_i64, _c_i64,
                               _vp, _c_i64, _vp]),
    "fm_omp_workspace_size": (_c_int, [_c_int, _c_i64, _c_i64, _c_int, ctypes.POINTER(_c_i64)]),
    "fm_omp_step": (_c_int, [_c_int, _c_int, _c_i64, _c_i64, _c_i64, _c_int, _c_int, _vp, _vp,
                             _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_i64, _vp, _vp,
                             _vp]),
    "fm_stela_prox": (_c_int, [_c_int, _c_int, _c_i64, _c_i64, _vp, _vp, _vp, ctypes.c_double,
                               _vp, _vp, _vp, _c_i64, _vp]),
    "fm_stela_line": (_c_int, [_c_int, _c_i64, _c_i64, _vp, _vp, _vp, _c_i64, ctypes.c_double,
                               ctypes.c_double, _vp, _vp, _vp]),
    "fm_stela_update": (_c_int, [_c_int, _c_i64, _c_i64, _c_i64, _vp, _vp, _vp, _vp, _vp, _vp,
                                 _vp]),
    "fm_stela_select": (_c_int, [_c_int, _c_i64, _c_i64, _vp, _vp, _vp, _vp]),
    "fm_profiler_enable": (_c_int, [_c_int]),
    "fm_profiler_reset": (_c_int, []),
    "fm_profiler_count": (_c_int, [ctypes.POINTER(_c_int)]),
    "fm_profiler_entry": (_c_int, [_c_int, ctypes.c_char_p, _c_int, ctypes.POINTER(_c_i64),
                                   ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_double)]),
}

EXPORTED = tuple(_PROTOTYPES)

_lib = None
_lock = threading.Lock()


class FastmatAmdError(RuntimeError):
    pass


def load():
    """Load the HIP library (idempotent).  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        import torch  # noqa: F401  (binds the process to torch's HIP runtime first)
        if not os.path.exists(LIB_PATH):
            raise FastmatAmdError(
                "libfastmat_amd.so is not built (expected at %s); run "
                "`python -m fastmat_amd.build`" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_LOCAL)
        for name, (res, args) in _PROTOTYPES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if lib.fm_abi_version() != 1:
            raise FastmatAmdError("libfastmat_amd ABI version mismatch")
        _lib = lib
    return _lib


def check(status):
    """Map a status code to the exception types fastmat raises."""
    if status == 0:
        return
    msg = (_lib.fm_get_last_error() or b"").decode(errors="replace")
    if status == 1:
        raise ValueError(msg)
    if status == 2:
        raise TypeError(msg)
    if status == 4:
        raise MemoryError(msg)
    if status == 5:
        raise NotImplementedError(msg)
    raise FastmatAmdError(msg or "fastmat_amd error %d" % status)


def call(name, *args):
    lib = load()
    check(getattr(lib, name)(*args))


def stream_handle(device=None):
    """hipStream_t of torch's current stream on `device`, as a c_void_p."""
    import torch
    s = torch.cuda.current_stream(device)
    return ctypes.c_void_p(s.cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def dvec(z):
    """Complex scalar as a ctypes double[2]."""
    z = complex(z)
    return (ctypes.c_double * 2)(z.real, z.imag)


class Profiler:
    """Context manager around the library's per-kernel event timing.

    >>> with Profiler() as p:
    ...     op.forward(x)
    >>> p.results  # {kernel label: dict(launches, ms, bytes, flops, pass_bytes)}
    """

    def __init__(self):
        self.results = {}

    def __enter__(self):
        lib = load()
        check(lib.fm_profiler_reset())
        check(lib.fm_profiler_enable(1))
        return self

    def __exit__(self, *exc):
        lib = load()
        check(lib.fm_profiler_enable(0))
        n = ctypes.c_int()
        check(lib.fm_profiler_count(ctypes.byref(n)))
        buf = ctypes.create_string_buffer(128)
        for i in range(n.value):
            la, ms = _c_i64(), ctypes.c_double()
            by, fl, pb = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
            check(lib.fm_profiler_entry(i, buf, 128, ctypes.byref(la), ctypes.byref(ms),
                                        ctypes.byref(by), ctypes.byref(fl), ctypes.byref(pb)))
            self.results[buf.value.decode()] = dict(launches=la.value, ms=ms.value,
                                                    bytes=by.value, flops=fl.value,
                                                    pass_bytes=pb.value)
        check(lib.fm_profiler_reset())
        return False

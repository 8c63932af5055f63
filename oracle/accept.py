"""The reference's acceptance rule for transform tests (test infrastructure).

Restates compareResults (inspect/test.py:104-202):

    error    = max|out - ref| / max|ref|                      (:186-188)
    tolError = 5 * 2 * maxEps * (2 * sqrt(maxDim)) ** TOL_POWER  (:181-183)

maxEps is the largest machine epsilon among the input dtype, the instance
dtype and the reference dtype (each promoted with int8, :166-174), optionally
raised to TOL_MINEPS.  A result also passes when max|ref| <= tolError (:189).
"""
import numpy as np


def type_eps(dtype):
    dt = np.dtype(dtype)
    if np.issubdtype(dt, np.integer) or dt == np.bool_:
        return 0.0
    return float(np.finfo(dt).eps)


def tolerance(max_dim, tol_power=1.0, eps_dtypes=(), min_eps=0.0):
    max_eps = max([min_eps] + [type_eps(d) for d in eps_dtypes])
    dynamics = 2.0
    return 5 * dynamics * max_eps * (dynamics * np.sqrt(max_dim)) ** tol_power


def max_rel_error(out, ref):
    out = np.asarray(out)
    ref = np.asarray(ref)
    max_ref = float(np.amax(np.abs(ref))) if ref.size else 0.0
    max_diff = float(np.amax(np.abs(out.astype(np.complex128) - ref))) if ref.size else 0.0
    return max_diff / max_ref if max_ref != 0 else max_diff


def passes(out, ref, max_dim, tol_power=1.0, eps_dtypes=(), min_eps=0.0):
    tol = tolerance(max_dim, tol_power, eps_dtypes, min_eps)
    max_ref = float(np.amax(np.abs(np.asarray(ref)))) if np.asarray(ref).size else 0.0
    err = max_rel_error(out, ref)
    return (err <= tol) or (max_ref <= tol), err, tol

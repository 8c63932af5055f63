"""CPU restatement of fastmat's sparse-recovery solvers (test infrastructure).

ISTA follows algorithms/ISTA.py:113-168, FISTA algorithms/FISTA.py:111-178
and STELA algorithms/STELA.py:128-267 line by line in numpy, on a dense
matrix.  ISTA / FISTA: the step size is 1 / sigma_max^2 with sigma_max from
numpy's SVD of that dense matrix (the reference takes it from scipy's svds,
Matrix.pyx:895-1006, which agrees to working precision); the iterate starts
as promote(float32, b.dtype) and widens to whatever the products return.
STELA works in promote(float64, b.dtype) with d = 1 / colNorms^2.
"""
import numpy as np


def soft_threshold(x, alpha):
    """ISTA.softThreshold (ISTA.py:113-123): x * m / (m + alpha),
    m = max(|x| - alpha, 0)."""
    m = np.maximum(np.abs(x) - alpha, 0)
    return np.multiply(m / (m + alpha), x)


def largest_singular_value(A):
    return float(np.linalg.svd(np.atleast_2d(A), compute_uv=False)[0])


def _prep(A, b, num_max_steps):
    b = np.asarray(b)
    if b.ndim > 2 or b.ndim < 1:
        raise ValueError("Only n x m arrays are supported")
    if num_max_steps <= 0:
        raise ValueError("at least one step")
    b2 = b.reshape((-1, 1)) if b.ndim == 1 else b
    L = 1.0 / largest_singular_value(A) ** 2
    x = np.zeros((A.shape[1], b2.shape[1]), dtype=np.promote_types(np.float32, b2.dtype))
    return b2, L, x


def ista(A, b, num_lambda=0.1, num_max_steps=100):
    """Returns (result, iterate x, last step)."""
    A = np.asarray(A)
    b2, L, x = _prep(A, b, num_max_steps)
    step = x
    for _ in range(num_max_steps):
        step = x - L * A.conj().T.dot(A.dot(x) - b2)
        x = soft_threshold(step, L * num_lambda * 0.5)
    return np.where(x != 0, step, x), x, step


def fista(A, b, num_lambda=0.1, num_max_steps=100):
    A = np.asarray(A)
    b2, L, x = _prep(A, b, num_max_steps)
    t = 1
    y = np.copy(x)
    step = x
    for _ in range(num_max_steps):
        x_old = np.copy(x)
        step = y - L * A.conj().T.dot(A.dot(y) - b2)
        x = soft_threshold(step, L * num_lambda * 0.5)
        t_old = t
        # a Python float, as numpy < 2 (the reference's pinned versions)
        # treated np.sqrt's scalar: the momentum does not widen a float32
        # iterate (under numpy 2 rules np.float64 would promote it)
        t = float((1 + np.sqrt(1 + 4 * t ** 2)) / 2)
        y = x + ((t_old - 1) / t) * (x - x_old)
    return np.where(x != 0, step, x), x, step


def fista_betas(num_max_steps):
    """Momentum (t_k - 1) / t_{k+1} of each FISTA step (FISTA.py:166-172)."""
    t = 1.0
    out = []
    for _ in range(num_max_steps):
        t_old = t
        t = (1 + np.sqrt(1 + 4 * t ** 2)) / 2
        out.append((t_old - 1) / t)
    return np.array(out)


def col_norms(A):
    """Column norms of the dense matrix (Matrix.pyx:1048-1087 takes them from
    the operator applied to the identity)."""
    return np.linalg.norm(np.atleast_2d(A), axis=0)


def stela(A, b, num_lambda=0.1, num_max_steps=100, num_max_error=1e-6, col_norms_=None):
    """STELA (algorithms/STELA.py:128-267) step by step on a dense matrix.
    Returns (x, steps run, gammas of the last step).  `col_norms_` replaces
    the dense column norms (to feed the operator's closed-form norms)."""
    A = np.asarray(A)
    b = np.asarray(b)
    if b.ndim > 2 or b.ndim < 1:
        raise ValueError("Only n x m arrays are supported for STELA")
    b2 = b.reshape((-1, 1)) if b.ndim == 1 else b
    if num_max_steps <= 0:
        raise ValueError("STELA would like to do at least one step for you")
    dt = np.promote_types(np.float64, b2.dtype)
    gamma = np.zeros(b2.shape[1], dtype=dt)
    x = np.zeros((A.shape[1], b2.shape[1]), dtype=dt)
    bx = np.zeros_like(x)
    res = (-b2).astype(dt)
    abxx = np.zeros_like(b2, dtype=dt)
    z = A.conj().T.dot(res)
    norms = col_norms(A) if col_norms_ is None else np.asarray(col_norms_, dtype=np.float64)
    d = (1.0 / norms ** 2).reshape((-1, 1))
    lam = num_lambda
    for step in range(num_max_steps):
        grad = d * x - z
        diff = np.maximum(np.minimum(z.real - x.real, +lam), -lam)
        if dt == complex:
            diff = diff + 1j * np.maximum(np.minimum(z.imag - x.imag, +lam), -lam)
        stop = np.linalg.norm(z - diff, axis=0)
        act = stop > num_max_error
        if np.sum(act) == 0:
            return x, step, gamma
        with np.errstate(all="ignore"):
            # a complex value stored into the real arrays keeps its real
            # part, as the reference's assignment does
            v = soft_threshold(grad[:, act], lam) / d
            bx[:, act] = _into(bx, v)
            abxx[:, act] = _into(abxx, A.dot(bx[:, act] - x[:, act]))
            gamma[act] = np.maximum(np.minimum(
                -(np.real(np.sum(np.conj(res[:, act]) * abxx[:, act], axis=0))
                  + lam * np.sum(np.abs(bx[:, act]) - np.abs(x[:, act]), axis=0))
                / np.sum(np.abs(abxx[:, act]) ** 2, axis=0), 1), 0)
        x[:, act] += (bx[:, act] - x[:, act]).dot(np.diag(gamma[act]))
        res[:, act] += gamma[act] * abxx[:, act]
        z[:, act] = A.conj().T.dot(res[:, act])
    return x, num_max_steps, gamma


def _into(dst, v):
    return v.real if (np.iscomplexobj(v) and not np.iscomplexobj(dst)) else v


def stela_objective(A, b, x, lam):
    """0.5 ||A x - b||^2 + lam ||x||_1 per column (the function STELA's line
    search decreases)."""
    A = np.asarray(A)
    r = A.dot(x) - np.asarray(b).reshape(A.shape[0], -1)
    return 0.5 * np.sum(np.abs(r) ** 2, axis=0) + lam * np.sum(np.abs(x), axis=0)


def omp(A, b, num_max_steps):
    """OMP (algorithms/OMP.pyx:122-240) on a dense matrix, the reference's
    rank-1 pseudo-inverse update step by step.  Returns (x, support)."""
    A = np.asarray(A)
    b = np.asarray(b)
    if b.ndim > 2 or b.ndim < 1:
        raise ValueError("Only n x m arrays are supported for OMP")
    b2 = b.reshape((-1, 1)) if b.ndim == 1 else b
    if num_max_steps <= 0:
        raise ValueError("OMP would like to do at least one step for you")
    s = int(num_max_steps)
    n, m, L = A.shape[0], A.shape[1], b2.shape[1]
    Cn = A * (1.0 / col_norms(A))[None, :]
    rt = np.promote_types(np.promote_types(Cn.dtype, b2.dtype), np.float64)
    xt = np.zeros((s, L), dtype=rt)
    res = b2.astype(rt, copy=True)
    sup = np.empty((s, L), dtype=np.intp)
    pinv = np.zeros((s, n, L), dtype=rt)
    acols = np.zeros((n, s, L), dtype=rt)
    for ii in range(s):
        c = np.abs(Cn.conj().T.dot(res))
        new = np.argmax(c, axis=0)
        sup[ii, :] = new
        cols = A[:, new]
        acols[:, ii, :] = cols
        if ii == 0:
            v2 = cols
            v2n = (v2 / np.linalg.norm(v2, axis=0) ** 2).conj()
            v2y = np.einsum('ji,ji->i', v2n, b2)
            xt[0, :] = v2y
            pinv[0, :, :] = v2n
        else:
            v1 = np.einsum('ijk,jk->ik', pinv[:ii, :, :], cols)
            v2 = cols - np.einsum('ijk,jk->ik', acols[:, :ii, :], v1)
            v2n = (v2 / np.linalg.norm(v2, axis=0) ** 2).conj()
            v2y = np.einsum('ji,ji->i', v2n, b2)
            xt[:ii, :] -= v2y * v1
            xt[ii, :] += v2y
            pinv[:ii, :, :] -= np.einsum('ik,jk->jik', v2n, v1)
            pinv[ii, :, :] = v2n
        res -= v2y * v2
    x = np.zeros((m, L), dtype=rt)
    x[sup, np.arange(L)] = xt
    return x, sup

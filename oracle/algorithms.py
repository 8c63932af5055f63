"""CPU restatement of fastmat's shrinkage solvers (test infrastructure).

ISTA follows algorithms/ISTA.py:113-168 and FISTA algorithms/FISTA.py:111-178
line by line in numpy, on a dense matrix; the step size is 1 / sigma_max^2
with sigma_max from numpy's SVD of that dense matrix (the reference takes it
from scipy's svds, Matrix.pyx:895-1006, which agrees to working precision).
The working dtype follows the reference: the iterate starts as
promote(float32, b.dtype) and widens to whatever the products return.
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

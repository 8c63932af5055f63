"""Closed-form dense matrices of the hot-path operators (test infrastructure).

Each function restates one class's ``_reference()`` of the reference
(fastmat 0.2.2.post0): the dense matrix built without any fast algorithm.
These are the oracle at small sizes; ``oracle.fast64`` covers large sizes.
All float results are FP64 / complex128.
"""
import numpy as np


def fourier(order):
    """Unnormalised DFT matrix, F[j, l] = exp(-2 pi i j l / N).

    Follows Fourier.pyx:241-246.  The product j*l is reduced modulo N in int64
    before forming the angle (the same matrix; fewer rounding errors at large N
    than the reference's direct ``j*l*-2j*pi/N``).
    """
    n = int(order)
    idx = np.arange(n, dtype=np.int64)
    prod = np.mod(np.outer(idx, idx), n)
    return np.exp(prod * (-2j * np.pi / n))


def circulant(vecC):
    """C[i, j] = c[(i - j) mod N], the 1-D branch of Circulant._refRecursion
    (Circulant.pyx:414-425)."""
    c = np.asarray(vecC)
    n = c.size
    i = np.arange(n)
    return c[np.mod(i[:, None] - i[None, :], n)]


def toeplitz_tensor(tenT, numRows):
    """1-D Toeplitz from the defining vector tenT = [vecC (m), vecR (n-1)].

    Follows Toeplitz._refRecursion, single-level branch (Toeplitz.pyx:719-732):
    the lower triangle holds tenT[:m] down the columns, the upper triangle holds
    ``tenT[m:][::-1]`` along the rows, i.e. T[i, j] = tenT[(i - j) mod (m+n-1)].
    """
    t = np.asarray(tenT)
    m = int(numRows)
    n = t.size - m + 1
    T = np.zeros((m, n), dtype=np.result_type(t.dtype, np.int8))
    for ii in range(min(m, n)):
        T[ii:m, ii] = t[:m - ii]
    vecR = t[m:][::-1]
    for ii in range(min(m, n - 1)):
        T[ii, ii + 1:n] = vecR[:n - ii - 1]
    return T


def toeplitz(vecC, vecR):
    """Toeplitz(vecC, vecR): tenT = hstack(vecC, vecR), split = len(vecC)
    (Toeplitz.pyx:182-201)."""
    vecC = np.asarray(vecC)
    vecR = np.asarray(vecR)
    return toeplitz_tensor(np.hstack((vecC, vecR)), vecC.size)


def circulant_multilevel(tenC):
    """d-level circulant matrix of a d-dimensional tensor, built block by block
    like Circulant._refRecursion (Circulant.pyx:349-412): level 0 holds
    N_0 x N_0 blocks, block (r, c) is the (d-1)-level matrix of
    tenC[(r - c) mod N_0]; a 1-D tensor gives ``circulant``."""
    t = np.asarray(tenC)
    if t.ndim == 1:
        return circulant(t)
    n0 = t.shape[0]
    subs = [circulant_multilevel(t[p]) for p in range(n0)]
    b = subs[0].shape[0]
    out = np.zeros((n0 * b, n0 * b), dtype=subs[0].dtype)
    for r in range(n0):
        for c in range(n0):
            out[r * b:(r + 1) * b, c * b:(c + 1) * b] = subs[(r - c) % n0]
    return out


def toeplitz_multilevel(tenT, split):
    """d-level Toeplitz matrix of a d-dimensional tensor with per-level split
    points (rows per level), built block by block like Toeplitz._refRecursion
    (Toeplitz.pyx:637-719): level 0 has split[0] x (shape[0] - split[0] + 1)
    blocks, block (r, c) is the (d-1)-level matrix of
    tenT[(r - c) mod shape[0]]; a 1-D tensor gives ``toeplitz_tensor``."""
    t = np.asarray(tenT)
    split = [int(v) for v in np.atleast_1d(split)]
    if t.ndim == 1:
        return toeplitz_tensor(t, split[0])
    d0, r0 = t.shape[0], split[0]
    c0 = d0 - r0 + 1
    subs = [toeplitz_multilevel(t[p], split[1:]) for p in range(d0)]
    br, bc = subs[0].shape
    out = np.zeros((r0 * br, c0 * bc), dtype=subs[0].dtype)
    for r in range(r0):
        for c in range(c0):
            out[r * br:(r + 1) * br, c * bc:(c + 1) * bc] = subs[(r - c) % d0]
    return out


def hadamard(order, dtype=np.int64):
    """Sylvester Hadamard matrix H[i, j] = (-1)^popcount(i & j)
    (Hadamard.pyx:242-248 builds it with scipy.linalg.hadamard)."""
    n = 1 << int(order)
    i = np.arange(n, dtype=np.int64)
    a = np.bitwise_and(i[:, None], i[None, :])
    # parity of popcount via xor-folding
    p = a.copy()
    for s in (32, 16, 8, 4, 2, 1):
        p ^= p >> s
    return (1 - 2 * (p & 1)).astype(dtype)


def permutation(sigma):
    """P = eye(N)[sigma, :] so that P x = x[sigma] (Permutation.pyx:121-129)."""
    sigma = np.asarray(sigma)
    return np.eye(sigma.size, dtype=np.int8)[sigma, :]


def diag(vecD):
    """Diag.reference: np.diag(vecD) (Diag.pyx:170-171)."""
    return np.diag(np.asarray(vecD))


def kron(*refs):
    """np.kron chain, first factor = slowest index (Kron.pyx:344-353)."""
    dt = np.promote_types(np.asarray(refs[0]).dtype, np.float64)
    res = np.asarray(refs[0]).astype(dt)
    for r in refs[1:]:
        res = np.kron(res, r)
    return res


def product(refs, scalar=1):
    """Product reference: scalar * R0 @ R1 @ ... (Product.pyx:243-257); the
    scalar is widened to at least double precision first, as the reference
    widens it to promote(float64, dtype), so no step rounds to single."""
    sc = np.asarray(scalar)
    sc = sc.astype(np.promote_types(np.float64, sc.dtype))
    res = np.asarray(refs[-1]) * sc[()]
    for r in refs[-2::-1]:
        res = np.asarray(r).dot(res)
    return res


def sum_(refs):
    """Sum reference (Sum.pyx:178-186)."""
    res = np.zeros_like(np.asarray(refs[0]), dtype=np.result_type(*refs))
    for r in refs:
        res = res + r
    return res


def partial(ref, rows=None, cols=None):
    """Partial reference: full[rows, :][:, cols] (Partial.pyx:297-303).

    ``rows``/``cols`` may be int index arrays or bool masks, as in
    Partial.pyx:141-166."""
    a = np.asarray(ref)
    if rows is not None:
        a = a[np.asarray(rows), :]
    if cols is not None:
        a = a[:, np.asarray(cols)]
    return a


def block_diag(refs):
    """BlockDiag reference: the blocks along the diagonal, zeros elsewhere
    (BlockDiag.pyx:165-181)."""
    refs = [np.asarray(r) for r in refs]
    dt = np.result_type(*refs)
    res = np.zeros((sum(r.shape[0] for r in refs), sum(r.shape[1] for r in refs)), dtype=dt)
    i = j = 0
    for r in refs:
        res[i:i + r.shape[0], j:j + r.shape[1]] = r
        i += r.shape[0]
        j += r.shape[1]
    return res


def blocks(grid):
    """Blocks reference: block (r, c) of the result is grid[r][c]
    (Blocks.pyx:291-312)."""
    grid = [[np.asarray(t) for t in row] for row in grid]
    dt = np.result_type(*[t for row in grid for t in row])
    heights = [row[0].shape[0] for row in grid]
    widths = [t.shape[1] for t in grid[0]]
    res = np.zeros((sum(heights), sum(widths)), dtype=dt)
    i = 0
    for row, h in zip(grid, heights):
        j = 0
        for t, w in zip(row, widths):
            res[i:i + h, j:j + w] = t
            j += w
        i += h
    return res


def lfsr_sequence(polynomial, start):
    """LFSRCirculant's register sequence in pure Python (the reference's
    _reference loop, LFSRCirculant.pyx:409-430): order from the polynomial's
    highest bit, output +1 for an even state and -1 for an odd one, next state
    = (state | mask * parity(state & polynomial)) >> 1, one period."""
    order, mask = 0, 1
    while (~mask & polynomial) > mask:
        mask <<= 1
        order += 1
    mask = 1 << order
    state = first = start & (mask - 1)
    out = []
    while True:
        out.append(-1 if state & 1 else 1)
        par = bin(state & polynomial).count("1") & 1
        if par:
            state |= mask
        state >>= 1
        if state == first or len(out) >= mask:
            break
    return np.array(out, dtype=np.int8)


def lfsr_circulant(polynomial, start):
    """LFSRCirculant reference: column j is the sequence rolled by j
    (LFSRCirculant.pyx:431-438)."""
    s = lfsr_sequence(polynomial, start)
    return np.stack([np.roll(s, j) for j in range(s.size)], axis=1)

"""FP64 / exact-integer evaluators of the hot-path operators at large N
(test infrastructure).

They compute the same y = A x as ``oracle.dense`` (pinned against it at small
sizes in tests/test_oracle_fast64.py) without forming A:

* Fourier: numpy's FFT in complex128 (the reference's own leaf,
  Fourier.pyx:214), or direct FP64 summation on sampled bins with the phase
  (j*l mod N) reduced exactly in int64 (independent of any FFT).
* Circulant / Toeplitz: circular convolution in complex128 with any embedding
  length L >= 2N-1 (Circulant) or L >= m+n-1 (Toeplitz).
* Hadamard: exact int64 / FP64 fast Walsh-Hadamard transform in natural
  (Sylvester) order; integer results wrap like the reference's C loop.
* Kron: mode-wise application with the np.kron index convention.

Column-major semantics: every function takes x of shape (n,) or (n, k) and
returns (m,) or (m, k).
"""
import numpy as np


def _as2d(x):
    x = np.asarray(x)
    return (x[:, None], True) if x.ndim == 1 else (x, False)


def _ret(y, was1d):
    return y[:, 0] if was1d else y


def next_pow2(n):
    n = int(n)
    return 1 if n <= 1 else 1 << (n - 1).bit_length()


# ----------------------------------------------------------------- Fourier
def fourier(x, adjoint=False):
    """y = F x (adjoint: F^H x = conj(F conj(x)), Fourier.pyx:235-238)."""
    x2, w = _as2d(x)
    x2 = x2.astype(np.complex128)
    n = x2.shape[0]
    y = np.fft.ifft(x2, axis=0) * n if adjoint else np.fft.fft(x2, axis=0)
    return _ret(y, w)


def fourier_bins(x, bins, adjoint=False):
    """Direct FP64 summation of selected output bins of F x (no FFT).

    y[j] = sum_l exp(-+2 pi i (j*l mod N)/N) x[l], with j*l reduced in int64.
    """
    x2, w = _as2d(x)
    x2 = x2.astype(np.complex128)
    n = x2.shape[0]
    l = np.arange(n, dtype=np.int64)
    sign = 1.0 if adjoint else -1.0
    out = np.empty((len(bins), x2.shape[1]), dtype=np.complex128)
    for i, j in enumerate(bins):
        ph = np.mod(np.int64(j) * l, n)
        row = np.exp(sign * 2j * np.pi * ph / n)
        out[i] = row @ x2
    return _ret(out, w)


# -------------------------------------------------------------- Circulant
def circulant(c, x, adjoint=False, L=None):
    """y = C(c) x, C[i, j] = c[(i - j) mod N] (Circulant.pyx:418-423).

    Evaluated as a circular convolution of length L (L == N, or any
    L >= 2N - 1 with the embedding [c, 0, ..., 0, c[1:]])."""
    c = np.asarray(c).astype(np.complex128)
    n = c.size
    if L is None:
        L = n
    if L != n:
        assert L >= 2 * n - 1
        cc = np.zeros(L, dtype=np.complex128)
        cc[:n] = c
        cc[L - n + 1:] = c[1:]
    else:
        cc = c
    return _conv_embedded(cc, x, n, n, adjoint)


def _conv_embedded(t, x, m, n, adjoint):
    """Top-left m x n block of the L x L circulant of generator t applied to x."""
    x2, w = _as2d(x)
    x2 = x2.astype(np.complex128)
    L = t.size
    th = np.fft.fft(t)
    if not adjoint:
        xp = np.zeros((L, x2.shape[1]), dtype=np.complex128)
        xp[:n] = x2
        y = np.fft.ifft(th[:, None] * np.fft.fft(xp, axis=0), axis=0)[:m]
    else:
        xp = np.zeros((L, x2.shape[1]), dtype=np.complex128)
        xp[:m] = x2
        y = np.fft.ifft(np.conj(th)[:, None] * np.fft.fft(xp, axis=0),
                        axis=0)[:n]
    return _ret(y, w)


def toeplitz_embed(tenT, m, L):
    """Embed tenT = [vecC (m), vecR (n-1)] into a length-L circulant generator
    by inserting zeros at the split point m (Toeplitz.pyx:357-366)."""
    t = np.asarray(tenT).astype(np.complex128)
    d = t.size
    assert L >= d
    out = np.zeros(L, dtype=np.complex128)
    out[:m] = t[:m]
    out[L - (d - m):] = t[m:]
    return out


def toeplitz(tenT, m, x, adjoint=False, L=None):
    """y = T x for the m x n Toeplitz with T[i, j] = tenT[(i - j) mod (m+n-1)]
    (Toeplitz.pyx:719-732), via circulant embedding of length L >= m+n-1."""
    d = np.asarray(tenT).size
    n = d - m + 1
    if L is None:
        L = next_pow2(d)
    return _conv_embedded(toeplitz_embed(tenT, m, L), x, m, n, adjoint)


# -------------------------------------------------------------- Hadamard
def conv_generator_grad(kind, m, n, x, g):
    """Gradient of a real loss with respect to the generator of a Circulant
    (c, length n) or Toeplitz (tenT = [vecC (m), vecR (n-1)]) matrix, given
    the forward input x (n, k) and the output gradient g = dLoss/dy (m, k),
    in PyTorch's convention for complex tensors (grad = dL/dRe + i dL/dIm):

        grad[q] = sum_cols sum_{(i, j): idx(i, j) = q} g[i] conj(x[j])

    with idx(i, j) = (i - j) mod n for a Circulant (C[i, j] = c[(i - j) mod N],
    Circulant.pyx:418-423) and (i - j) mod (m + n - 1) for a Toeplitz
    (T[i, j] = tenT[(i - j) mod (m + n - 1)], Toeplitz.pyx:719-732).  Computed
    as FP64 FFT correlations of length n (Circulant) or m + n - 1 (Toeplitz,
    g and x zero-padded), i.e. without any power-of-two embedding."""
    x2, _ = _as2d(x)
    g2, _ = _as2d(g)
    x2 = x2.astype(np.complex128)
    g2 = g2.astype(np.complex128)
    d = n if kind == "circulant" else m + n - 1
    X = np.fft.fft(x2, n=d, axis=0)
    G = np.fft.fft(g2, n=d, axis=0)
    return np.fft.ifft(np.sum(G * np.conj(X), axis=1))


def _wrap_dtype(dtype):
    return np.dtype(dtype)


def hadamard(x, out_dtype=None):
    """Natural-order fast Walsh-Hadamard transform along axis 0.

    Integers are computed exactly in int64 and then wrapped to ``out_dtype``
    (two's-complement, like the reference's C loop Hadamard.pyx:36-61);
    floats are computed in FP64 (complex128 for complex input).
    """
    x2, w = _as2d(x)
    n = x2.shape[0]
    order = n.bit_length() - 1
    assert (1 << order) == n
    if out_dtype is None:
        out_dtype = x2.dtype
    out_dtype = np.dtype(out_dtype)
    if np.issubdtype(out_dtype, np.integer):
        y = x2.astype(np.int64)
        with np.errstate(over='ignore'):
            y = _fwht_stages(y)
        y = y.astype(out_dtype)  # wraps modulo 2^bits (two's complement)
    else:
        work = np.complex128 if np.iscomplexobj(x2) else np.float64
        y = _fwht_stages(x2.astype(work))
    return _ret(y, w)


def _fwht_stages(y):
    n, k = y.shape
    h = 1
    y = y.copy()
    while h < n:
        v = y.reshape(n // (2 * h), 2, h, k)
        a = v[:, 0].copy()
        b = v[:, 1]
        v[:, 0] = a + b
        v[:, 1] = a - b
        h *= 2
    return y


# ------------------------------------------------------------------ Kron
def kron_apply(mode_ops, dims, x):
    """Apply Kron(A_1, ..., A_d) to x mode-wise (np.kron convention: factor 1
    is the slowest index, Kron.pyx:267-303).

    ``mode_ops[i]`` maps a (dims[i], batch) complex128 array to its image under
    A_i (e.g. a dense matrix product or ``fourier``)."""
    x2, w = _as2d(x)
    k = x2.shape[1]
    t = x2.astype(np.complex128).reshape(tuple(dims) + (k,))
    for i, op in enumerate(mode_ops):
        t = np.moveaxis(t, i, 0)
        shp = t.shape
        t = op(t.reshape(shp[0], -1)).reshape(shp)
        t = np.moveaxis(t, 0, i)
    return _ret(t.reshape(-1, k), w)

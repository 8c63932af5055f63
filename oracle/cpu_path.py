"""Restatement of the reference's CPU forward/backward (test infrastructure).

This is the timed CPU baseline (``cpu_baseline.kind == "port"``): the same
numpy calls in the same order as fastmat 0.2.2's CPU path, so that the time
it takes on the host is what fastmat would take for the same call.

* Fourier radix path: ``np.fft.fft(X, axis=0)`` (Fourier.pyx:213-214); for
  complex64 input numpy 2.x returns complex64.
* Fourier Bluestein path: set-up Fourier.pyx:127-156 (chirp angle
  ``n**2 * pi / N`` in float64, wrapped convolution kernel, its FFT) and
  apply Fourier.pyx:215-231 in a complex128 buffer of length L.
* Circulant / Toeplitz: ``Partial(Product(F.H, Diag(fft(c')/L), F))``
  (Circulant.pyx:104-137, Toeplitz.pyx:211-316): zero-pad scatter
  (Partial.pyx:268-280), ``astype`` (Product.pyx:215), fft, diagonal multiply
  (Diag.pyx:149-157), conj/fft/conj (Fourier.pyx:235-238), row gather.
* Hadamard: the C restatement of the typed radix-2 loop (oracle/fwht.c).
"""
import numpy as np

from . import heuristics


class FourierCPU:
    def __init__(self, order, optimize=True, max_stage=4):
        self.order = N = int(order)
        self.L = heuristics.fourier_bluestein_length(N, optimize, max_stage)
        if self.L > 0:
            L = self.L
            samples = np.linspace(0, N - 1, N)
            arg = (samples ** 2) * np.pi / N
            conv = np.empty(L, dtype=np.complex128)
            pre = np.empty(N, dtype=np.complex128)
            buf = np.cos(arg)
            conv.real[:N] = buf
            pre.real[:] = buf
            np.sin(arg, out=buf)
            conv.imag[:N] = buf
            pre.imag[:] = -buf
            conv[L - N + 1:] = conv[1:N][::-1]
            conv[N:L - N + 1] = 0
            self.pre = pre
            self.conv_hat = np.fft.fft(conv)

    def forward(self, X):
        if self.L == 0:
            return np.fft.fft(X, axis=0)
        N, L = self.order, self.L
        res = np.empty((L, X.shape[1]),
                       dtype=np.promote_types(np.complex128, X.dtype), order="F")
        res[N:] = 0
        res[:N] = (self.pre.T * X.T).T
        res = (self.conv_hat.T * np.fft.fft(res, axis=0).T).T
        return (self.pre.T * np.fft.ifft(res, axis=0)[:N, :].T).T

    def backward(self, X):
        res = self.forward(np.conj(X))
        np.conjugate(res, out=res)
        return res


class ConvCPU:
    """Partial(Product(F.H, Diag(fft(t')/L), F)) with rows/cols selection."""

    def __init__(self, t_embedded, m, n, fourier_optimize=True):
        self.L = L = t_embedded.size
        self.m, self.n = m, n
        self.d = np.fft.fft(t_embedded.astype(np.complex128)) / L
        self.F = FourierCPU(L, optimize=fourier_optimize)

    def _product(self, X, d):
        X = X.astype(np.promote_types(X.dtype, np.complex128))
        X = self.F.forward(X)
        X = (d[:, None] * X)
        return self.F.backward(X)

    def forward(self, X):
        L, m, n = self.L, self.m, self.n
        if n != L:
            Xf = np.zeros((L, X.shape[1]), dtype=X.dtype, order="F")
            Xf[:n] = X
        else:
            Xf = X
        Y = self._product(Xf, self.d)
        return Y[:m] if m != L else Y

    def backward(self, X):
        L, m, n = self.L, self.m, self.n
        if m != L:
            Xf = np.zeros((L, X.shape[1]), dtype=X.dtype, order="F")
            Xf[:m] = X
        else:
            Xf = X
        # Product._backward applies factors first-to-last with their backward:
        # (F.H).backward = F.forward, Diag.backward = conj(d), F.backward.
        X = Xf.astype(np.promote_types(Xf.dtype, np.complex128))
        X = self.F.forward(X)
        X = np.conj(self.d)[:, None] * X
        Y = self.F.backward(X)
        return Y[:n] if n != L else Y


def circulant(c, optimize=True):
    c = np.asarray(c)
    N = c.size
    L = heuristics.circulant_embedding_length(N, optimize)
    if L != N:
        t = np.concatenate([c, np.zeros(L - (2 * N - 1), dtype=c.dtype), c[1:]])
    else:
        t = c
    return ConvCPU(t, N, N)


def toeplitz(vecC, vecR, optimize=True):
    vecC = np.asarray(vecC)
    vecR = np.asarray(vecR)
    m, n = vecC.size, vecR.size + 1
    d = m + n - 1
    L = heuristics.toeplitz_embedding_length(m, n, optimize)
    t = np.concatenate([vecC, np.zeros(L - d, dtype=vecC.dtype), vecR])
    # Toeplitz builds its Fourier with optimize=False (Toeplitz.pyx:219-220)
    return ConvCPU(t, m, n, fourier_optimize=False)


def hadamard_forward(X):
    from . import fwht_c
    return fwht_c.hadamard(X)

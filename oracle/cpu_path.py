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
* Kron: the term loop of Kron.pyx:267-339 with its reshapes; dense terms are
  ``ndarray.dot`` (Matrix.pyx:1831-1842, 1924-1935).
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


class DenseCPU:
    """Matrix._forward / _backward (Matrix.pyx:1831-1842, 1924-1935):
    ``ndarray.dot`` with the array and its conjugate transpose."""

    def __init__(self, A):
        self.A = np.asarray(A)
        self.AH = self.A.T.conj()

    def forward(self, X):
        return self.A.dot(X)

    def backward(self, X):
        return self.AH.dot(X)


class KronCPU:
    """Kron._forward / _backward (Kron.pyx:267-339): for every term in order,
    reshape the data to (head, rest) in C order, to (term rows, rest') in
    Fortran order, apply the term, reshape back in Fortran order; a last C
    order reshape to (N, k).  ``terms`` have forward/backward on 2-D arrays."""

    def __init__(self, terms, dims):
        self.terms = list(terms)
        self.dims = tuple(int(d) for d in dims)
        self.N = int(np.prod(self.dims))

    def _run(self, X, backward):
        N, k = self.N, X.shape[1]
        data = X
        head = 1
        for term, d in zip(self.terms, self.dims):
            head *= d
            data = np.reshape(data, (head, k * N // head), order="C")
            data = np.reshape(data, (d, k * N // d), order="F")
            data = term.backward(data) if backward else term.forward(data)
            data = np.reshape(data, (head, k * N // head), order="F")
        return np.reshape(data, (N, k), order="C")

    def forward(self, X):
        return self._run(X, False)

    def backward(self, X):
        return self._run(X, True)


def hadamard_forward(X):
    from . import fwht_c
    return fwht_c.hadamard(X)


class CirculantMultilevelCPU:
    """Multi-level Circulant with power-of-two levels (no embedding):
    Product(K.H, Diag(fftn(t)/L), K), K = Kron(Fourier(L_1), ..., Fourier(L_d))
    (Circulant.pyx:138-215), each Kron factor applied along its mode with
    np.fft.fft (Kron.pyx:244-300, Fourier.pyx:213-214) and F.H as
    conj(fft(conj(.))) (Fourier.pyx:235-238)."""

    def __init__(self, tenC):
        t = np.asarray(tenC)
        self.shape = t.shape
        self.d = np.fft.fftn(t.astype(np.complex128)) / t.size

    def _k(self, Z, inverse):
        for ax in range(len(self.shape)):
            Z = np.conj(np.fft.fft(np.conj(Z), axis=ax)) if inverse else np.fft.fft(Z, axis=ax)
        return Z

    def _apply(self, X, d):
        k = X.shape[1]
        Z = X.astype(np.promote_types(X.dtype, np.complex128)).reshape(self.shape + (k,))
        Z = self._k(Z, False)
        Z = d[..., None] * Z
        return self._k(Z, True).reshape(X.shape)

    def forward(self, X):
        return self._apply(X, self.d)

    def backward(self, X):
        return self._apply(X, np.conj(self.d))


class LFSRCirculantCPU:
    """LFSRCirculant._core (LFSRCirculant.pyx:316-396): scatter the rows to
    the register states (zeros elsewhere), the typed radix-2 Hadamard loop
    (the C restatement), gather the tap-side states; `states` and `taps` are
    one period of each sequence."""

    def __init__(self, order, states, taps):
        self.order = order
        n = len(states)
        neg = (-np.arange(n)) % n
        self.sf, self.sb = np.asarray(states)[neg], np.asarray(states)
        self.gf, self.gb = np.asarray(taps), np.asarray(taps)[neg]

    def _core(self, X, sc, ga):
        from . import fwht_c
        Z = np.zeros((1 << self.order, X.shape[1]), dtype=X.dtype)
        Z[sc] = X
        return fwht_c.hadamard(Z)[ga]

    def forward(self, X):
        return self._core(X, self.sf, self.gf)

    def backward(self, X):
        return self._core(X, self.sb, self.gb)

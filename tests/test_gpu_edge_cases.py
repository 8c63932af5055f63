"""Edge cases of the GPU path: empty blocks (zero columns), size-1 operators,
the largest transform lengths the tile engine supports, and the error paths
(wrong shapes, unsupported lengths) that must raise instead of launching.

Oracle: the closed forms of oracle.dense at small sizes and np.fft in
complex128 (oracle.fast64) at large sizes; max-norm relative error <= 1e-5
for complex64 results, exact for integer results.
"""
import sys

import numpy as np
import pytest

from oracle import accept, dense, fast64

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def fm():
    import fastmat_amd
    from fastmat_amd import _lib
    _lib.load()
    return fastmat_amd


def operators(fm, rng):
    c = (rng.standard_normal(12) + 1j * rng.standard_normal(12)).astype(np.complex64)
    A = rng.standard_normal((5, 7)).astype(np.float32)
    return [
        ("Fourier(64)", fm.Fourier(64), dense.fourier(64)),
        ("Fourier(35)", fm.Fourier(35), dense.fourier(35)),
        ("Circulant(12)", fm.Circulant(c), dense.circulant(c)),
        ("Toeplitz(5x8)", fm.Toeplitz(c[:5], c[5:]), dense.toeplitz(c[:5], c[5:])),
        ("Hadamard(5)", fm.Hadamard(5), dense.hadamard(5, np.int8)),
        ("Diag", fm.Diag(c), dense.diag(c)),
        ("Permutation", fm.Permutation(rng.permutation(12)), None),
        ("Matrix(5x7)", fm.Matrix(A), A),
        ("Kron", fm.Kron(fm.Fourier(4), fm.Hadamard(2)), dense.kron(dense.fourier(4),
                                                                   dense.hadamard(2, np.int8))),
    ]


def test_zero_columns(fm):
    rng = np.random.default_rng(11)
    for name, op, _ in operators(fm, rng):
        say("zero columns", name)
        for adjoint in (False, True):
            n_in = op.numRows if adjoint else op.numCols
            n_out = op.numCols if adjoint else op.numRows
            x = np.zeros((n_in, 0), dtype=np.complex64)
            y = op.backward(x) if adjoint else op.forward(x)
            assert y.shape == (n_out, 0), (name, adjoint, y.shape)
            assert y.dtype == np.promote_types(np.promote_types(op.dtype, np.int8), x.dtype)


def test_size_one_operators(fm):
    # (Hadamard has no size-one instance: order 0 is rejected like the
    # reference, Hadamard.pyx:116-117)
    with pytest.raises(ValueError):
        fm.Hadamard(0)
    rng = np.random.default_rng(12)
    c = np.array([1.5 - 0.5j], dtype=np.complex64)
    cases = [("Fourier(1)", fm.Fourier(1), np.ones((1, 1))),
             ("Circulant(1)", fm.Circulant(c), c.reshape(1, 1)),
             ("Toeplitz(1x1)", fm.Toeplitz(c, np.zeros(0, np.complex64)), c.reshape(1, 1)),
             ("Diag(1)", fm.Diag(c), c.reshape(1, 1)),
             ("Permutation(1)", fm.Permutation(np.array([0])), np.ones((1, 1)))]
    for name, op, ref in cases:
        say("size one", name)
        x = (rng.standard_normal((1, 3)) + 1j * rng.standard_normal((1, 3))).astype(np.complex64)
        y = op.forward(x)
        assert y.shape == (1, 3)
        assert accept.max_rel_error(y, ref.astype(np.complex128) @ x) < 1e-6, name
        z = op.backward(x)
        assert accept.max_rel_error(z, ref.astype(np.complex128).conj().T @ x) < 1e-6, name
        v = op.forward(x[:, 0])  # 1-D input gives 1-D output
        assert v.shape == (1,)


def test_vector_input_and_integer_exactness(fm):
    rng = np.random.default_rng(13)
    H = fm.Hadamard(10)
    x = rng.integers(-100, 100, 1024).astype(np.int16)
    y = H.forward(x)
    assert y.shape == (1024,) and y.dtype == np.int16
    np.testing.assert_array_equal(y, dense.hadamard(10, np.int64).dot(x.astype(np.int64))
                                  .astype(np.int16))
    P = fm.Permutation(rng.permutation(1000))
    xi = rng.integers(-2 ** 31, 2 ** 31 - 1, (1000, 3), dtype=np.int64).astype(np.int32)
    np.testing.assert_array_equal(P.forward(xi), xi[P.sigma])


@pytest.mark.parametrize("logn", [22, 24])
def test_large_radix_lengths(fm, logn):
    """Four-step lengths 2^22 (2048 x 2048) and 2^24 (4096 x 4096) against the
    complex128 FFT."""
    n = 1 << logn
    say("large radix", n)
    rng = np.random.default_rng(logn)
    x = (rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))).astype(np.complex64)
    F = fm.Fourier(n)
    y = F.forward(x)
    assert accept.max_rel_error(y, fast64.fourier(x)) < 1e-5
    z = F.backward(x)
    assert accept.max_rel_error(z, n * np.fft.ifft(x.astype(np.complex128), axis=0)) < 1e-5


def test_largest_radix_length_sampled_bins(fm):
    """2^27 = 8192 x 16384: the 16384-point pass (the longest tile); checked
    on sampled output bins by direct FP64 summation with the phase reduced
    exactly in int64."""
    n = 1 << 27
    say("largest radix", n)
    rng = np.random.default_rng(27)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    F = fm.Fourier(n)
    bins = [0, 1, 12345, n // 2 + 7, n - 1]
    y = F.forward(x)
    ref = fast64.fourier_bins(x, bins)
    scale = np.abs(ref).max()
    assert np.abs(y[bins] - ref).max() / scale < 1e-5
    del y
    z = F.backward(x)
    refz = fast64.fourier_bins(x, bins[:3], adjoint=True)
    assert np.abs(z[bins[:3]] - refz).max() / np.abs(refz).max() < 1e-5


def test_large_circulant_and_toeplitz(fm):
    rng = np.random.default_rng(21)
    n = 300000  # embedding L = 2^20: three passes
    c = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    x = (rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))).astype(np.complex64)
    say("large circulant", n)
    C = fm.Circulant(c)
    assert accept.max_rel_error(C.forward(x), fast64.circulant(c, x)) < 1e-5
    m = 200000
    t = (rng.standard_normal(m + n - 1) + 1j * rng.standard_normal(m + n - 1)).astype(np.complex64)
    say("large toeplitz", m, n)
    T = fm.Toeplitz(t[:m], t[m:])
    assert accept.max_rel_error(T.forward(x), fast64.toeplitz(t, m, x)) < 1e-5
    g = (rng.standard_normal((m, 2)) + 1j * rng.standard_normal((m, 2))).astype(np.complex64)
    assert accept.max_rel_error(T.backward(g), fast64.toeplitz(t, m, g, adjoint=True)) < 1e-5


def test_error_paths_raise(fm):
    F = fm.Fourier(64)
    say("error paths")
    with pytest.raises(ValueError):
        F.forward(np.zeros((63, 2), np.complex64))
    with pytest.raises(ValueError):
        F.forward(np.zeros((64, 2, 2), np.complex64))
    with pytest.raises(ValueError):
        fm.Hadamard(4).backward(np.zeros(15))
    # lengths beyond the two-level engine (2^28 complex64) are rejected by the
    # plan, before any launch
    with pytest.raises((NotImplementedError, ValueError, MemoryError)):
        fm.Fourier(1 << 30).plan_info()

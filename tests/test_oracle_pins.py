"""Pins for the CPU oracle (runs without a GPU).

Each oracle function is checked against something that does not restate its
own formula: a library routine (scipy.linalg / numpy), a closed form or
invariant, brute force on tiny inputs, or a statement in the reference's own
test definitions.  A plausible mistake (wrong sign, transposed operand, index
off by one, dropped term) fails at least one of these.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import accept, cpu_path, dense, fast64, fwht_c, heuristics

RNG = np.random.default_rng(20261202)


def crandn(*shape, rng=RNG):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# ------------------------------------------------------------------ Fourier
@pytest.mark.parametrize("n", [1, 2, 3, 8, 35, 127, 128])
def test_dense_fourier_vs_numpy_fft(n):
    x = crandn(n, 3)
    np.testing.assert_allclose(dense.fourier(n) @ x, np.fft.fft(x, axis=0),
                               rtol=1e-12, atol=1e-12 * n)


@pytest.mark.parametrize("n", [5, 64, 97])
def test_dense_fourier_invariants(n):
    F = dense.fourier(n)
    # F^H F = N I (Fourier.pyx:188-189) and F symmetric
    np.testing.assert_allclose(F.conj().T @ F, n * np.eye(n), atol=1e-9 * n)
    np.testing.assert_allclose(F, F.T, atol=1e-12)
    # impulse -> all ones; F F x = N x[(-n) mod N]
    e0 = np.zeros(n); e0[0] = 1
    np.testing.assert_allclose(F @ e0, np.ones(n), atol=1e-12)
    x = crandn(n)
    np.testing.assert_allclose(F @ (F @ x), n * x[np.mod(-np.arange(n), n)],
                               atol=1e-9 * n)


def test_fast64_fourier_matches_direct_bins_large():
    n = 1 << 14
    x = crandn(n, 2)
    bins = [0, 1, 7, 4095, n // 2, n - 1]
    y = fast64.fourier(x)
    yb = fast64.fourier_bins(x, bins)
    np.testing.assert_allclose(y[bins], yb, rtol=1e-10, atol=1e-9)
    ya = fast64.fourier(x, adjoint=True)
    yba = fast64.fourier_bins(x, bins, adjoint=True)
    np.testing.assert_allclose(ya[bins], yba, rtol=1e-10, atol=1e-9)
    # Parseval: ||F x||^2 = N ||x||^2
    np.testing.assert_allclose(np.linalg.norm(y, axis=0) ** 2,
                               n * np.linalg.norm(x, axis=0) ** 2, rtol=1e-10)


@pytest.mark.parametrize("n", [35, 127, 1000])
def test_fast64_fourier_vs_dense(n):
    x = crandn(n, 2)
    F = dense.fourier(n)
    np.testing.assert_allclose(fast64.fourier(x), F @ x, atol=1e-9 * n)
    np.testing.assert_allclose(fast64.fourier(x, adjoint=True), F.conj().T @ x,
                               atol=1e-9 * n)


# ----------------------------------------------------------------- Circulant
@pytest.mark.parametrize("n", [1, 2, 7, 41, 64])
def test_dense_circulant_vs_scipy(n):
    c = crandn(n)
    np.testing.assert_array_equal(dense.circulant(c), sla.circulant(c))


@pytest.mark.parametrize("n,L", [(7, 7), (7, 13), (7, 16), (41, 96), (64, 64),
                                 (64, 128), (100, 256)])
def test_fast64_circulant_vs_dense(n, L):
    c = crandn(n)
    x = crandn(n, 3)
    C = sla.circulant(c)
    np.testing.assert_allclose(fast64.circulant(c, x, L=L), C @ x, atol=1e-10 * n)
    np.testing.assert_allclose(fast64.circulant(c, x, adjoint=True, L=L),
                               C.conj().T @ x, atol=1e-10 * n)


# ------------------------------------------------------------------ Toeplitz
@pytest.mark.parametrize("m,n", [(4, 3), (4, 5), (4, 41), (1, 1), (1, 6), (6, 1),
                                 (17, 9)])
def test_dense_toeplitz_vs_scipy(m, n):
    vecC = crandn(m)
    vecR = crandn(n - 1)
    # vecR is indexed backwards: the first row is [c0, vecR[-1], ..., vecR[0]]
    ref = sla.toeplitz(vecC, np.r_[vecC[0], vecR[::-1]])
    np.testing.assert_array_equal(dense.toeplitz(vecC, vecR), ref)


@pytest.mark.parametrize("m,n,L", [(4, 3, 6), (4, 5, 8), (4, 41, 64), (17, 9, 25),
                                   (300, 500, 1024), (1, 1, 1)])
def test_fast64_toeplitz_vs_dense(m, n, L):
    vecC = crandn(m)
    vecR = crandn(n - 1)
    T = sla.toeplitz(vecC, np.r_[vecC[0], vecR[::-1]])
    t = np.hstack([vecC, vecR])
    x = crandn(n, 2)
    yv = crandn(m, 2)
    np.testing.assert_allclose(fast64.toeplitz(t, m, x, L=L), T @ x, atol=1e-10 * L)
    np.testing.assert_allclose(fast64.toeplitz(t, m, yv, adjoint=True, L=L),
                               T.conj().T @ yv, atol=1e-10 * L)


# ------------------------------------------------------------------ Hadamard
@pytest.mark.parametrize("order", [0, 1, 2, 3, 6, 10])
def test_dense_hadamard_vs_scipy(order):
    np.testing.assert_array_equal(dense.hadamard(order), sla.hadamard(1 << order))


@pytest.mark.parametrize("order", [1, 4, 6, 10])
@pytest.mark.parametrize("dtype", [np.int8, np.int16, np.int32, np.int64])
def test_hadamard_int_wraps_like_c(order, dtype):
    n = 1 << order
    info = np.iinfo(dtype)
    x = RNG.integers(info.min, info.max, size=(n, 3), endpoint=True).astype(dtype)
    H = sla.hadamard(n).astype(object)
    exact = (H.dot(x.astype(object)))
    bits = np.dtype(dtype).itemsize * 8
    wrapped = np.array([[((int(v) + (1 << (bits - 1))) % (1 << bits)) - (1 << (bits - 1))
                         for v in row] for row in exact], dtype=dtype)
    np.testing.assert_array_equal(fast64.hadamard(x), wrapped)
    np.testing.assert_array_equal(fwht_c.hadamard(x), wrapped)


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex64, np.complex128])
def test_fwht_c_float(dtype):
    n = 1 << 9
    x = RNG.standard_normal((n, 2)).astype(dtype)
    if np.iscomplexobj(x):
        x = x + 1j * RNG.standard_normal((n, 2)).astype(dtype)
    y = fwht_c.hadamard(x)
    assert y.dtype == x.dtype
    ref = sla.hadamard(n).astype(np.float64) @ x.astype(np.complex128)
    err = accept.max_rel_error(y, ref)
    real = np.float32 if dtype in (np.float32, np.complex64) else np.float64
    assert err < 50 * np.finfo(real).eps
    # integer-valued floats are exact in any order -> bit-exact vs int64 path
    xi = RNG.integers(-8, 9, size=(n, 2))
    np.testing.assert_array_equal(fwht_c.hadamard(xi.astype(dtype)).real,
                                  fast64.hadamard(xi.astype(np.int64)))


def test_hadamard_involution():
    n = 1 << 12
    x = RNG.integers(-8, 9, size=(n, 2)).astype(np.int64)
    np.testing.assert_array_equal(fast64.hadamard(fast64.hadamard(x)), n * x)


# ---------------------------------------------------------------------- Kron
def test_dense_kron_index_formula():
    A = RNG.standard_normal((3, 3))
    B = crandn(4, 4)
    K = dense.kron(A, B)
    for i1 in range(3):
        for i2 in range(4):
            for j1 in range(3):
                for j2 in range(4):
                    assert K[i1 * 4 + i2, j1 * 4 + j2] == A[i1, j1] * B[i2, j2]


def test_fast64_kron_modewise_vs_dense():
    dims = (4, 6, 5)
    mats = [crandn(d, d) for d in dims]
    x = crandn(int(np.prod(dims)), 3)
    ops = [lambda v, M=M: M @ v for M in mats]
    np.testing.assert_allclose(fast64.kron_apply(ops, dims, x),
                               dense.kron(*mats) @ x, atol=1e-10)
    # a Fourier factor through fast64.fourier gives the same as its dense matrix
    ops[0] = lambda v: fast64.fourier(v)
    mats[0] = dense.fourier(4)
    np.testing.assert_allclose(fast64.kron_apply(ops, dims, x),
                               dense.kron(*mats) @ x, atol=1e-10)


# ---------------------------------------------------------------- heuristics
@pytest.mark.parametrize("n,val", [(7, 63), (16, 176), (96, 1728), (41, 1763),
                                   (35, 525), (72, 1368), (127, 16383),
                                   (256, 5376), (1 << 20, (1 << 20) * 51)])
def test_fft_complexity_hand_values(n, val):
    # hand-evaluated from core/cmath.pyx:156-214 (e.g. 72 = 4*2*9 and the
    # trial division stops at ii*ii < nn, so 9 costs 10: 72*(5+3+10+1))
    assert float(heuristics.fft_complexity(n)) == float(val)


def test_circulant_padding_first_at_41():
    # Circulant.pyx:436-437: "7 is just any number that causes no padding",
    # "41 is the first size bluestein is faster for"
    assert heuristics.circulant_embedding_length(7) == 7
    for n in range(1, 41):
        assert heuristics.circulant_embedding_length(n) == n, n
    L = heuristics.circulant_embedding_length(41)
    assert L >= 81 and L != 41


def test_fourier_path_choice():
    # Fourier.pyx:256-257: "127 is a size for which bluestein is faster"
    assert heuristics.fourier_bluestein_length(127) > 0
    assert heuristics.fourier_bluestein_length(127, optimize=False) == 0
    assert heuristics.fourier_bluestein_length(1 << 20) == 0
    assert heuristics.fourier_bluestein_length(1000003) == 2097152


@pytest.mark.parametrize("order", [1, 2, 13, 64, 65, 81, 1000, 2 * 1000003 - 1])
def test_optimal_size_properties(order):
    L = heuristics.find_optimal_fft_size(order, 4)
    assert L >= order
    r = L
    for p in (2, 3):
        while r % p == 0:
            r //= p
    assert r == 1  # maxStage=4 only allows factors 2, 3 and 4


# ------------------------------------------------------------------ cpu path
@pytest.mark.parametrize("n", [35, 127, 128])
def test_cpu_path_fourier(n):
    x = crandn(n, 2)
    f = cpu_path.FourierCPU(n)
    F = dense.fourier(n)
    np.testing.assert_allclose(f.forward(x), F @ x, atol=1e-9 * n)
    np.testing.assert_allclose(f.backward(x), F.conj().T @ x, atol=1e-9 * n)


def test_cpu_path_fourier_complex64_dtype():
    # numpy 2.x keeps complex64 through np.fft: the radix path returns it
    x = crandn(64, 2).astype(np.complex64)
    assert cpu_path.FourierCPU(64).forward(x).dtype == np.complex64
    # the Bluestein path computes in a complex128 buffer (Fourier.pyx:216-220)
    x127 = crandn(127, 1).astype(np.complex64)
    assert cpu_path.FourierCPU(127).forward(x127).dtype == np.complex128


@pytest.mark.parametrize("n", [7, 41, 64])
def test_cpu_path_circulant(n):
    c = crandn(n)
    x = crandn(n, 2)
    op = cpu_path.circulant(c)
    C = sla.circulant(c)
    np.testing.assert_allclose(op.forward(x), C @ x, atol=1e-10 * n)
    np.testing.assert_allclose(op.backward(x), C.conj().T @ x, atol=1e-10 * n)


@pytest.mark.parametrize("m,n", [(4, 3), (4, 41), (30, 20)])
def test_cpu_path_toeplitz(m, n):
    vecC = crandn(m)
    vecR = crandn(n - 1)
    T = sla.toeplitz(vecC, np.r_[vecC[0], vecR[::-1]])
    op = cpu_path.toeplitz(vecC, vecR)
    x = crandn(n, 2)
    yv = crandn(m, 2)
    np.testing.assert_allclose(op.forward(x), T @ x, atol=1e-10 * (m + n))
    np.testing.assert_allclose(op.backward(yv), T.conj().T @ yv, atol=1e-10 * (m + n))


def test_acceptance_rule_rejects_planted_error():
    n = 64
    ref = crandn(n, 2)
    tol = accept.tolerance(n, 3, [np.complex64])
    ok, err, _ = accept.passes(ref + 0.1 * tol * np.abs(ref).max(), ref, n, 3,
                               [np.complex64])
    assert ok
    ok, err, _ = accept.passes(ref + 10 * tol * np.abs(ref).max(), ref, n, 3,
                               [np.complex64])
    assert not ok


# ------------------------------------------------------ multi-level Circulant / Toeplitz
def _shift_power(rows, cols, d, p):
    """rows x cols 0/1 matrix with ones where (i - j) mod d == p."""
    i = np.arange(rows)[:, None]
    j = np.arange(cols)[None, :]
    return (np.mod(i - j, d) == p).astype(float)


@pytest.mark.parametrize("shape", [(3, 4), (3, 4, 5), (2, 1, 3)])
def test_dense_circulant_multilevel_vs_kron_sum(shape):
    # C = sum_p t[p] * kron_k S_{d_k}^{p_k}: an expansion over shift matrices,
    # independent of the block placement the oracle restates
    t = crandn(*shape)
    expect = np.zeros((int(np.prod(shape)),) * 2, dtype=complex)
    for p in np.ndindex(*shape):
        term = np.ones((1, 1))
        for d, pk in zip(shape, p):
            term = np.kron(term, _shift_power(d, d, d, pk))
        expect += t[p] * term
    np.testing.assert_allclose(dense.circulant_multilevel(t), expect, atol=1e-13)


@pytest.mark.parametrize("shape", [(3, 4), (4, 2, 5)])
def test_dense_circulant_multilevel_diagonalised_by_kron_dft(shape):
    # K C K^-1 = diag(fftn(t)) with K = kron of DFT matrices (C order)
    t = crandn(*shape)
    K = np.ones((1, 1))
    for d in shape:
        K = np.kron(K, np.fft.fft(np.eye(d), axis=0))
    C = dense.circulant_multilevel(t)
    lhs = K @ C @ np.linalg.inv(K)
    np.testing.assert_allclose(lhs, np.diag(np.fft.fftn(t).ravel()), atol=1e-10)


def test_dense_multilevel_reduce_to_one_level():
    c = crandn(9)
    np.testing.assert_array_equal(dense.circulant_multilevel(c), sla.circulant(c))
    t = crandn(11)
    np.testing.assert_array_equal(dense.toeplitz_multilevel(t, [4]),
                                  dense.toeplitz_tensor(t, 4))
    # a level of size 1 is a Kronecker factor of [t0]
    t2 = crandn(1, 7)
    np.testing.assert_array_equal(dense.toeplitz_multilevel(t2, [1, 3]),
                                  dense.toeplitz_tensor(t2[0], 3))


@pytest.mark.parametrize("shape,split", [((3, 3, 41), (2, 1, 4)), ((4, 5), (2, 3)),
                                         ((5, 3), (5, 1))])
def test_dense_toeplitz_multilevel_vs_kron_sum(shape, split):
    t = crandn(*shape)
    rows = [int(s) for s in split]
    cols = [d - s + 1 for d, s in zip(shape, split)]
    expect = np.zeros((int(np.prod(rows)), int(np.prod(cols))), dtype=complex)
    for p in np.ndindex(*shape):
        term = np.ones((1, 1))
        for d, r, c, pk in zip(shape, rows, cols, p):
            term = np.kron(term, _shift_power(r, c, d, pk))
        expect += t[p] * term
    np.testing.assert_allclose(dense.toeplitz_multilevel(t, split), expect, atol=1e-13)


def test_dense_toeplitz_multilevel_embeds_in_circulant():
    # a 2-level Toeplitz is the leading block of the 2-level circulant of its
    # defining tensor with zeros inserted at the split points
    t = crandn(4, 5)
    split = (2, 3)
    rows = [2, 3]
    cols = [3, 3]
    pad = np.zeros((8, 8), dtype=complex)
    ii = np.r_[0:2, 8 - 2:8]
    jj = np.r_[0:3, 8 - 2:8]
    pad[np.ix_(ii, jj)] = t
    C = dense.circulant_multilevel(pad)
    r_idx = (np.arange(rows[0])[:, None] * 8 + np.arange(rows[1])[None, :]).ravel()
    c_idx = (np.arange(cols[0])[:, None] * 8 + np.arange(cols[1])[None, :]).ravel()
    np.testing.assert_allclose(C[np.ix_(r_idx, c_idx)], dense.toeplitz_multilevel(t, split),
                               atol=1e-13)


# ------------------------------------------------------ Product / Sum / Partial
def test_dense_product_entrywise_and_scalar_precision():
    rng = np.random.default_rng(31)
    a = rng.integers(-3, 4, (5, 7)).astype(np.int16)
    b = rng.integers(-3, 4, (7, 5)).astype(np.int32)
    c = rng.integers(-3, 4, (5, 7)).astype(np.int16)
    s = np.complex64(1.3 - 0.7j)
    P = dense.product([a, b, c], scalar=s)
    # brute force: triple sum in exact integers, then one multiply by the
    # scalar's exact complex128 value
    exact = np.zeros((5, 7), dtype=np.int64)
    for i in range(5):
        for j in range(7):
            exact[i, j] = sum(int(a[i, p]) * int(b[p, q]) * int(c[q, j])
                              for p in range(7) for q in range(5))
    expect = exact * np.complex128(s)
    assert P.dtype == np.complex128
    np.testing.assert_allclose(P, expect, rtol=1e-15, atol=1e-12)
    # a scalar of 1 with one factor is that factor
    np.testing.assert_array_equal(dense.product([a]), a)


def test_dense_sum_and_partial():
    rng = np.random.default_rng(32)
    mats = [rng.integers(-5, 6, (6, 4)) for _ in range(3)]
    S = dense.sum_(mats)
    assert all(S[i, j] == mats[0][i, j] + mats[1][i, j] + mats[2][i, j]
               for i in range(6) for j in range(4))
    A = np.arange(48).reshape(6, 8)
    rows, cols = np.array([4, 0, 5]), np.array([True, False, True, True, False, False, True, False])
    Pa = dense.partial(A, rows=rows, cols=cols)
    assert Pa.shape == (3, 4)
    assert Pa[0, 0] == A[4, 0] and Pa[2, 3] == A[5, 6] and Pa[1, 1] == A[0, 2]
    np.testing.assert_array_equal(dense.partial(A), A)


# ------------------------------------------------------------ Permutation / Diag
def test_dense_permutation_brute_force():
    # P x = x[sigma] (Permutation.pyx:121-122) checked entry by entry, P^T is
    # the inverse (backward uses tau = argsort(sigma), :124-125)
    rng = np.random.default_rng(33)
    sigma = rng.permutation(9)
    P = dense.permutation(sigma)
    x = rng.standard_normal(9)
    y = [sum(P[i, j] * x[j] for j in range(9)) for i in range(9)]
    assert all(y[i] == x[sigma[i]] for i in range(9))
    np.testing.assert_array_equal(P.T @ P, np.eye(9))
    tau = np.argsort(sigma)
    np.testing.assert_array_equal(P.T @ x, x[tau])


def test_dense_diag_and_cpu_hadamard():
    d = crandn(6)
    D = dense.diag(d)
    x = crandn(6)
    np.testing.assert_allclose(D @ x, d * x, rtol=1e-15, atol=0)
    assert np.count_nonzero(D - np.diag(np.diag(D))) == 0
    # the timed CPU Hadamard leg equals the dense Sylvester matrix product
    xi = RNG.integers(-8, 9, (64, 3)).astype(np.int32)
    np.testing.assert_array_equal(cpu_path.hadamard_forward(xi),
                                  (sla.hadamard(64).astype(np.int64) @ xi).astype(np.int32))


# ------------------------------------------------------------ shrinkage solvers
from oracle import algorithms as oalg  # noqa: E402


def test_soft_threshold_closed_form():
    # x m/(m+a) with m = max(|x|-a, 0) equals x (1 - a/|x|) for |x| > a, else 0
    x = np.array([-3.0, -0.5, 0.0, 0.2, 0.5, 2.0])
    a = 0.5
    expect = np.where(np.abs(x) > a, x * (1 - a / np.maximum(np.abs(x), 1e-300)), 0.0)
    np.testing.assert_allclose(oalg.soft_threshold(x, a), expect, rtol=1e-15, atol=0)
    z = np.array([3 + 4j, 0.1 + 0.1j])
    np.testing.assert_allclose(oalg.soft_threshold(z, 1.0), [z[0] * 0.8, 0], rtol=1e-15)


def test_ista_identity_reaches_shrunk_measurement():
    # A = I: sigma = 1, L = 1, one step gives soft(b, lambda/2), a fixed point
    b = np.array([[2.0], [-0.1], [0.7]])
    res, x, _ = oalg.ista(np.eye(3), b, num_lambda=1.0, num_max_steps=5)
    np.testing.assert_allclose(x, oalg.soft_threshold(b, 0.5))


def test_ista_diagonal_converges_to_separable_minimiser():
    # min sum (d_i x_i - b_i)^2 + lam |x_i|: x_i = soft(d_i b_i, lam/2) / d_i^2
    d = np.array([2.0, 1.0, 0.5, 1.5])
    b = np.array([3.0, -0.2, 1.0, -2.0])
    lam = 0.4
    _, x, _ = oalg.ista(np.diag(d), b.astype(np.float64), num_lambda=lam, num_max_steps=4000)
    expect = oalg.soft_threshold(d * b, lam / 2) / d ** 2
    np.testing.assert_allclose(x[:, 0], expect, atol=1e-10)
    _, xf, _ = oalg.fista(np.diag(d), b.astype(np.float64), num_lambda=lam, num_max_steps=4000)
    np.testing.assert_allclose(xf[:, 0], expect, atol=1e-10)


def test_ista_objective_decreases_and_fista_betas():
    rng = np.random.default_rng(34)
    A = rng.standard_normal((20, 40))
    b = rng.standard_normal(20)
    lam = 0.3
    L = 1.0 / oalg.largest_singular_value(A) ** 2
    x = np.zeros(40)
    prev = np.inf
    for _ in range(50):
        x = oalg.soft_threshold(x - L * A.T @ (A @ x - b), L * lam * 0.5)
        obj = np.sum((A @ x - b) ** 2) + lam * np.sum(np.abs(x))
        assert obj <= prev + 1e-12
        prev = obj
    # t_{k+1} = (1 + sqrt(1 + 4 t_k^2)) / 2 from t_0 = 1: beta_0 = 0, beta_k -> 1
    bet = oalg.fista_betas(200)
    assert bet[0] == 0.0 and np.all(np.diff(bet) > 0) and bet[-1] < 1
    t = 1.0
    for k in range(5):
        t1 = (1 + np.sqrt(1 + 4 * t * t)) / 2
        assert abs(bet[k] - (t - 1) / t1) < 1e-15
        t = t1
    # the oracle's sigma is numpy's SVD value
    assert abs(oalg.largest_singular_value(np.diag([3.0, -5.0, 1.0])) - 5.0) < 1e-14


def test_stela_identity_one_step_and_stop():
    # A = I: d = 1, z = -b, one step gives soft(b, lambda) (gamma = 1), after
    # which z - clamp(z - x) = 0 for every column: the loop returns at step 1
    b = np.array([[2.0, 0.1], [-0.1, -3.0], [0.7, 0.0]])
    x, steps, gamma = oalg.stela(np.eye(3), b, num_lambda=0.5, num_max_steps=10)
    np.testing.assert_allclose(x, oalg.soft_threshold(b, 0.5), rtol=1e-15, atol=1e-16)
    assert steps == 1
    np.testing.assert_allclose(gamma, [1.0, 1.0])


def test_stela_unitary_complex_one_step():
    # a unitary A (normalised DFT): ||A x - b|| = ||x - A^H b||, so the first
    # step is soft(A^H b, lambda); a dropped conjugate or transpose fails
    n = 16
    F = np.fft.fft(np.eye(n)) / np.sqrt(n)
    rng = np.random.default_rng(51)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x, _, gamma = oalg.stela(F, b, num_lambda=0.3, num_max_steps=1)
    np.testing.assert_allclose(x[:, 0], oalg.soft_threshold(F.conj().T @ b, 0.3), atol=1e-14)
    np.testing.assert_allclose(gamma, [1.0], rtol=1e-13)
    # a fixed point after that (the stopping value clamps real and imaginary
    # parts separately, so for complex data it need not reach zero)
    x5, _, _ = oalg.stela(F, b, num_lambda=0.3, num_max_steps=5)
    np.testing.assert_allclose(x5, x, atol=1e-13)


def test_stela_single_column_closed_form():
    # one column a with ||a|| >= 1: the exact line search lands on the
    # minimiser of 0.5 ||a x - b||^2 + lambda |x|, soft(a^H b, lambda) / ||a||^2,
    # in the first step (gamma = 1 / ||a||^4 along bx = soft(a^H b) ||a||^2)
    rng = np.random.default_rng(52)
    a = (rng.standard_normal(7) + 1j * rng.standard_normal(7)).reshape(7, 1)
    b = rng.standard_normal(7) + 1j * rng.standard_normal(7)
    lam = 0.2
    x, _, gamma = oalg.stela(a, b, num_lambda=lam, num_max_steps=1)
    na2 = np.sum(np.abs(a) ** 2)
    g = np.vdot(a[:, 0], b)
    np.testing.assert_allclose(x[0, 0], oalg.soft_threshold(g, lam) / na2, rtol=1e-13)
    np.testing.assert_allclose(gamma[0], 1 / na2 ** 2, rtol=1e-13)


def test_stela_objective_monotone():
    # the step minimises an upper bound of 0.5||Ax-b||^2 + lambda||x||_1
    # that is tight at gamma = 0, so the objective never increases
    rng = np.random.default_rng(53)
    for cplx in (False, True):
        A = rng.standard_normal((12, 30))
        b = rng.standard_normal((12, 2))
        if cplx:
            A = A + 1j * rng.standard_normal((12, 30))
            b = b + 1j * rng.standard_normal((12, 2))
        prev = oalg.stela_objective(A, b, np.zeros((30, 2)), 0.2)
        for s in range(1, 25):
            x, _, _ = oalg.stela(A, b, num_lambda=0.2, num_max_steps=s, num_max_error=0.0)
            obj = oalg.stela_objective(A, b, x, 0.2)
            assert np.all(obj <= prev + 1e-12), (cplx, s)
            prev = obj


def test_block_compositions_against_scipy_and_numpy():
    import scipy.linalg
    rng = np.random.default_rng(71)
    a = rng.standard_normal((3, 4))
    b = (rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2)))
    c = rng.integers(-3, 4, (5, 1)).astype(np.int16)
    np.testing.assert_array_equal(dense.block_diag([a, b, c]), scipy.linalg.block_diag(a, b, c))
    g = [[a, rng.standard_normal((3, 2))], [rng.standard_normal((1, 4)), b[:1, :]]]
    np.testing.assert_array_equal(dense.blocks(g), np.block(g))
    # a block diagonal matrix is the grid with zero off-diagonal blocks
    z = [[a, np.zeros((3, 2))], [np.zeros((2, 4)), b]]
    np.testing.assert_array_equal(dense.block_diag([a, b]), dense.blocks(z))


def test_lfsr_docstring_sequence_and_periods():
    # the sequence printed in LFSRCirculant's docstring (LFSRCirculant.pyx:
    # 95-112) for polynomial 0b11001, start 0b1010
    s = dense.lfsr_sequence(0b11001, 0b1010)
    np.testing.assert_array_equal(
        s, [+1, -1, +1, -1, -1, +1, +1, -1, +1, +1, +1, -1, -1, -1, -1])
    # the periods its test table states (LFSRCirculant.pyx:440-452)
    assert dense.lfsr_sequence(0x19, 0xD).size == 15
    assert dense.lfsr_sequence(0x17, 0x1).size == 7
    # the benchmark's polynomials are maximal: period 2^c - 1
    # (LFSRCirculant.pyx:460-480)
    db = [0x3, 0x7, 0xB, 0x19, 0x25, 0x61, 0x89, 0x1C3, 0x211, 0x481, 0xB03, 0x1C11]
    for c, p in enumerate(db, start=1):
        if c == 1:
            continue
        assert dense.lfsr_sequence(p, 0xFFFFFFFF).size == 2 ** c - 1, hex(p)


def test_lfsr_circulant_structure():
    # circulant: C[i, j] = s[(i - j) mod N]; an m-sequence has the two-valued
    # autocorrelation N, -1, so C^T C = (N + 1) I - J
    C = dense.lfsr_circulant(0x19, 0xD).astype(np.int64)
    s = dense.lfsr_sequence(0x19, 0xD).astype(np.int64)
    n = s.size
    i = np.arange(n)
    np.testing.assert_array_equal(C, s[(i[:, None] - i[None, :]) % n])
    np.testing.assert_array_equal(C.T @ C, (n + 1) * np.eye(n, dtype=np.int64) - 1)


def test_cpu_path_multilevel_circulant_matches_dense():
    rng = np.random.default_rng(81)
    t = rng.standard_normal((4, 8)) + 1j * rng.standard_normal((4, 8))
    op = cpu_path.CirculantMultilevelCPU(t)
    M = dense.circulant_multilevel(t)
    x = rng.standard_normal((32, 3)) + 1j * rng.standard_normal((32, 3))
    np.testing.assert_allclose(op.forward(x), M @ x, atol=1e-12)
    np.testing.assert_allclose(op.backward(x), M.conj().T @ x, atol=1e-12)


def _lfsr_states_taps(poly, start):
    order, mask = 0, 1
    while (~mask & poly) > mask:
        mask <<= 1
        order += 1
    mask = 1 << order
    s = st0 = start & (mask - 1)
    states = []
    while True:
        states.append(s)
        if bin(s & poly).count("1") & 1:
            s |= mask
        s >>= 1
        if s == st0:
            break
    taps, b = [], 1
    for _ in states:
        taps.append(b)
        b <<= 1
        if b & mask:
            b ^= poly | mask
    return order, states, taps


@pytest.mark.parametrize("poly,start", [(0x19, 0xD), (0x17, 0x1), (0x25, 0x1F)])
def test_cpu_path_lfsr_hadamard_factorisation(poly, start):
    # scatter -> Hadamard -> gather equals the dense circulant of the
    # sequence (the GF(2)-linearity identity the reference's _core relies on)
    order, states, taps = _lfsr_states_taps(poly, start)
    op = cpu_path.LFSRCirculantCPU(order, states, taps)
    C = dense.lfsr_circulant(poly, start).astype(np.int64)
    x = np.random.default_rng(poly).integers(-3, 4, (C.shape[0], 4)).astype(np.int64)
    np.testing.assert_array_equal(op.forward(x), C @ x)
    np.testing.assert_array_equal(op.backward(x), C.T @ x)


def test_omp_exact_recovery_reference_test():
    # test/algorithm.py:23-70: 40 x 100 Gaussian, 6-sparse: support and
    # values recovered (rtol 1e-12 as the reference's own test states)
    rng = np.random.default_rng(95)
    A = rng.standard_normal((40, 100))
    x = np.zeros(100)
    supp = rng.choice(100, 6, replace=False)
    x[supp] = rng.standard_normal(6)
    y, sup = oalg.omp(A, A @ x, 6)
    assert sorted(sup[:, 0]) == sorted(supp)
    np.testing.assert_allclose(y[:, 0], x, rtol=1e-12, atol=1e-15)


def test_omp_least_squares_and_orthogonal_residual():
    # after s steps the coefficients are the least-squares fit on the
    # selected columns and the residual is orthogonal to them; the first
    # pick is the largest normalised correlation
    rng = np.random.default_rng(96)
    for cplx in (False, True):
        A = rng.standard_normal((30, 50)) * rng.uniform(0.5, 3, 50)
        b = rng.standard_normal((30, 2))
        if cplx:
            A = A + 1j * rng.standard_normal((30, 50))
            b = b + 1j * rng.standard_normal((30, 2))
        y, sup = oalg.omp(A, b, 5)
        for l in range(2):
            S = sup[:, l]
            assert sup[0, l] == np.argmax(np.abs(A.conj().T @ b[:, l]) / np.linalg.norm(A, axis=0))
            xs = np.linalg.lstsq(A[:, S], b[:, l], rcond=None)[0]
            np.testing.assert_allclose(y[S, l], xs, rtol=1e-10, atol=1e-12)
            r = b[:, l] - A @ y[:, l]
            assert np.max(np.abs(A[:, S].conj().T @ r)) < 1e-10 * np.linalg.norm(b[:, l])


@pytest.mark.parametrize("kind,m,n", [("circulant", 7, 7), ("circulant", 16, 16),
                                      ("toeplitz", 4, 5), ("toeplitz", 6, 3),
                                      ("toeplitz", 1, 4)])
def test_conv_generator_grad_brute_force_and_fd(kind, m, n):
    """fast64.conv_generator_grad against (1) the brute-force sum over the
    matrix entries that hold each generator element (index maps of
    Circulant.pyx:418-423 / Toeplitz.pyx:719-732, read off the dense oracle
    matrices) and (2) FP64 central differences of the loss
    sum |A(gen) x - y|^2 (PyTorch convention dL/dRe + i dL/dIm)."""
    rng = np.random.default_rng(hash((kind, m, n)) % 2**32)
    k = 3
    d = n if kind == "circulant" else m + n - 1
    gen = rng.standard_normal(d) + 1j * rng.standard_normal(d)
    x = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    y = rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))

    def mat(t):
        if kind == "circulant":
            return dense.circulant(t)
        return dense.toeplitz(t[:m], t[m:])

    # which generator element sits at (i, j): dense matrix of indices
    idx = np.rint(mat(np.arange(d, dtype=np.float64)).real).astype(int)
    g = 2 * (mat(gen) @ x - y)
    brute = np.zeros(d, dtype=np.complex128)
    for i in range(m):
        for j in range(n):
            brute[idx[i, j]] += np.sum(g[i] * np.conj(x[j]))
    got = fast64.conv_generator_grad(kind, m, n, x, g)
    np.testing.assert_allclose(got, brute, rtol=0, atol=1e-12 * np.abs(brute).max())

    def loss(t):
        return float(np.sum(np.abs(mat(t) @ x - y) ** 2))

    h = 1e-6
    fd = np.zeros(d, dtype=np.complex128)
    for q in range(d):
        for unit in (1.0, 1j):
            tp, tm = gen.copy(), gen.copy()
            tp[q] += h * unit
            tm[q] -= h * unit
            fd[q] += unit * (loss(tp) - loss(tm)) / (2 * h)
    np.testing.assert_allclose(got, fd, rtol=0, atol=1e-6 * np.abs(fd).max())


def test_kron_cpu_path_matches_dense_kron():
    """oracle.cpu_path.KronCPU (the reference's term loop with its C / Fortran
    order reshapes) against np.kron of the dense factors, forward and
    backward, on Fourier(4) x Circulant(8) x dense 5 x 5."""
    import scipy.linalg as sla
    from oracle import cpu_path
    rng = np.random.default_rng(1)
    A = rng.standard_normal((5, 5)).astype(np.float32)
    c = (rng.standard_normal(8) + 1j * rng.standard_normal(8)).astype(np.complex64)
    K = cpu_path.KronCPU([cpu_path.FourierCPU(4), cpu_path.circulant(c), cpu_path.DenseCPU(A)],
                         (4, 8, 5))
    ref = np.kron(np.kron(dense.fourier(4), sla.circulant(c.astype(np.complex128))),
                  A.astype(np.float64))
    x = rng.standard_normal((160, 3)) + 1j * rng.standard_normal((160, 3))
    y = ref @ x
    assert np.abs(K.forward(x) - y).max() <= 1e-12 * np.abs(y).max()
    yb = ref.conj().T @ x
    assert np.abs(K.backward(x) - yb).max() <= 1e-12 * np.abs(yb).max()

"""GPU parity of every operator on the hot path against the CPU oracle.

Integer and float Hadamard results are compared bit for bit with the C
restatement of the reference's loop (oracle/fwht.c).  Floating-point
transforms are compared with FP64 evaluations (oracle.fast64 / oracle.dense)
by max|y - y64| / max|y64| (inspect/test.py:186-188), bounded by:
  complex64  : 1e-5 (FP32 butterflies, FP64-generated twiddles/chirps)
  complex128 : 1e-12
"""
import sys

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import accept, dense, fast64, fwht_c

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
TOL64 = 1e-12


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def crandn(rng, *shape, dtype=np.complex128):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)


@pytest.fixture(scope="module")
def fm():
    import fastmat_amd
    from fastmat_amd import _lib
    _lib.load()
    return fastmat_amd


def relerr(y, ref):
    return accept.max_rel_error(y, ref)


# ------------------------------------------------------------------ Hadamard
@pytest.mark.parametrize("dtype", [np.int8, np.int16, np.int32, np.int64])
@pytest.mark.parametrize("order", [1, 2, 3, 4, 6, 9, 12, 15])
def test_hadamard_int_bitexact(fm, dtype, order):
    rng = np.random.default_rng(order * 10 + np.dtype(dtype).itemsize)
    info = np.iinfo(dtype)
    x = rng.integers(info.min, info.max, size=(1 << order, 3), endpoint=True).astype(dtype)
    say("hadamard int", np.dtype(dtype).name, order)
    y = fm.Hadamard(order).forward(x)
    assert y.dtype == x.dtype
    np.testing.assert_array_equal(y, fwht_c.hadamard(x))
    np.testing.assert_array_equal(fm.Hadamard(order).backward(x), fwht_c.hadamard(x))


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex64, np.complex128])
@pytest.mark.parametrize("order", [1, 5, 10, 13, 16])
def test_hadamard_float_bitexact(fm, dtype, order):
    rng = np.random.default_rng(100 + order)
    x = rng.standard_normal((1 << order, 2)).astype(dtype)
    if np.iscomplexobj(x):
        x = x + 1j * rng.standard_normal((1 << order, 2)).astype(dtype)
    say("hadamard float", np.dtype(dtype).name, order)
    y = fm.Hadamard(order).forward(x)
    assert y.dtype == x.dtype
    # same stage order as the reference loop -> identical IEEE results
    np.testing.assert_array_equal(y, fwht_c.hadamard(x))


def test_hadamard_order20_int32_config(fm):
    rng = np.random.default_rng(20261202 + 2)
    x = rng.integers(-8, 9, size=(1 << 20, 4)).astype(np.int32)
    say("hadamard order 20 int32 k=4")
    y = fm.Hadamard(20).forward(x)
    np.testing.assert_array_equal(y, fast64.hadamard(x.astype(np.int64)).astype(np.int32))
    # H H = N I
    np.testing.assert_array_equal(fm.Hadamard(20).forward(y), x.astype(np.int32) << 20)


def test_hadamard_vector_and_strided_input(fm):
    rng = np.random.default_rng(7)
    x = rng.integers(-2, 3, size=(64, 5)).astype(np.int16)
    say("hadamard 1-D and C-order input")
    np.testing.assert_array_equal(fm.Hadamard(6).forward(x[:, 1]), fwht_c.hadamard(x[:, 1]))
    xc = np.ascontiguousarray(x)
    np.testing.assert_array_equal(fm.Hadamard(6).forward(xc[:, ::2]),
                                  fwht_c.hadamard(np.asfortranarray(xc[:, ::2])))


# ------------------------------------------------------------------- Fourier
FOURIER_SIZES = [1, 2, 4, 8, 16, 64, 512, 4096, 8192, 16384, 65536, 3, 5, 35, 127, 1000, 4097,
                 5000, 40000]


@pytest.mark.parametrize("n", FOURIER_SIZES)
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_fourier_parity(fm, n, dtype):
    rng = np.random.default_rng(n)
    k = 3
    x = crandn(rng, n, k, dtype=dtype)
    say("fourier", n, np.dtype(dtype).name)
    F = fm.Fourier(n)
    y = F.forward(x)
    assert y.dtype == dtype
    tol = TOL32 if dtype == np.complex64 else TOL64
    assert relerr(y, fast64.fourier(x)) < tol
    yb = F.backward(x)
    assert relerr(yb, fast64.fourier(x, adjoint=True)) < tol


def test_fourier_2pow20_radix(fm):
    rng = np.random.default_rng(20261202 + 1)
    n = 1 << 20
    x = crandn(rng, n, 2, dtype=np.complex64)
    say("fourier 2^20 c64 k=2")
    F = fm.Fourier(n)
    y = F.forward(x)
    e = relerr(y, fast64.fourier(x))
    say("  max rel err", e)
    assert e < TOL32
    assert relerr(F.backward(x), fast64.fourier(x, adjoint=True)) < TOL32


def test_fourier_bluestein_prime_1000003(fm):
    rng = np.random.default_rng(20261202 + 1)
    n = 1000003
    x = crandn(rng, n, 1, dtype=np.complex64)
    say("fourier 1000003 c64 bluestein")
    F = fm.Fourier(n)
    L, blue, passes = F.plan_info()
    assert blue and L == 2097152
    y = F.forward(x)
    e = relerr(y, fast64.fourier(x))
    say("  max rel err", e)
    assert e < TOL32
    # an FFT-independent check on sampled bins (exact integer phase reduction)
    bins = [0, 1, 2, 999, 500001, n - 1]
    np.testing.assert_allclose(y[bins, 0], fast64.fourier_bins(x, bins)[:, 0],
                               rtol=0, atol=TOL32 * np.abs(y).max())


def test_fourier_dense_and_1d(fm):
    F = fm.Fourier(35)
    say("fourier dense reference")
    np.testing.assert_allclose(F.reference(), dense.fourier(35), atol=1e-12)
    x = np.arange(35, dtype=np.float32)
    y = F.forward(x)
    assert y.shape == (35,) and y.dtype == np.complex64
    assert relerr(y, np.fft.fft(x.astype(np.float64))) < TOL32


# ------------------------------------------------------- Circulant / Toeplitz
@pytest.mark.parametrize("n", [1, 2, 7, 41, 64, 1000, 4096, 5000, 20000])
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_circulant_parity(fm, n, dtype):
    rng = np.random.default_rng(1000 + n)
    c = crandn(rng, n, dtype=dtype)
    x = crandn(rng, n, 4, dtype=dtype)
    say("circulant", n, np.dtype(dtype).name)
    C = fm.Circulant(c)
    tol = TOL32 if dtype == np.complex64 else TOL64
    y = C.forward(x)
    assert y.dtype == dtype
    assert relerr(y, fast64.circulant(c, x)) < tol
    assert relerr(C.backward(x), fast64.circulant(c, x, adjoint=True)) < tol


def test_circulant_config1_fp64_dense_oracle(fm):
    rng = np.random.default_rng(20261202 + 0)
    n = 4096
    c = crandn(rng, n)
    x = crandn(rng, n, 32)
    say("circulant 4096 c128 k=32 vs dense")
    C = fm.Circulant(c)
    Cd = sla.circulant(c)
    y = C.forward(x)
    assert y.dtype == np.complex128
    assert relerr(y, Cd @ x) < TOL64
    assert relerr(C.backward(x), Cd.conj().T @ x) < TOL64


@pytest.mark.parametrize("m,n", [(4, 3), (4, 5), (4, 41), (1, 1), (1, 7), (7, 1), (300, 500),
                                 (5000, 3000), (3000, 9000)])
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_toeplitz_parity(fm, m, n, dtype):
    rng = np.random.default_rng(m * 7 + n)
    vecC = crandn(rng, m, dtype=dtype)
    vecR = crandn(rng, n - 1, dtype=dtype)
    x = crandn(rng, n, 3, dtype=dtype)
    yv = crandn(rng, m, 3, dtype=dtype)
    say("toeplitz", m, n, np.dtype(dtype).name)
    T = fm.Toeplitz(vecC, vecR)
    t = np.hstack([vecC, vecR])
    tol = TOL32 if dtype == np.complex64 else TOL64
    assert relerr(T.forward(x), fast64.toeplitz(t, m, x)) < tol
    assert relerr(T.backward(yv), fast64.toeplitz(t, m, yv, adjoint=True)) < tol


def test_toeplitz_tensor_split_matches_vectors(fm):
    rng = np.random.default_rng(3)
    t = rng.standard_normal(9)
    T1 = fm.Toeplitz(t)                 # square: split (9+1)//2 = 5
    T2 = fm.Toeplitz(t[:5], t[5:])
    assert T1.shape == T2.shape == (5, 5)
    np.testing.assert_array_equal(T1.reference(), T2.reference())
    np.testing.assert_array_equal(T1.reference(), dense.toeplitz(t[:5], t[5:]))
    x = rng.standard_normal((5, 2))
    say("toeplitz tensor/split")
    np.testing.assert_allclose(T1.forward(x), T2.forward(x), rtol=0, atol=1e-12)


# ----------------------------------------------------------------- dense/MFMA
@pytest.mark.parametrize("m,n,k", [(37, 53, 5), (128, 128, 33), (1, 1, 1), (200, 70, 130)])
def test_dense_mfma(fm, m, n, k):
    rng = np.random.default_rng(m + n + k)
    A = rng.standard_normal((m, n)).astype(np.float32)
    M = fm.Matrix(A)
    say("dense mfma", m, n, k)
    for xdt in (np.float32, np.complex64):
        x = rng.standard_normal((n, k)).astype(xdt)
        if xdt == np.complex64:
            x = x + 1j * rng.standard_normal((n, k)).astype(np.float32)
        y = M.forward(x)
        assert y.dtype == xdt
        assert relerr(y, A.astype(np.float64) @ x.astype(np.complex128)) < 1e-5
        xb = rng.standard_normal((m, k)).astype(xdt)
        assert relerr(M.backward(xb), A.T.astype(np.float64) @ xb.astype(np.complex128)) < 1e-5
    Ac = (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))).astype(np.complex64)
    Mc = fm.Matrix(Ac)
    x = crandn(rng, n, k, dtype=np.complex64)
    assert relerr(Mc.forward(x), Ac.astype(np.complex128) @ x) < 1e-5
    xb = crandn(rng, m, k, dtype=np.complex64)
    assert relerr(Mc.backward(xb), Ac.conj().T.astype(np.complex128) @ xb) < 1e-5


@pytest.mark.parametrize("dtype", [np.int8, np.int16, np.int32, np.int64])
def test_dense_integer_wraps_like_numpy_dot(fm, dtype):
    """Integer dense products wrap modulo 2^bits exactly as numpy's integer
    ndarray.dot (Matrix.pyx:1842) does, overflow included."""
    rng = np.random.default_rng(np.dtype(dtype).itemsize)
    info = np.iinfo(dtype)
    m, n, k = 70, 129, 5
    A = rng.integers(info.min, info.max, size=(m, n), endpoint=True).astype(dtype)
    x = rng.integers(info.min, info.max, size=(n, k), endpoint=True).astype(dtype)
    xb = rng.integers(info.min, info.max, size=(m, k), endpoint=True).astype(dtype)
    M = fm.Matrix(A)
    say("dense integer", np.dtype(dtype).name)
    y = M.forward(x)
    assert y.dtype == dtype
    with np.errstate(over="ignore"):
        ref = A.dot(x)
        refb = A.T.dot(xb)
    np.testing.assert_array_equal(y, ref)
    np.testing.assert_array_equal(M.backward(xb), refb)
    # the products do overflow the width (the test exercises the wrap)
    big = np.abs(A.astype(object).dot(x.astype(object))).max()
    assert big > info.max


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_dense_fp64(fm, dtype):
    rng = np.random.default_rng(9)
    m, n, k = 97, 61, 7
    A = rng.standard_normal((m, n)).astype(dtype)
    x = rng.standard_normal((n, k)).astype(dtype)
    if np.dtype(dtype).kind == "c":
        A = A + 1j * rng.standard_normal((m, n))
        x = x + 1j * rng.standard_normal((n, k))
    xb = x[:, :3].copy()
    xb = np.resize(xb, (m, 3))
    M = fm.Matrix(A)
    say("dense fp64", np.dtype(dtype).name)
    assert relerr(M.forward(x), A @ x) < 1e-13
    assert relerr(M.backward(xb), A.conj().T @ xb) < 1e-13
    # mixed: int32 matrix with complex64 data promotes to complex128
    Ai = rng.integers(-5, 6, size=(m, n)).astype(np.int32)
    xc = crandn(rng, n, 2, dtype=np.complex64)
    yi = fm.Matrix(Ai).forward(xc)
    assert yi.dtype == np.complex128
    assert relerr(yi, Ai @ xc.astype(np.complex128)) < 1e-13


# ------------------------------------------------------------------- Kron
def test_kron_small_mixed(fm):
    rng = np.random.default_rng(11)
    A = rng.standard_normal((5, 5)).astype(np.float32)
    c = crandn(rng, 8, dtype=np.complex64)
    K = fm.Kron(fm.Fourier(4), fm.Circulant(c), fm.Matrix(A))
    ref = np.kron(np.kron(dense.fourier(4), sla.circulant(c)), A)
    x = crandn(rng, 160, 3, dtype=np.complex64)
    say("kron F4 x C8 x M5")
    assert K.dtype == np.complex64
    assert relerr(K.forward(x), ref @ x) < 1e-5
    assert relerr(K.backward(x), ref.conj().T @ x) < 1e-5


def test_kron_config4_shape_small_k(fm):
    rng = np.random.default_rng(20261202 + 3)
    A = (rng.standard_normal((128, 128)) / np.sqrt(128)).astype(np.float32)
    c = rng.standard_normal(128)
    K = fm.Kron(fm.Fourier(64), fm.Circulant(c.astype(np.complex64)), fm.Matrix(A))
    x = crandn(rng, 1 << 20, 2, dtype=np.complex64)
    say("kron config4 k=2")
    ops = [lambda v: fast64.fourier(v), lambda v: sla.circulant(c) @ v,
           lambda v: A.astype(np.float64) @ v]
    ref = fast64.kron_apply(ops, (64, 128, 128), x)
    assert relerr(K.forward(x), ref) < 1e-5
    opsb = [lambda v: fast64.fourier(v, adjoint=True), lambda v: sla.circulant(c).T @ v,
            lambda v: A.T.astype(np.float64) @ v]
    refb = fast64.kron_apply(opsb, (64, 128, 128), x)
    assert relerr(K.backward(x), refb) < 1e-5


@pytest.mark.parametrize("lead,k", [((), 3), ((4,), 1), ((2, 8), 5)])
def test_kron_fused_circulant_dense_tail(fm, lead, k):
    """Kron products ending in Circulant(128) x dense 128 x 128 float32 run
    those two modes as one pass (fm_kron_conv_dense_apply): forward and
    backward against the FP64 mode-wise oracle, with a complex generator and
    0, 1 or 2 leading factors (Fourier, Hadamard)."""
    import torch
    from fastmat_amd import kron as fkron
    rng = np.random.default_rng(40 + len(lead))
    A = (rng.standard_normal((128, 128)) / np.sqrt(128)).astype(np.float32)
    c = crandn(rng, 128, dtype=np.complex64)
    heads = [fm.Fourier(lead[0])] if len(lead) >= 1 else []
    if len(lead) == 2:
        heads = [fm.Hadamard(1), fm.Fourier(lead[1])]
    K = fm.Kron(*(heads + [fm.Circulant(c), fm.Matrix(A)])) if heads else \
        fm.Kron(fm.Circulant(c), fm.Matrix(A))
    assert fkron._fused_tail(K._content, torch.complex64)
    dims = tuple((2,) if len(lead) == 2 else ()) + tuple(lead[-1:]) + (128, 128)
    n = int(np.prod(dims))
    x = crandn(rng, n, k, dtype=np.complex64)
    say("kron fused tail", dims, "k", k)
    C64 = sla.circulant(c.astype(np.complex128))
    lead_ops = []
    if len(lead) == 2:
        lead_ops.append(lambda v: dense.hadamard(1) @ v)
    if lead:
        lead_ops.append(lambda v: fast64.fourier(v))
    ops = lead_ops + [lambda v: C64 @ v, lambda v: A.astype(np.float64) @ v]
    assert relerr(K.forward(x), fast64.kron_apply(ops, dims, x)) < 1e-5
    lead_b = []
    if len(lead) == 2:
        lead_b.append(lambda v: dense.hadamard(1) @ v)
    if lead:
        lead_b.append(lambda v: fast64.fourier(v, adjoint=True))
    opsb = lead_b + [lambda v: C64.conj().T @ v, lambda v: A.T.astype(np.float64) @ v]
    got = K.backward(x)
    want = fast64.kron_apply(opsb, dims, x)
    err = relerr(got, want)
    if err >= 1e-5:
        # name the wrong 128 x 128 slabs (rows // 16384) and columns
        bad = np.abs(got - want) > 1e-4 * np.abs(want).max()
        rows, cols = np.nonzero(bad)
        print("[diag] wrong elements", rows.size, "slabs", sorted(set((rows // 16384).tolist()))[:20],
              "columns", sorted(set(cols.tolist())), flush=True)
    assert err < 1e-5


# ---------------------------------------------- composition and elementwise
def test_composition(fm):
    rng = np.random.default_rng(5)
    n = 32
    d = crandn(rng, n, dtype=np.complex64)
    sigma = rng.permutation(n)
    c = crandn(rng, n, dtype=np.complex64)
    F = fm.Fourier(n)
    P = fm.Permutation(sigma)
    Dg = fm.Diag(d)
    C = fm.Circulant(c)
    H = fm.Hadamard(5)
    x = crandn(rng, n, 4, dtype=np.complex64)
    Fd = dense.fourier(n)
    Pd = np.eye(n)[sigma]
    Hd = sla.hadamard(n)
    Cd = sla.circulant(c)
    cases = [
        ("product", F * Dg * P, Fd @ np.diag(d) @ Pd),
        ("sum", F + C, Fd + Cd),
        ("difference", C - H, Cd - Hd),
        ("scaled", (2.5 - 1j) * F / 3.0, (2.5 - 1j) * Fd / 3.0),
        ("hermitian", C.H * F, Cd.conj().T @ Fd),
        ("transpose", F.T * C.T, Fd.T @ Cd.T),
        ("conj", C.conj, Cd.conj()),
        ("partial", fm.Partial(F, rows=np.arange(5, 20), cols=np.arange(n) % 3 == 0),
         Fd[5:20][:, np.arange(n) % 3 == 0]),
        ("nested", fm.Product(fm.Sum(F, C), fm.Product(Dg, H)), (Fd + Cd) @ np.diag(d) @ Hd),
    ]
    for name, op, ref in cases:
        say("composition", name)
        xin = crandn(rng, op.numCols, 3, dtype=np.complex64)
        assert relerr(op.forward(xin), ref @ xin) < 1e-5, name
        xb = crandn(rng, op.numRows, 3, dtype=np.complex64)
        assert relerr(op.backward(xb), ref.conj().T @ xb) < 1e-5, name
        np.testing.assert_allclose(op.reference(), ref, rtol=1e-5, atol=1e-5)


def test_permutation_exact(fm):
    rng = np.random.default_rng(9)
    n = 1000
    sigma = rng.permutation(n)
    P = fm.Permutation(sigma)
    x = rng.integers(-100, 100, size=(n, 7)).astype(np.int64)
    say("permutation")
    np.testing.assert_array_equal(P.forward(x), x[sigma])
    np.testing.assert_array_equal(P.backward(x), x[np.argsort(sigma)])


def test_torch_inputs_stay_on_device(fm):
    import torch
    rng = np.random.default_rng(1)
    x = torch.from_numpy(crandn(rng, 256, 4, dtype=np.complex64)).cuda()
    say("torch in/out")
    F = fm.Fourier(256)
    y = F.forward(x)
    assert isinstance(y, torch.Tensor) and y.is_cuda and y.shape == (256, 4)
    assert relerr(y.cpu().numpy(), fast64.fourier(x.cpu().numpy())) < TOL32

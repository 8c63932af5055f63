"""The reference's own randomized transform tests, replayed on the GPU path.

Mirrors inspect/test.py:318-368 (forward vs reference() @ x, backward vs
reference().T.conj() @ x, input not modified) with each class's _getTest()
parameters (Fourier.pyx:249-270, Circulant.pyx:428-467 including the
multi-level shapes, Toeplitz.pyx:737-800 including the multi-level case,
Hadamard.pyx:251-267, Kron.pyx:356-385, Product.pyx:279-317,
Sum.pyx:189-219, Partial.pyx:306-362, Permutation.pyx:132-145): inputs drawn
like arrTestDist (inspect/common.py:369-422), columns {1, 5}
(inspect/test.py:969), all 8 dtypes (:970), F / C / strided layouts (:973),
acceptance by compareResults (inspect/test.py:104-202: max-norm relative
error within 5*2*eps*(2*sqrt(maxDim))**TOL_POWER, and the result dtype equal
to promote(instance dtype, input dtype)).  The dense matrices come from the
oracle (oracle/dense.py), not from the classes' own reference().
"""
import sys

import numpy as np
import pytest

from oracle import accept, dense

pytestmark = pytest.mark.gpu

ALLTYPES = [np.int8, np.int16, np.int32, np.int64, np.float32, np.float64, np.complex64,
            np.complex128]
FEWTYPES = [np.int32, np.float32, np.complex64]
RNG = np.random.default_rng(20261202 + 99)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def test_dist(shape, dtype, rng=RNG):
    """arrTestDist: ints from {-2,-1,1,2}, floats |x| in [1,2] with random
    sign, the largest element bumped (inspect/common.py:369-422)."""
    dtype = np.dtype(dtype)

    def draw(s):
        return rng.uniform(1.0, 2.0, size=s) * rng.choice([-1, 1], s)

    if np.issubdtype(dtype, np.integer):
        r = rng.choice([-2, -1, 1, 2], shape).astype(dtype)
    elif np.issubdtype(dtype, np.complexfloating):
        r = (draw(shape) + 1j * draw(shape)).astype(dtype)
    else:
        r = draw(shape).astype(dtype)
    idx = np.unravel_index(np.abs(r).argmax(), r.shape)
    if np.issubdtype(dtype, np.integer):
        r[idx] += np.sign(r[idx])
    else:
        r[idx] *= 1.5
    return r


test_dist.__test__ = False  # not a pytest test


def layouts(x):
    yield "F", np.asfortranarray(x)
    yield "C", np.ascontiguousarray(x)
    big = np.zeros((x.shape[0], 3 * x.shape[1]), dtype=x.dtype)
    big[:, ::3] = x
    yield "strided", big[:, ::3]


def check_transforms(fm_mat, ref, tol_power, min_eps=0.0, name=""):
    """Forward and backward of fm_mat against the dense `ref` for every input
    dtype, column count and layout; returns the number of queries run."""
    n_rows, n_cols = ref.shape
    ref64 = ref.astype(np.complex128)
    queries = 0
    for dt in ALLTYPES:
        for k in (1, 5):
            for adjoint in (False, True):
                n_in = n_rows if adjoint else n_cols
                x0 = test_dist((n_in, k), dt)
                for lname, x in layouts(x0):
                    if k == 1 and lname != "F":
                        continue
                    keep = x.copy()
                    out = fm_mat.backward(x) if adjoint else fm_mat.forward(x)
                    assert np.array_equal(x, keep), "input modified"
                    expect = (ref64.conj().T if adjoint else ref64) @ x0.astype(np.complex128)
                    exp_dtype = np.promote_types(
                        np.promote_types(fm_mat.dtype, np.int8), x0.dtype)
                    assert out.dtype == exp_dtype, (name, dt, out.dtype, exp_dtype)
                    max_dim = max(ref.shape + fm_mat.shape)
                    eps_types = [x0.dtype, np.promote_types(np.int8, fm_mat.dtype),
                                 np.promote_types(np.int8, ref.dtype)]
                    ok, err, tol = accept.passes(out, expect, max_dim, tol_power, eps_types,
                                                 min_eps)
                    assert ok, (name, np.dtype(dt).name, k, lname, adjoint, err, tol)
                    queries += 1
    return queries


@pytest.fixture(scope="module")
def fm():
    import fastmat_amd
    from fastmat_amd import _lib
    _lib.load()
    return fastmat_amd


@pytest.mark.parametrize("n", [35, 127])
@pytest.mark.parametrize("optimize", [False, True])
def test_fourier(fm, n, optimize):
    say("reference suite Fourier", n, optimize)
    F = fm.Fourier(n, optimize=optimize)
    check_transforms(F, dense.fourier(n), 3.0, name="Fourier")


@pytest.mark.parametrize("shape", [(7,), (41,), (3, 4), (3, 4, 5)])
@pytest.mark.parametrize("mtype", FEWTYPES)
def test_circulant(fm, shape, mtype):
    say("reference suite Circulant", shape, np.dtype(mtype).name)
    c = test_dist(shape, mtype)
    C = fm.Circulant(c)
    check_transforms(C, dense.circulant_multilevel(c), 2.0, float(np.finfo(np.float64).eps),
                     name="Circulant")


@pytest.mark.parametrize("ncols", [3, 5, 41])
@pytest.mark.parametrize("th", FEWTYPES)
@pytest.mark.parametrize("tv", FEWTYPES)
def test_toeplitz(fm, ncols, th, tv):
    say("reference suite Toeplitz", 4, ncols, np.dtype(th).name, np.dtype(tv).name)
    vecH = test_dist((4,), th)
    vecV = test_dist((ncols - 1,), tv)
    T = fm.Toeplitz(vecH, vecV)
    check_transforms(T, dense.toeplitz(vecH, vecV), 2.0, float(np.finfo(np.float64).eps),
                     name="Toeplitz")


@pytest.mark.parametrize("optimize", [False, True])
@pytest.mark.parametrize("mtype", FEWTYPES)
def test_toeplitz_multilevel(fm, mtype, optimize):
    say("reference suite Toeplitz multi-level", np.dtype(mtype).name, optimize)
    shape, split = np.array([3, 3, 41]), np.array([2, 1, 4])
    tenT = test_dist(tuple(shape), mtype)
    T = fm.Toeplitz(tenT, split=split, optimize=optimize)
    check_transforms(T, dense.toeplitz_multilevel(tenT, split), 2.0,
                     float(np.finfo(np.float64).eps), name="Toeplitz ML")


@pytest.mark.parametrize("order", [4, 6])
def test_hadamard(fm, order):
    say("reference suite Hadamard", order)
    H = fm.Hadamard(order)
    check_transforms(H, dense.hadamard(order, np.int8), 1.0, name="Hadamard")


def test_permutation(fm):
    say("reference suite Permutation")
    sigma = RNG.permutation(35)
    P = fm.Permutation(sigma)
    check_transforms(P, dense.permutation(sigma), 1.0, name="Permutation")


@pytest.mark.parametrize("t1", ALLTYPES)
@pytest.mark.parametrize("t2", ALLTYPES)
def test_kron_dense_factors(fm, t1, t2):
    say("reference suite Kron", np.dtype(t1).name, np.dtype(t2).name)
    a1, a2, a3 = test_dist((5, 5), t1), test_dist((4, 4), t2), test_dist((3, 3), t1)
    K = fm.Kron(fm.Matrix(a1), fm.Matrix(a2), fm.Matrix(a3))
    check_transforms(K, dense.kron(a1, a2, a3), 4.0, name="Kron")


@pytest.mark.parametrize("ncols", [7, 5])
@pytest.mark.parametrize("t1", ALLTYPES)
@pytest.mark.parametrize("t2", FEWTYPES)
def test_product(fm, ncols, t1, t2):
    say("reference suite Product", ncols, np.dtype(t1).name, np.dtype(t2).name)
    for st in FEWTYPES:
        s = test_dist((1,), st)[0]
        a1, a2, a3 = test_dist((5, ncols), t1), test_dist((ncols, 5), t2), \
            test_dist((5, ncols), t1)
        P = fm.Product(s, fm.Matrix(a1), fm.Matrix(a2), fm.Matrix(a3))
        ref = dense.product([a1, a2, a3], scalar=s)
        check_transforms(P, ref, 3.0, name="Product")


@pytest.mark.parametrize("ncols", [17, 25])
@pytest.mark.parametrize("t1", ALLTYPES)
@pytest.mark.parametrize("t2", FEWTYPES)
def test_sum(fm, ncols, t1, t2):
    say("reference suite Sum", ncols, np.dtype(t1).name, np.dtype(t2).name)
    a1, a2, a3 = (test_dist((25, ncols), t1), test_dist((25, ncols), t2),
                  test_dist((25, ncols), t1))
    S = fm.Sum(fm.Matrix(a1), fm.Matrix(a2), fm.Matrix(a3))
    check_transforms(S, dense.sum_([a1, a2, a3]), 3.0, name="Sum")


SUBCOLS = [np.array([1, 2, 3, 11, 12]), np.array([6]),
           np.array([1, 1, 0, 0, 1, 0, 0, 1, 0, 1, 1, 0, 0, 1, 0, 1], dtype=bool)]
SUBROWS = [np.array([7, 8, 9, 13]), np.array([10]),
           np.array([0, 1, 1, 0, 0, 1, 0, 1, 1, 1, 0, 0, 1, 0, 0, 1], dtype=bool)]


@pytest.mark.parametrize("ci", range(3))
@pytest.mark.parametrize("ri", range(3))
def test_partial(fm, ci, ri):
    say("reference suite Partial", ci, ri)
    P = fm.Partial(fm.Hadamard(4), rows=SUBROWS[ri], cols=SUBCOLS[ci])
    ref = dense.partial(dense.hadamard(4, np.int8), SUBROWS[ri], SUBCOLS[ci])
    check_transforms(P, ref, 1.0, name="Partial")


@pytest.mark.parametrize("t1", ALLTYPES)
@pytest.mark.parametrize("t2", FEWTYPES)
def test_blockdiag(fm, t1, t2):
    """BlockDiag.pyx:192-230: Matrix(arr1), -2 Matrix(arr2), 2 Matrix(arr3),
    5 x 5 each, TOL_POWER 3."""
    say("reference suite BlockDiag", np.dtype(t1).name, np.dtype(t2).name)
    a1, a2, a3 = test_dist((5, 5), t1), test_dist((5, 5), t2), test_dist((5, 5), t1)
    B = fm.BlockDiag(fm.Matrix(a1), -2. * fm.Matrix(a2), 2. * fm.Matrix(a3))
    ref = dense.block_diag([a1, dense.product([a2], scalar=-2.), dense.product([a3], scalar=2.)])
    check_transforms(B, ref, 3.0, name="BlockDiag")


@pytest.mark.parametrize("t1", ALLTYPES)
@pytest.mark.parametrize("t2", FEWTYPES)
def test_blocks(fm, t1, t2):
    """Blocks.pyx:314-352: [[Matrix(arr1), Matrix(arr2)], [Matrix(arr3),
    Matrix(arr4)]], 4 x 4 each, types (t1, t2; t2, t1)."""
    say("reference suite Blocks", np.dtype(t1).name, np.dtype(t2).name)
    a1, a2, a3, a4 = (test_dist((4, 4), t1), test_dist((4, 4), t2), test_dist((4, 4), t2),
                      test_dist((4, 4), t1))
    B = fm.Blocks([[fm.Matrix(a1), fm.Matrix(a2)], [fm.Matrix(a3), fm.Matrix(a4)]])
    check_transforms(B, dense.blocks([[a1, a2], [a3, a4]]), 1.0, name="Blocks")


@pytest.mark.parametrize("c", [64, 100])
def test_block_benchmark_compositions(fm, c):
    """The reference's benchmark compositions (BlockDiag.pyx:232-252,
    Blocks.pyx:354-372): structured blocks, so the 1-D Circulant blocks of the
    BlockDiag run through one batch bound to slices of the caller's block and
    the others through their leading-dimension applies."""
    say("block benchmark compositions", c)
    rng = np.random.default_rng(c)
    c1, c2, d = rng.standard_normal(c), rng.standard_normal(c), rng.standard_normal(c)
    B = fm.BlockDiag(fm.Circulant(c1), fm.Circulant(c2), fm.Fourier(c), fm.Diag(d))
    ref = dense.block_diag([dense.circulant(c1), dense.circulant(c2), dense.fourier(c),
                            dense.diag(d)])
    check_transforms(B, ref, 3.0, name="BlockDiag-structured")
    G = fm.Blocks([[fm.Circulant(c1), fm.Circulant(c2)], [fm.Fourier(c), fm.Diag(d)]])
    ref = dense.blocks([[dense.circulant(c1), dense.circulant(c2)],
                        [dense.fourier(c), dense.diag(d)]])
    check_transforms(G, ref, 3.0, name="Blocks-structured")
    # mixed: Toeplitz and Hadamard blocks, nested BlockDiag, a Blocks term
    T = fm.Toeplitz(rng.standard_normal(5), rng.standard_normal(8))
    H = fm.Hadamard(3)
    N = fm.BlockDiag(T, fm.BlockDiag(H, fm.Circulant(c1)), G)
    refN = dense.block_diag([dense.toeplitz(T.tenT[:5], T.tenT[5:]), dense.hadamard(3),
                             dense.circulant(c1), ref])
    check_transforms(N, refN, 3.0, name="BlockDiag-nested")


@pytest.mark.parametrize("poly", [0x19, 0x17])
@pytest.mark.parametrize("start", [0xD, 0x1])
def test_lfsr_circulant(fm, poly, start):
    """LFSRCirculant.pyx:440-458: polynomials 0x19 (period 15) and 0x17
    (period 7), starts 0xD and 0x1."""
    say("reference suite LFSRCirculant", hex(poly), hex(start))
    L = fm.LFSRCirculant(poly, start)
    check_transforms(L, dense.lfsr_circulant(poly, start), 1.0, name="LFSRCirculant")


def test_lfsr_circulant_order_20(fm):
    """The benchmark's order-20 register (polynomial 0x120001, start
    0xFFFFFFFF, N = 2^20 - 1): integer data exactly against the circular
    convolution (float64 FFT, rounded: all sums are integers below 2^22),
    float32 within 1e-5 of it, forward and backward."""
    import torch
    L = fm.LFSRCirculant(0x120001, 0xFFFFFFFF)
    n = L.numRows
    assert n == 2 ** 20 - 1
    s = dense.lfsr_sequence(0x120001, 0xFFFFFFFF).astype(np.float64)
    S = np.fft.fft(s)
    rng = np.random.default_rng(77)
    x = rng.choice([-2, -1, 1, 2], (n, 3)).astype(np.int32)
    say("LFSRCirculant order 20")
    fwd = np.real(np.fft.ifft(S[:, None] * np.fft.fft(x.astype(np.float64), axis=0), axis=0))
    bwd = np.real(np.fft.ifft(np.conj(S)[:, None] * np.fft.fft(x.astype(np.float64), axis=0),
                              axis=0))
    y = L.forward(x)
    assert y.dtype == np.int32
    np.testing.assert_array_equal(y, np.rint(fwd).astype(np.int32))
    np.testing.assert_array_equal(L.backward(x), np.rint(bwd).astype(np.int32))
    xf = x.astype(np.float32)
    yf = L.forward(torch.from_numpy(np.ascontiguousarray(xf.T)).cuda().t())
    assert accept.max_rel_error(yf.cpu().numpy(), fwd) < 1e-5


@pytest.mark.parametrize("cdt", [np.complex64, np.complex128])
def test_two_level_plan(fm, cdt):
    """Two-level Circulant / Toeplitz large enough for the three-pass
    two-level plan: power-of-two levels against the FP64 2-D circular
    convolution, embedded levels against the oracle's dense matrices."""
    from fastmat_amd.circulant import _TwoLevelConv
    rng = np.random.default_rng(91)
    tol = 1e-5 if cdt == np.complex64 else 1e-12
    # power-of-two levels 64 x 256: no embedding
    t = (rng.standard_normal((64, 256)) + 1j * rng.standard_normal((64, 256))).astype(cdt)
    C = fm.Circulant(t)
    say("two-level plan pow2", np.dtype(cdt).name)
    x = rng.standard_normal((64 * 256, 3)) + 1j * rng.standard_normal((64 * 256, 3))
    x = x.astype(cdt)
    T = np.fft.fft2(t.astype(np.complex128))
    X3 = x.astype(np.complex128).T.reshape(3, 64, 256)
    fwd = np.fft.ifft2(T[None] * np.fft.fft2(X3)).reshape(3, -1).T
    bwd = np.fft.ifft2(np.conj(T)[None] * np.fft.fft2(X3)).reshape(3, -1).T
    assert accept.max_rel_error(C.forward(x), fwd) < tol
    assert accept.max_rel_error(C.backward(x), bwd) < tol
    assert isinstance(C._multilevel_chain(np.dtype(cdt)).content[0], _TwoLevelConv)
    # embedded levels: Circulant 40 x 100 -> 128 x 256, Toeplitz (63, 255)
    # split (32, 128) -> 64 x 256
    say("two-level plan embedded", np.dtype(cdt).name)
    t2 = (rng.standard_normal((40, 100)) + 1j * rng.standard_normal((40, 100))).astype(cdt)
    C2 = fm.Circulant(t2)
    M2 = dense.circulant_multilevel(t2.astype(np.complex128))
    x2 = (rng.standard_normal((4000, 2)) + 1j * rng.standard_normal((4000, 2))).astype(cdt)
    assert accept.max_rel_error(C2.forward(x2), M2 @ x2.astype(np.complex128)) < tol
    assert accept.max_rel_error(C2.backward(x2), M2.conj().T @ x2.astype(np.complex128)) < tol
    assert isinstance(C2._multilevel_chain(np.dtype(cdt)).content[0], _TwoLevelConv)
    t3 = (rng.standard_normal((63, 255)) + 1j * rng.standard_normal((63, 255))).astype(cdt)
    T3 = fm.Toeplitz(t3, split=np.array([32, 128]))
    M3 = dense.toeplitz_multilevel(t3.astype(np.complex128), np.array([32, 128]))
    x3 = (rng.standard_normal((M3.shape[1], 2)) +
          1j * rng.standard_normal((M3.shape[1], 2))).astype(cdt)
    assert accept.max_rel_error(T3.forward(x3), M3 @ x3.astype(np.complex128)) < tol
    g3 = (rng.standard_normal((M3.shape[0], 2)) +
          1j * rng.standard_normal((M3.shape[0], 2))).astype(cdt)
    assert accept.max_rel_error(T3.backward(g3), M3.conj().T @ g3.astype(np.complex128)) < tol

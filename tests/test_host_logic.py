"""Host-side logic of the package, checked without a GPU.

Covers the constructor behaviour the reference tests exactly
(test/classes.py:275-320 for Toeplitz), the dense closed forms behind
``reference()`` / ``array`` against the oracle, and the multi-level
embedding (padding per level, row/column selection) evaluated in numpy FP64
against the oracle's dense matrices.
"""
import warnings

import numpy as np
import pytest

from oracle import dense

RNG = np.random.default_rng(20261202 + 17)


def crandn(*shape):
    return RNG.standard_normal(shape) + 1j * RNG.standard_normal(shape)


@pytest.fixture(scope="module")
def fm():
    import fastmat_amd
    return fastmat_amd


def embedding_dense(op):
    """The multi-level chain Partial(Product(K^H, Diag(fftn(t')/L), K)) built
    from the operator's embedding description, in numpy FP64."""
    Ls, t, rows, cols = op._multi
    K = np.ones((1, 1))
    for L in Ls:
        K = np.kron(K, np.fft.fft(np.eye(L), axis=0))
    spec = np.fft.fftn(t).ravel() / np.prod(Ls)
    full = K.conj().T @ np.diag(spec) @ K
    return full[np.ix_(rows, cols)]


def test_toeplitz_constructor_checks(fm):
    # test/classes.py:275-320, dense comparisons through reference()
    arrC = RNG.standard_normal((4, 10, 10))
    arrR = RNG.standard_normal((6, 10, 10))
    tenT = np.vstack([arrC, arrR])
    split = arrC.shape[0]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", FutureWarning)
        T1A = fm.Toeplitz(arrC[:, 0, 0], arrR[:, 0, 0])
        np.testing.assert_array_equal(T1A.vecC, arrC[:, 0, 0])
        np.testing.assert_array_equal(T1A.vecR, arrR[:, 0, 0])
        with pytest.raises(ValueError):
            fm.Toeplitz(arrC[:, 0, 0], arrR[:, 0, 0], split=5)
        with pytest.raises(ValueError):
            fm.Toeplitz(arrC[:, :, 0], arrR[:, :, 0])
        with pytest.raises(ValueError):
            fm.Toeplitz()
        with pytest.raises(TypeError):
            fm.Toeplitz('abc')
        with pytest.raises(ValueError):
            fm.Toeplitz('abc', 'def')
        with pytest.raises(ValueError):
            fm.Toeplitz('abc', 'def', 'ghi')
        T1B = fm.Toeplitz(tenT[:, 0, 0], split=split)
        np.testing.assert_array_equal(T1B.tenT, tenT[:, 0, 0])
        np.testing.assert_array_equal(T1A[...], T1B[...])
        np.testing.assert_array_equal(T1A[...], T1A.reference())
        T2 = fm.Toeplitz(tenT[:, :, 0], split=2 * (split,))
        np.testing.assert_array_equal(T2[...], T2.reference())
        T3A = fm.Toeplitz(tenT, split=3 * (split,))
        T3B = fm.Toeplitz(tenT[:9, :9, :9])
        T3C = fm.Toeplitz(tenT, split=3 * (tenT.shape[0] - split,))
        for T in (T3A, T3B, T3C):
            np.testing.assert_array_equal(T[...], T.reference())
        np.testing.assert_array_equal(T3A.vecC, arrC)
        np.testing.assert_array_equal(T3A.vecR, arrR)
        np.testing.assert_array_equal(T3A.tenT, tenT)
        np.testing.assert_array_equal(T3B.tenT, tenT[:9, :9, :9])
        np.testing.assert_array_equal(T3C.tenT, tenT)
        with pytest.raises(ValueError):
            fm.Toeplitz(tenT)
        with pytest.raises(ValueError):
            fm.Toeplitz(tenT, split=np.zeros((3, 3)))
        with pytest.raises(ValueError):
            fm.Toeplitz(tenT, split=4 * (split,))
    # shapes of the multi-level cases
    assert T3A.shape == (4 ** 3, 7 ** 3)
    assert T3C.shape == (6 ** 3, 5 ** 3)
    # the 3-level reference equals the oracle's block construction
    np.testing.assert_allclose(T3B.reference(),
                               dense.toeplitz_multilevel(tenT[:9, :9, :9], (5, 5, 5)),
                               rtol=0, atol=0)


@pytest.mark.parametrize("shape", [(7,), (41,), (3, 4), (3, 4, 5), (8, 2)])
def test_circulant_reference_and_embedding(fm, shape):
    t = crandn(*shape)
    C = fm.Circulant(t)
    n = int(np.prod(shape))
    assert C.shape == (n, n)
    ref = dense.circulant_multilevel(t)
    np.testing.assert_array_equal(C.reference(), ref)
    if len(shape) > 1:
        Ls = C._multi[0]
        assert all((L & (L - 1)) == 0 for L in Ls)
        np.testing.assert_allclose(embedding_dense(C), ref, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("shape,split", [((3, 3, 41), (2, 1, 4)), ((4, 5), (2, 3)),
                                         ((9, 9, 9), (5, 5, 5)), ((8, 3), (8, 1))])
def test_toeplitz_multilevel_reference_and_embedding(fm, shape, split):
    t = crandn(*shape)
    T = fm.Toeplitz(t, split=np.array(split))
    ref = dense.toeplitz_multilevel(t, split)
    assert T.shape == ref.shape
    np.testing.assert_array_equal(T.reference(), ref)
    np.testing.assert_allclose(embedding_dense(T), ref, atol=1e-12 * np.abs(ref).max())


def test_multilevel_dtype_follows_generator(fm):
    assert fm.Circulant(RNG.standard_normal((3, 4)).astype(np.float32)).dtype == np.complex64
    assert fm.Circulant(RNG.standard_normal((3, 4))).dtype == np.complex128
    for it in (np.int8, np.int16, np.int32):
        T = fm.Toeplitz(np.ones((3, 5), it), split=np.array([2, 3]))
        assert T.dtype == np.promote_types(it, np.complex64)


@pytest.mark.parametrize("shape,rows,cols", [
    ((9,), (4,), (6,)), ((7,), (1,), (7,)), ((7,), (7,), (1,)),
    ((5, 7), (3, 2), (3, 6)), ((3, 3, 41), (2, 1, 4), (2, 3, 38)), ((4, 6), (4, 6), (4, 6)),
])
def test_multilevel_column_norms_match_dense(shape, rows, cols):
    # the window-sum closed form of the Toeplitz / Circulant column norms
    # against the norms of the dense multi-level matrix
    from fastmat_amd.circulant import _multilevel_col_norms, _multilevel_dense
    rng = np.random.default_rng(sum(shape))
    t = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    M = _multilevel_dense(t, rows, cols)
    np.testing.assert_allclose(_multilevel_col_norms(t, rows, cols),
                               np.linalg.norm(M, axis=0), rtol=1e-14)


def test_lfsr_host_sequences_and_errors():
    """LFSRCirculant's constructor runs the library's host-side register
    routine (no device work): periods, sequence and the reference's errors
    (LFSRCirculant.pyx:196-222)."""
    import fastmat_amd as fm
    L = fm.LFSRCirculant(0b11001, 0b1010)
    assert L.shape == (15, 15) and L.order == 4 and L.dtype == np.int8
    np.testing.assert_array_equal(L.vecC, dense.lfsr_sequence(0b11001, 0b1010))
    assert L.states.dtype == np.uint32 and L.states[0] == 0b1010
    np.testing.assert_array_equal(L.reference(), dense.lfsr_circulant(0b11001, 0b1010))
    assert fm.LFSRCirculant(0x17, 0xD).period == 7
    big = fm.LFSRCirculant(0x120001, 0xFFFFFFFF)
    assert big.period == 2 ** 20 - 1
    with pytest.raises(ValueError, match="order 1 to 31"):
        fm.LFSRCirculant(0x1, 0x1)
    with pytest.raises(ValueError, match="non-zero"):
        fm.LFSRCirculant(0x19, 0x10)
    with pytest.raises(ValueError, match="invalid sequence"):
        fm.LFSRCirculant(0x10, 0x3)
    with pytest.warns(FutureWarning):
        assert L.taps == 0b11001


def test_block_composition_constructor_errors():
    """BlockDiag.pyx:110-127 and Blocks.pyx:88-170 argument checks (host
    only)."""
    import fastmat_amd as fm
    M = fm.Matrix(np.eye(2))
    B = fm.BlockDiag(M, fm.Matrix(np.ones((3, 1), dtype=np.int16)))
    assert B.shape == (5, 3) and B.dtype == np.float64
    with pytest.raises(ValueError, match="at least one matrix"):
        fm.BlockDiag()
    with pytest.raises(ValueError, match="Only fastmat matrices"):
        fm.BlockDiag(M, np.eye(2))
    G = fm.Blocks([[M, fm.Matrix(np.zeros((2, 3)))], [fm.Matrix(np.ones((1, 2))),
                                                     fm.Matrix(np.ones((1, 3)))]])
    assert G.shape == (3, 5)
    with pytest.raises(ValueError, match="Not a nested list"):
        fm.Blocks(M)
    with pytest.raises(ValueError, match="Contains no matrices"):
        fm.Blocks([])
    with pytest.raises(TypeError, match="row\\(0\\)"):
        fm.Blocks([M])
    with pytest.raises(TypeError, match="not a fastmat Matrix"):
        fm.Blocks([[M, np.eye(2)]])
    with pytest.raises(ValueError, match="incompatible number of entries"):
        fm.Blocks([[M, M], [M]])
    with pytest.raises(ValueError, match="row height"):
        fm.Blocks([[M, fm.Matrix(np.ones((3, 2)))]])
    with pytest.raises(ValueError, match="column width"):
        fm.Blocks([[M, M], [M, fm.Matrix(np.ones((2, 3)))]])


def test_algorithm_parameter_handling():
    """Algorithm.pyx:28-176 / test/algorithm.py:23-90: abstract base,
    parameter updates, trace type, constructor type checks (host only)."""
    import fastmat_amd as fm
    import fastmat_amd.algorithms as fma
    with pytest.raises(NotImplementedError):
        fma.Algorithm()
    with pytest.raises(TypeError):
        fma.ISTA(np.eye(3))
    with pytest.raises(TypeError):
        fma.STELA(None)
    alg = fma.FISTA(fm.Matrix(np.eye(3)), numLambda=0.5)
    assert alg.numLambda == 0.5 and alg.numMaxSteps == 100
    with pytest.raises(AttributeError):
        alg.updateParameters(numBla=123)
    alg.updateParameters(numMaxSteps=7)
    assert alg.numMaxSteps == 7
    for bad in (None, (), {}, 1):
        with pytest.raises(TypeError):
            alg.trace = bad
    alg.trace = []
    s = fma.STELA(fm.Matrix(np.eye(3)))
    assert s.numMaxError == 1e-6
    np.testing.assert_allclose(s.softThreshold(np.array([2.0, -0.1]), 0.5), [1.5, 0.0])


def test_partial_selection_follows_numpy_indexing(fm):
    """The reference's selection check (Partial.pyx:141-166) accepts
    [-1, 1, 3] for 3 columns (length 3, index sum equal to arange's); numpy
    then raises IndexError when the selection is applied.  In-range negative
    indices wrap as numpy's do."""
    from fastmat_amd.partial import _device_index
    P = fm.Partial(fm.Eye(3), cols=[-1, 1, 3])
    assert list(P.colSelection) == [-1, 1, 3]
    with pytest.raises(IndexError):
        _device_index(P.colSelection, 3)
    # [2, -1, 2] also passes the reference's check; numpy reads -1 as 2
    P2 = fm.Partial(fm.Eye(3), rows=[2, -1, 2])
    assert list(_device_index(P2.rowSelection, 3)) == [2, 2, 2]
    np.testing.assert_array_equal(P2._reference(), np.eye(3)[[2, -1, 2], :])
    with pytest.raises(ValueError):
        fm.Partial(fm.Eye(3), cols=[0, 5])


@pytest.mark.skipif(not __import__("os").path.exists(
    __import__("os").path.join(__import__("os").path.dirname(__import__("os").path.dirname(
        __import__("os").path.abspath(__file__))), "fastmat_amd", "libfastmat_amd.so")),
    reason="library not built")
def test_hadamard_tile_index_mapping_is_bijective():
    """The Hadamard pass kernel's tile index formulas (strided stage-0 load,
    every stage's LDS exchange, the strided re-stage store), run on the host
    by the library itself: every (vector, element) slot of a tile is hit
    exactly once for 256-, 512- and 1024-thread tiles.  A tile of 1024
    threads whose shape is derived as for 512 (log2 of the thread count taken
    as 9, the mistake that made the 1024-thread strided variant give wrong
    int32 results) is caught."""
    import ctypes
    from fastmat_amd import _lib
    lib = _lib.load()
    bad = ctypes.c_int64()
    checked = 0
    for loge in (3, 4, 5):
        for nt in (256, 512, 1024):
            for logn in range(1, 15):
                _lib.check(lib.fm_wht_tile_check(logn, nt, nt, loge, ctypes.byref(bad)))
                assert bad.value in (0, -1), (logn, nt, loge, bad.value)
                checked += bad.value == 0
    assert checked > 60
    # the order-20 strided pass with 1024 threads (2^10 elements, 32 per thread)
    _lib.check(lib.fm_wht_tile_check(10, 1024, 1024, 5, ctypes.byref(bad)))
    assert bad.value == 0
    _lib.check(lib.fm_wht_tile_check(10, 1024, 512, 5, ctypes.byref(bad)))
    assert bad.value > 0

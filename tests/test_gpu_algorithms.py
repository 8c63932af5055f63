"""GPU tests of the device-resident shrinkage solvers (ISTA, FISTA) and of
largestSingularValue.

Oracle: oracle.algorithms (the reference's ISTA / FISTA loops restated in
numpy on the dense matrix, step size from numpy's SVD).  Results are held to
max|x - x_ref| / max|x_ref| <= 1e-4 for single-precision problems (the
iterate is computed in float32 on both sides; the order of roundings
differs) and <= 1e-9 for double precision.  The graph-replayed loop must
equal the eager loop (callbacks force the eager path) bit for bit.
"""
import sys

import numpy as np
import pytest

from oracle import accept, algorithms as oalg

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


def tol_for(dt):
    return 1e-4 if np.dtype(dt) in (np.dtype(np.float32), np.dtype(np.complex64)) else 1e-9


def sparse_truth(rng, n, k, cols, dtype):
    x = np.zeros((n, cols), dtype=dtype)
    for c in range(cols):
        idx = rng.choice(n, k, replace=False)
        v = rng.uniform(1, 2, k) * rng.choice([-1, 1], k)
        if np.issubdtype(np.dtype(dtype), np.complexfloating):
            v = v * np.exp(2j * np.pi * rng.uniform(size=k))
        x[idx, c] = v
    return x


@pytest.mark.parametrize("typeA", [np.float32, np.float64, np.complex64, np.complex128])
@pytest.mark.parametrize("solver", ["ISTA", "FISTA"])
def test_reference_configuration(fm, typeA, solver):
    """ISTA.py:187-236 / FISTA's _getTest: Product(Matrix(U(-100, 100),
    64 x 64), Hadamard(6), typeExpansion=typeA), 6-sparse truth, lambda 1,
    10 steps."""
    import fastmat_amd.algorithms as fma
    rng = np.random.default_rng(41)
    say("reference config", solver, np.dtype(typeA).name)
    M = rng.uniform(-100, 100, (64, 64)).astype(typeA)
    A = fm.Product(fm.Matrix(M), fm.Hadamard(6), typeExpansion=typeA)
    # the dense matrix in the instance's dtype, so the oracle's products
    # return what A.forward returns (the reference's dtype flow)
    Ad = A.reference().astype(A.dtype)
    x0 = sparse_truth(rng, 64, 6, 2, typeA)
    b = (Ad @ x0).astype(np.promote_types(typeA, np.float32))
    alg = getattr(fma, solver)(A, numLambda=1.0, numMaxSteps=10)
    out = alg.process(b)
    ref, _, _ = getattr(oalg, solver.lower())(Ad, b, num_lambda=1.0, num_max_steps=10)
    assert out.shape == ref.shape
    assert out.dtype == ref.dtype, (out.dtype, ref.dtype)
    assert accept.max_rel_error(out, ref) < tol_for(ref.dtype)


@pytest.mark.parametrize("solver", ["ISTA", "FISTA"])
def test_circulant_deconvolution(fm, solver):
    """The docstring example of ISTA.py:44-66: a windowed harmonic as a
    Circulant kernel, 3 pulses, lambda 0.005, 100 steps (float64)."""
    import fastmat_amd.algorithms as fma
    rng = np.random.default_rng(42)
    n = 512
    t = np.linspace(0, 20 * np.pi, n)
    c = np.cos(2 * t) * np.exp(-t ** 2)
    C = fm.Circulant(c)
    x = np.zeros(n)
    x[rng.choice(n, 3, replace=False)] = 1
    b = C.reference() @ x
    say("deconvolution", solver)
    alg = getattr(fma, solver)(C, numLambda=0.005, numMaxSteps=100)
    out = alg.process(b)
    ref, _, _ = getattr(oalg, solver.lower())(C.reference(), b, num_lambda=0.005,
                                                 num_max_steps=100)
    assert out.shape == (n, 1)
    assert accept.max_rel_error(out, ref) < 1e-8


def test_graph_replay_equals_eager_and_device_io(fm):
    import torch
    import fastmat_amd.algorithms as fma
    rng = np.random.default_rng(43)
    n = 4096
    c = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    C = fm.Circulant(c)
    x0 = sparse_truth(rng, n, 20, 3, np.complex64)
    b = np.asarray(C.forward(x0))
    say("graph vs eager FISTA")
    graphed = fma.FISTA(C, numLambda=0.05, numMaxSteps=25).process(b)
    seen = []
    eager_alg = fma.FISTA(C, numLambda=0.05, numMaxSteps=25)
    eager_alg.cbStep = lambda a: seen.append(a.numStep)
    eager = eager_alg.process(b)
    assert seen == list(range(25))
    np.testing.assert_array_equal(graphed, eager)
    # device in -> device out, same values
    bd = torch.from_numpy(np.ascontiguousarray(b.T)).cuda().t()
    dev = fma.FISTA(C, numLambda=0.05, numMaxSteps=25).process(bd)
    assert torch.is_tensor(dev) and dev.is_cuda
    np.testing.assert_array_equal(dev.cpu().numpy(), graphed)
    ref, _, _ = oalg.fista(C.reference(), b, num_lambda=0.05, num_max_steps=25)
    assert accept.max_rel_error(graphed, ref) < 1e-4


def test_largest_singular_value(fm):
    rng = np.random.default_rng(44)
    say("largest singular value")
    M = rng.standard_normal((40, 40))
    P = fm.Product(fm.Matrix(M), fm.Matrix(M.T))
    ref = np.linalg.svd(M @ M.T, compute_uv=False)[0]
    assert abs(P.largestSingularValue - ref) <= 1e-6 * ref
    T = fm.Toeplitz(rng.standard_normal(30), rng.standard_normal(45))
    ref = np.linalg.svd(T.reference(), compute_uv=False)[0]
    assert abs(T.largestSingularValue - ref) <= 1e-6 * ref
    K = fm.Kron(fm.Fourier(8), fm.Circulant(rng.standard_normal(6)))
    ref = np.linalg.svd(K.reference(), compute_uv=False)[0]
    assert abs(K.largestSingularValue - ref) <= 1e-9 * ref
    with pytest.raises(ValueError):
        import fastmat_amd.algorithms as fma
        fma.ISTA(fm.Fourier(8), numMaxSteps=0).process(np.ones(8))
    with pytest.raises(TypeError):
        import fastmat_amd.algorithms as fma
        fma.ISTA(np.eye(3))


# ------------------------------------------------------------------ STELA
@pytest.mark.parametrize("typeA", [np.float32, np.float64, np.complex64, np.complex128])
def test_stela_reference_configuration(fm, typeA):
    """STELA.py:268-330: Product(Matrix(U(-100, 100)) 64 x 64, Hadamard(6),
    typeExpansion=typeA), 6-sparse truth, lambda 0.1, 3 steps.  STELA works
    in float64 / complex128, so the oracle gets the exact dense product."""
    import fastmat_amd.algorithms as fma
    rng = np.random.default_rng(61)
    say("STELA reference config", np.dtype(typeA).name)
    M = rng.uniform(-100, 100, (64, 64)).astype(typeA)
    A = fm.Product(fm.Matrix(M), fm.Hadamard(6), typeExpansion=typeA)
    Ad = A.reference()
    x0 = sparse_truth(rng, 64, 6, 2, typeA)
    b = (Ad @ x0).astype(typeA)
    out = fma.STELA(A, numLambda=0.1, numMaxSteps=3).process(b)
    ref, _, _ = oalg.stela(Ad, b, num_lambda=0.1, num_max_steps=3)
    assert out.shape == ref.shape and out.dtype == ref.dtype, (out.dtype, ref.dtype)
    assert accept.max_rel_error(out, ref) < 1e-9


@pytest.mark.parametrize("typeB", [np.float64, np.complex128])
def test_stela_circulant_deconvolution(fm, typeB):
    """The docstring example of STELA.py:41-66 (lambda 0.005, 100 steps),
    real data and a complex variant."""
    import fastmat_amd.algorithms as fma
    rng = np.random.default_rng(62)
    n = 512
    t = np.linspace(0, 20 * np.pi, n)
    c = np.cos(2 * t) * np.exp(-t ** 2)
    if typeB == np.complex128:
        c = c + 0.5j * np.sin(3 * t) * np.exp(-t ** 2)
    C = fm.Circulant(c)
    x = np.zeros(n, dtype=typeB)
    x[rng.choice(n, 3, replace=False)] = 1
    b = C.reference() @ x
    say("STELA deconvolution", np.dtype(typeB).name)
    out = fma.STELA(C, numLambda=0.005, numMaxSteps=100).process(b)
    ref, _, _ = oalg.stela(C.reference(), b, num_lambda=0.005, num_max_steps=100,
                           col_norms_=C.colNorms)
    assert out.shape == (n, 1) and out.dtype == ref.dtype
    assert accept.max_rel_error(out, ref) < 1e-8


def test_stela_complex_operator_real_data(fm):
    """A complex operator with real measurements: z = A^H r is complex while
    the iterate stays real (the reference keeps the real part on
    assignment)."""
    import fastmat_amd.algorithms as fma
    rng = np.random.default_rng(63)
    F = fm.Fourier(64)
    b = rng.standard_normal((64, 3))
    say("STELA complex operator, real data")
    out = fma.STELA(F, numLambda=0.5, numMaxSteps=6).process(b)
    ref, _, _ = oalg.stela(F.reference(), b, num_lambda=0.5, num_max_steps=6)
    assert out.dtype == np.float64 and ref.dtype == np.float64
    assert accept.max_rel_error(out, ref) < 1e-9


def test_stela_graph_equals_eager_and_early_stop(fm):
    import torch
    import fastmat_amd.algorithms as fma
    rng = np.random.default_rng(64)
    n = 2048
    c = rng.standard_normal(n)
    C = fm.Circulant(c)
    x0 = sparse_truth(rng, n, 10, 3, np.float64)
    b = np.asarray(C.forward(x0))
    say("STELA graph vs eager")
    graphed = fma.STELA(C, numLambda=0.05, numMaxSteps=20).process(b)
    seen = []
    alg = fma.STELA(C, numLambda=0.05, numMaxSteps=20)
    alg.cbStep = lambda a: seen.append(a.numStep)
    eager = alg.process(b)
    np.testing.assert_array_equal(graphed, eager)
    ref, steps, _ = oalg.stela(C.reference(), b, num_lambda=0.05, num_max_steps=20,
                               col_norms_=C.colNorms)
    assert seen == list(range(steps))
    assert accept.max_rel_error(graphed, ref) < 1e-8
    # device in -> device out
    bd = torch.from_numpy(np.ascontiguousarray(b.T)).cuda().t()
    dev = fma.STELA(C, numLambda=0.05, numMaxSteps=20).process(bd)
    assert torch.is_tensor(dev) and dev.is_cuda
    np.testing.assert_array_equal(dev.cpu().numpy(), graphed)
    # an identity problem converges in one step: the eager loop returns at
    # step 1 before its callbacks, like the reference; the graph replays
    # that follow change nothing
    E = fm.Diag(np.ones(16))
    bb = rng.standard_normal((16, 2))
    seen.clear()
    alg = fma.STELA(E, numLambda=0.5, numMaxSteps=10)
    alg.cbStep = lambda a: seen.append(a.numStep)
    out_e = alg.process(bb)
    out_g = fma.STELA(E, numLambda=0.5, numMaxSteps=10).process(bb)
    assert seen == [0]
    np.testing.assert_array_equal(out_e, out_g)
    np.testing.assert_allclose(out_g, oalg.soft_threshold(bb, 0.5), rtol=1e-15, atol=1e-16)


def test_column_norms_and_scipy_operator(fm):
    rng = np.random.default_rng(65)
    say("colNorms / scipyLinearOperator")
    cases = [
        fm.Fourier(24), fm.Hadamard(5), fm.Circulant(rng.standard_normal(12)),
        fm.Circulant(rng.standard_normal((3, 4))),
        fm.Toeplitz(rng.standard_normal(7), rng.standard_normal(10)),
        fm.Toeplitz(rng.standard_normal((5, 7)), split=np.array([3, 2])),
        fm.Diag(rng.standard_normal(9)), fm.Eye(6),
        fm.Permutation(rng.permutation(8)),
        fm.Kron(fm.Fourier(4), fm.Diag(rng.standard_normal(3))),
        fm.Partial(fm.Fourier(16), cols=np.array([1, 4, 9])),
        fm.Partial(fm.Fourier(16), rows=np.array([0, 3, 5, 6])),
        fm.Product(fm.Matrix(rng.standard_normal((10, 12))),
                   fm.Matrix(rng.standard_normal((12, 7)))),
        fm.Sum(fm.Matrix(rng.standard_normal((8, 8))), fm.Circulant(rng.standard_normal(8))),
        fm.Matrix(rng.standard_normal((5, 3)) + 1j * rng.standard_normal((5, 3))),
    ]
    for A in cases:
        ref = np.linalg.norm(A.reference().astype(np.complex128), axis=0)
        np.testing.assert_allclose(np.asarray(A.colNorms, dtype=np.float64), ref,
                                   rtol=1e-6 if A.dtype in (np.float32, np.complex64)
                                   else 1e-12, err_msg=repr(A))
    # scipy LinearOperator over forward / backward (Matrix.pyx:998-1006)
    C = fm.Circulant(rng.standard_normal(16) + 1j * rng.standard_normal(16))
    L = C.scipyLinearOperator
    v = rng.standard_normal(16) + 1j * rng.standard_normal(16)
    V = rng.standard_normal((16, 3))
    Cd = C.reference()
    assert L.shape == (16, 16) and L.dtype == C.dtype
    assert accept.max_rel_error(L.matvec(v), Cd @ v) < 1e-5
    assert accept.max_rel_error(L.rmatvec(v), Cd.conj().T @ v) < 1e-5
    assert accept.max_rel_error(L.matmat(V), Cd @ V) < 1e-5
    assert C.scipyLinearOperator is L


def test_fista_on_blockdiag_captures(fm):
    """A solver over a BlockDiag of Circulant blocks: the eager first step
    applies the blocks as one batch, the captured step block by block; both
    agree with the oracle."""
    import fastmat_amd.algorithms as fma
    from oracle import dense
    rng = np.random.default_rng(66)
    c1, c2 = rng.standard_normal(256), rng.standard_normal(128)
    B = fm.BlockDiag(fm.Circulant(c1), fm.Circulant(c2))
    ref_m = dense.block_diag([dense.circulant(c1), dense.circulant(c2)])
    x0 = sparse_truth(rng, 384, 5, 2, np.float64)
    b = ref_m @ x0
    say("FISTA over BlockDiag")
    out = fma.FISTA(B, numLambda=0.05, numMaxSteps=30).process(b)
    ref, _, _ = oalg.fista(ref_m, b, num_lambda=0.05, num_max_steps=30)
    assert accept.max_rel_error(out, ref) < 1e-8


def test_matrix_properties_rows_normalized_gram(fm):
    """rowNorms, colNormalized / rowNormalized, gram, getCols / getRows,
    nbytes and the deprecated aliases (Matrix.pyx:316-1340)."""
    import warnings
    import fastmat_amd.algorithms as fma
    rng = np.random.default_rng(67)
    say("matrix properties")
    cases = [fm.Fourier(16), fm.Hadamard(4), fm.Circulant(rng.standard_normal(10)),
             fm.Toeplitz(rng.standard_normal(6), rng.standard_normal(9)),
             fm.Diag(rng.standard_normal(7) + 0.5), fm.Permutation(rng.permutation(6)),
             fm.Kron(fm.Fourier(4), fm.Diag(rng.standard_normal(3) + 2)),
             fm.Partial(fm.Fourier(16), rows=np.array([0, 2, 5])),
             fm.Partial(fm.Fourier(16), cols=np.array([1, 3])),
             fm.LFSRCirculant(0x19, 0xD),
             fm.Product(fm.Matrix(rng.standard_normal((5, 8))), fm.Hadamard(3))]
    for A in cases:
        R = A.reference().astype(np.complex128)
        tol = 1e-6 if A.dtype in (np.float32, np.complex64) else 1e-12
        np.testing.assert_allclose(np.asarray(A.rowNorms, dtype=np.float64),
                                   np.linalg.norm(R, axis=1), rtol=tol, err_msg=repr(A))
        Cn = A.colNormalized.reference().astype(np.complex128)
        np.testing.assert_allclose(np.linalg.norm(Cn, axis=0), 1.0, rtol=tol, err_msg=repr(A))
        Rn = A.rowNormalized.reference().astype(np.complex128)
        np.testing.assert_allclose(np.linalg.norm(Rn, axis=1), 1.0, rtol=tol, err_msg=repr(A))
        G = A.gram
        assert accept.max_rel_error(G.forward(np.eye(A.numCols)), R.conj().T @ R) < tol * 10
        idx = np.array([0, A.numCols - 1])
        assert accept.max_rel_error(A.getCols(idx), R[:, idx]) < tol
        ridx = np.array([A.numRows - 1, 0])
        assert accept.max_rel_error(A.getRows(ridx), R[ridx, :]) < tol
        assert A.nbytes > 0
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        F = fm.Fourier(8)
        assert F.numN == 8 and F.numM == 8 and abs(F.largestSV - np.sqrt(8)) < 1e-12
        assert F.normalized is F.colNormalized
    with pytest.raises(ValueError, match="zero-norm column"):
        fm.Diag(np.array([1.0, 0.0])).colNormalized
    alg = fma.ISTA(fm.Fourier(8))
    assert alg.nbytes >= 0


# ------------------------------------------------------------------- OMP
ALLTYPES = [np.int8, np.int16, np.int32, np.int64, np.float32, np.float64, np.complex64,
            np.complex128]


@pytest.mark.parametrize("typeA", ALLTYPES)
def test_omp_reference_configuration(fm, typeA):
    """OMP.pyx:241-290: Product(Matrix(U(-100, 100)) 8 x 8, Hadamard(3),
    typeExpansion=typeA), 3-sparse truth, 3 steps, against the oracle on the
    exact dense product."""
    import fastmat_amd.algorithms as fma
    rng = np.random.default_rng(71 + ALLTYPES.index(typeA))
    say("OMP reference config", np.dtype(typeA).name)
    M = rng.uniform(-100, 100, (8, 8)).astype(typeA)
    A = fm.Product(fm.Matrix(M), fm.Hadamard(3), typeExpansion=typeA)
    Ad = A.reference()
    x0 = np.zeros((8, 2), dtype=np.promote_types(typeA, np.float64))
    for j in range(2):
        x0[rng.choice(8, 3, replace=False), j] = rng.uniform(1, 2, 3)
    b = Ad @ x0
    out = fma.OMP(A, numMaxSteps=3).process(b)
    ref, sup = oalg.omp(Ad, b, 3)
    assert out.shape == ref.shape and out.dtype == ref.dtype, (out.dtype, ref.dtype)
    assert accept.max_rel_error(out, ref) < 1e-9


def test_omp_reference_unit_test(fm):
    """test/algorithm.py:23-90: 40 x 100 Gaussian, 6-sparse: support and
    values recovered; argument errors."""
    import fastmat_amd.algorithms as fma
    rng = np.random.default_rng(97)
    say("OMP recovery")
    A = rng.standard_normal((40, 100))
    x = np.zeros(100)
    supp = rng.choice(100, 6, replace=False)
    x[supp] = rng.standard_normal(6)
    alg = fma.OMP(fm.Matrix(A), numMaxSteps=6)
    y = alg.process(A @ x)
    np.testing.assert_array_equal(np.sort(np.nonzero(y[:, 0])[0]), np.sort(supp))
    np.testing.assert_allclose(y[:, 0], x, rtol=1e-10, atol=1e-13)
    alg.process(rng.standard_normal((40, 7)))
    with pytest.raises(ValueError):
        alg.process(rng.standard_normal((40, 40, 7)))
    with pytest.raises(ValueError):
        alg.process(np.zeros(()))
    alg.updateParameters(numMaxSteps=0)
    with pytest.raises(ValueError):
        alg.process(rng.standard_normal(40))


def test_omp_structured_operators(fm):
    """Complex Circulant deconvolution (the docstring's example shape) and a
    Fourier-Hadamard product with several columns, against the oracle, and a
    device-resident call."""
    import torch
    import fastmat_amd.algorithms as fma
    rng = np.random.default_rng(98)
    n = 512
    t = np.linspace(0, 20 * np.pi, n)
    c = np.cos(2 * t) * np.exp(-t ** 2 / 50) + 0.3j * np.sin(3 * t) * np.exp(-t ** 2 / 40)
    C = fm.Circulant(c)
    x = np.zeros((n, 3), dtype=np.complex128)
    for j in range(3):
        x[rng.choice(n, 3, replace=False), j] = rng.uniform(1, 2, 3)
    b = C.reference() @ x
    say("OMP circulant")
    out = fma.OMP(C, numMaxSteps=3).process(b)
    ref, _ = oalg.omp(C.reference(), b, 3)
    assert accept.max_rel_error(out, ref) < 1e-9
    bd = torch.from_numpy(np.ascontiguousarray(b.T)).cuda().t()
    dev = fma.OMP(C, numMaxSteps=3).process(bd)
    assert torch.is_tensor(dev) and dev.is_cuda
    np.testing.assert_array_equal(dev.cpu().numpy(), out)


def test_omp_chunked_selection(fm):
    """More than 512 columns, so the selection runs per column chunk and then
    merges: a complex Circulant 4096 against the oracle, and an exact tie
    between columns of different chunks (Hadamard, integer correlations),
    where the first index must win as in np.argmax."""
    import fastmat_amd.algorithms as fma
    rng = np.random.default_rng(99)
    n = 4096
    c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    C = fm.Circulant(c)
    Cd = C.reference()
    x = np.zeros((n, 3), dtype=np.complex128)
    for j in range(3):
        x[rng.choice(n, 4, replace=False), j] = rng.uniform(1, 2, 4)
    b = Cd @ x
    say("OMP circulant 4096")
    out = fma.OMP(C, numMaxSteps=4).process(b)
    ref, _ = oalg.omp(Cd, b, 4)
    assert accept.max_rel_error(out, ref) < 1e-9
    H = fm.Hadamard(12)
    Hd = H.reference().astype(np.float64)
    bt = Hd[:, 3000] + Hd[:, 700] + 0.5 * Hd[:, 1500]
    say("OMP hadamard tie")
    alg = fma.OMP(H, numMaxSteps=3)
    got = alg.process(bt)
    ref, sup = oalg.omp(Hd, bt, 3)
    np.testing.assert_array_equal(alg.arrSupport[:, 0], [700, 3000, 1500])
    np.testing.assert_array_equal(np.asarray(sup).reshape(-1)[:3], [700, 3000, 1500])
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)

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

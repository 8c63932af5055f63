"""The allocation-free Circulant / Toeplitz apply (fm_conv_apply_noalloc,
through the C ABI) against the workspace apply (fm_conv_apply) on the same
plans: bit-identical results, on consecutive calls of which the later ones
are smaller, including a transform whose columns run in several chunks of
the 64 MiB scratch and a length whose single column does not fit it (a
clear ValueError)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def crandn(rng, *shape):
    return (rng.standard_normal(shape, dtype=np.float32) +
            1j * rng.standard_normal(shape, dtype=np.float32)).astype(np.complex64)


@pytest.fixture(scope="module")
def env():
    import torch
    import fastmat_amd as fm
    from fastmat_amd import _lib
    _lib.load()
    return torch, fm, _lib


def noalloc(lib_mod, h, adjoint, X, Y):
    from fastmat_amd import _device as D
    lib_mod.call("fm_conv_apply_noalloc", h, int(adjoint), lib_mod.ptr(X), X.shape[1],
                 lib_mod.ptr(Y), Y.shape[1], X.shape[0], D.stream())


def test_conv_apply_noalloc_matches_workspace_apply(env):
    torch, fm, L = env
    rng = np.random.default_rng(20261203)
    cases = [
        # (label, operator, columns): a chunked four-step first (2^22-point
        # embedding, 32 MiB per column: 3 columns run as 2 + 1), then smaller
        ("circulant n=1.5e6 L=2^22", fm.Circulant(crandn(rng, 1500000)), 3),
        ("toeplitz 9000x7000 L=2^14", fm.Toeplitz(crandn(rng, 9000), crandn(rng, 6999)), 5),
        ("toeplitz 9000x7000 L=2^14", None, 2),
        ("circulant n=700 one pass", fm.Circulant(crandn(rng, 700)), 4),
    ]
    prev = None
    for label, op, k in cases:
        op = op if op is not None else prev
        prev = op
        h = op.plan_handle(np.complex64)
        for adjoint in (False, True):
            n = op.numRows if adjoint else op.numCols
            m = op.numCols if adjoint else op.numRows
            say("noalloc", label, "k", k, "adjoint", adjoint)
            X = torch.from_numpy(crandn(rng, k, n)).cuda()
            ref = op._run_plan(h, X, adjoint)
            Y = torch.full((k, m), float("nan"), dtype=torch.complex64, device="cuda")
            noalloc(L, h, adjoint, X, Y)
            torch.cuda.synchronize()
            assert torch.equal(Y, ref), (label, adjoint)


def test_conv_apply_noalloc_rejects_oversize(env):
    torch, fm, L = env
    rng = np.random.default_rng(5)
    # 2^24-point embedding: 128 MiB per complex64 column > 64 MiB scratch
    op = fm.Circulant(crandn(rng, 5000000))
    X = torch.zeros((1, op.numCols), dtype=torch.complex64, device="cuda")
    Y = torch.zeros((1, op.numRows), dtype=torch.complex64, device="cuda")
    say("noalloc oversize")
    with pytest.raises(ValueError, match="scratch"):
        noalloc(L, op.plan_handle(np.complex64), False, X, Y)

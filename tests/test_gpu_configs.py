"""GPU parity tests at the benchmark configurations' full sizes:

* Fourier complex64 at N = 2^20 (radix) and N = 1,000,003 (Bluestein), 64
  columns each: every column of the forward and the adjoint against the FP64
  oracle (numpy's complex128 FFT), 1e-5 max-norm relative error.
* Hadamard order 20 in float32, 64 columns N(0, 1): bit-exact against the C
  restatement of the reference's radix-2 loop (Hadamard.pyx:164-230), hence
  also within 1e-5 of FP64.
* The uneven batch of 256 Circulant / Toeplitz operators through the packed
  autograd Function: the input gradient of every operator against
  2 A^H (A x - y) and the generator gradient of every operator against the
  FP64 closed-form correlation, both from the oracle in FP64.
"""
import sys

import numpy as np
import pytest

from oracle import accept, fast64, fwht_c

pytestmark = pytest.mark.gpu

TOL32 = 1e-5


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    from fastmat_amd import _lib
    _lib.load()
    return torch


def crandn(rng, *shape):
    return (rng.standard_normal(shape, dtype=np.float32) +
            1j * rng.standard_normal(shape, dtype=np.float32)).astype(np.complex64)


@pytest.mark.parametrize("n", [1 << 20, 1000003])
def test_fourier_config_64_columns(torch_mod, n):
    torch = torch_mod
    import fastmat_amd as fm
    rng = np.random.default_rng(20261202 + 1)
    k = 64
    x = crandn(rng, n, k)
    F = fm.Fourier(n)
    X = torch.from_numpy(np.ascontiguousarray(x.T)).cuda()
    say("fourier", n, "k", k)
    Y = F._apply(X, False).cpu().numpy().T
    Z = F._apply(X, True).cpu().numpy().T
    for c0 in range(0, k, 16):
        xs = x[:, c0:c0 + 16]
        assert accept.max_rel_error(Y[:, c0:c0 + 16], fast64.fourier(xs)) < TOL32, c0
        assert accept.max_rel_error(Z[:, c0:c0 + 16], fast64.fourier(xs, adjoint=True)) < TOL32
    # per column, too (the max-norm of one column is not dominated by another's)
    errs = [accept.max_rel_error(Y[:, c], fast64.fourier(x[:, c])) for c in range(0, k, 9)]
    say("  sampled column errors", max(errs))
    assert max(errs) < TOL32


def test_hadamard_order20_float32_config(torch_mod):
    import fastmat_amd as fm
    rng = np.random.default_rng(20261202 + 2)
    x = rng.standard_normal((1 << 20, 64), dtype=np.float32)
    say("hadamard order 20 float32 k=64")
    H = fm.Hadamard(20)
    y = H.forward(x)
    assert y.dtype == np.float32
    ref = fwht_c.hadamard(x)
    np.testing.assert_array_equal(y, ref)
    np.testing.assert_array_equal(H.backward(x), ref)
    assert accept.max_rel_error(y[:, :4], fast64.hadamard(x[:, :4].astype(np.float64))) < TOL32


def test_uneven_batch_config_gradients(torch_mod):
    """Loss sum_i ||A_i x_i - y_i||^2 over the 256-operator batch through the
    packed autograd Function (complex64): every x-gradient and every
    generator gradient against FP64."""
    torch = torch_mod
    import bench
    import fastmat_amd.autograd as fa
    from fastmat_amd.batch import ConvBatch
    spec = bench.make_uneven_batch()
    ops, ks = bench.build_batch_ops(spec)
    batch = ConvBatch(ops, ks)
    rng = np.random.default_rng(20261202 + 5)
    n_in, n_out = batch.packed_sizes()
    gens = [torch.from_numpy(np.asarray(e[4], dtype=np.complex64)).cuda().requires_grad_(True)
            for e in spec]
    xp = torch.from_numpy(crandn(rng, n_in)).cuda().requires_grad_(True)
    yp = torch.from_numpy(crandn(rng, n_out)).cuda()
    say("uneven batch autograd", len(ops), "ops", sum(ks), "cols")
    out = fa.batch_conv_matvec_packed(batch, gens, xp)
    ((out - yp).abs() ** 2).sum().backward()
    torch.cuda.synchronize()
    xv = batch.packed_views(xp.detach(), inputs=True)
    yv = batch.packed_views(yp, inputs=False)
    ov = batch.packed_views(out.detach(), inputs=False)
    gxv = batch.packed_views(xp.grad, inputs=True)
    worst = [0.0, 0.0, 0.0]
    for i, (kind, m, n, k, gen) in enumerate(spec):
        x = xv[i].cpu().numpy().T.astype(np.complex128)
        y = yv[i].cpu().numpy().T.astype(np.complex128)
        g64 = np.asarray(gen, dtype=np.complex64).astype(np.complex128)
        if kind == "circulant":
            ax = fast64.circulant(g64, x)
        else:
            ax = fast64.toeplitz(g64, m, x)
        r = ax - y
        if kind == "circulant":
            gx_ref = 2 * fast64.circulant(g64, r, adjoint=True)
        else:
            gx_ref = 2 * fast64.toeplitz(g64, m, r, adjoint=True)
        gg_ref = fast64.conv_generator_grad(kind, m, n, x, 2 * r)
        e = (accept.max_rel_error(ov[i].cpu().numpy().T, ax),
             accept.max_rel_error(gxv[i].cpu().numpy().T, gx_ref),
             accept.max_rel_error(gens[i].grad.cpu().numpy(), gg_ref))
        worst = [max(a, b) for a, b in zip(worst, e)]
        assert max(e) < TOL32, (i, kind, m, n, k, e)
    say("  max rel err: output %.2e, x-grad %.2e, generator grad %.2e" % tuple(worst))


def test_uneven_batch_config_least_squares(torch_mod):
    """The 256-operator batch through the fused least-squares Function (the
    benchmark's step): the loss and every x and generator gradient against
    FP64."""
    torch = torch_mod
    import bench
    import fastmat_amd.autograd as fa
    from fastmat_amd.batch import ConvBatch
    spec = bench.make_uneven_batch()
    ops, ks = bench.build_batch_ops(spec)
    batch = ConvBatch(ops, ks)
    rng = np.random.default_rng(20261202 + 6)
    n_in, n_out = batch.packed_sizes()
    gens = [torch.from_numpy(np.asarray(e[4], dtype=np.complex64)).cuda().requires_grad_(True)
            for e in spec]
    xp = torch.from_numpy(crandn(rng, n_in)).cuda().requires_grad_(True)
    yp = torch.from_numpy(crandn(rng, n_out)).cuda()
    say("uneven batch least squares", len(ops), "ops")
    loss = fa.batch_least_squares_packed(batch, gens, xp, yp)
    loss.backward()
    torch.cuda.synchronize()
    xv = batch.packed_views(xp.detach(), inputs=True)
    yv = batch.packed_views(yp, inputs=False)
    gxv = batch.packed_views(xp.grad, inputs=True)
    tot = 0.0
    worst = [0.0, 0.0]
    for i, (kind, m, n, k, gen) in enumerate(spec):
        x = xv[i].cpu().numpy().T.astype(np.complex128)
        y = yv[i].cpu().numpy().T.astype(np.complex128)
        g64 = np.asarray(gen, dtype=np.complex64).astype(np.complex128)
        if kind == "circulant":
            r = fast64.circulant(g64, x) - y
            gx_ref = 2 * fast64.circulant(g64, r, adjoint=True)
        else:
            r = fast64.toeplitz(g64, m, x) - y
            gx_ref = 2 * fast64.toeplitz(g64, m, r, adjoint=True)
        tot += float(np.sum(np.abs(r) ** 2))
        e = (accept.max_rel_error(gxv[i].cpu().numpy().T, gx_ref),
             accept.max_rel_error(gens[i].grad.cpu().numpy(),
                                  fast64.conv_generator_grad(kind, m, n, x, 2 * r)))
        worst = [max(a, b) for a, b in zip(worst, e)]
        assert max(e) < TOL32, (i, kind, m, n, k, e)
    say("  loss rel err %.2e, x-grad %.2e, generator grad %.2e"
        % (abs(float(loss) - tot) / tot, worst[0], worst[1]))
    assert abs(float(loss) - tot) < TOL32 * tot

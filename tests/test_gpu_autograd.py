"""GPU tests of the differentiable structured matvecs (pattern (a)): the
forward of ``conv_matvec`` / ``batch_conv_matvec_packed`` is the operator of
the generator tensor's *current* values, recomputed on the device every call
(the operator is its generator's spectrum, Circulant.pyx:130,
Toeplitz.pyx:253-255), and the generator gradient is checked against FP64
central differences of the loss computed by the *device* forward in
complex128, and against the FP64 closed-form correlation of
oracle.fast64.conv_generator_grad.
"""
import sys

import numpy as np
import pytest

from oracle import accept, fast64

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def crandn(rng, *shape, dtype=np.complex64):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    from fastmat_amd import _lib
    _lib.load()
    return torch


def oracle_apply(kind, m, gen, x, adjoint=False):
    if kind == "circulant":
        return fast64.circulant(gen, x, adjoint=adjoint)
    return fast64.toeplitz(gen, m, x, adjoint=adjoint)


def make_op(kind, m, gen):
    import fastmat_amd as fm
    return fm.Circulant(gen) if kind == "circulant" else fm.Toeplitz(gen[:m], gen[m:])


CASES = [("circulant", 300, 300), ("circulant", 5000, 5000), ("toeplitz", 200, 350),
         ("toeplitz", 3000, 5000)]


@pytest.mark.parametrize("kind,m,n", CASES)
def test_forward_follows_generator_updates(torch_mod, kind, m, n):
    """Updating the generator tensor in place (as an optimizer step does)
    changes the next forward to the new operator; the matrix object keeps
    its constructor's operator."""
    torch = torch_mod
    import fastmat_amd.autograd as fa
    rng = np.random.default_rng(m * 7 + n)
    d = n if kind == "circulant" else m + n - 1
    g0 = crandn(rng, d)
    delta = crandn(rng, d)
    x = crandn(rng, n, 3)
    A = make_op(kind, m, g0)
    gen = torch.from_numpy(g0.copy()).cuda().requires_grad_(True)
    xt = torch.from_numpy(x).cuda()
    say("generator update", kind, m, n)
    y1 = fa.conv_matvec(A, gen, xt).detach().cpu().numpy()
    assert accept.max_rel_error(y1, oracle_apply(kind, m, g0, x)) < 1e-5
    with torch.no_grad():
        gen.add_(torch.from_numpy(delta).cuda())
    y2 = fa.conv_matvec(A, gen, xt).detach().cpu().numpy()
    g1 = (g0.astype(np.complex128) + delta.astype(np.complex128))
    assert accept.max_rel_error(y2, oracle_apply(kind, m, g1, x)) < 1e-5
    # the matrix object is unchanged
    assert accept.max_rel_error(A.forward(x), oracle_apply(kind, m, g0, x)) < 1e-5
    # the gradient after the update belongs to the updated operator
    y = crandn(rng, m, 3)
    out = fa.conv_matvec(A, gen, xt)
    ((out - torch.from_numpy(y).cuda()).abs() ** 2).sum().backward()
    r = oracle_apply(kind, m, g1, x.astype(np.complex128)) - y
    ref = fast64.conv_generator_grad(kind, m, n, x, 2 * r)
    assert accept.max_rel_error(gen.grad.cpu().numpy(), ref) < 1e-5


@pytest.mark.parametrize("kind,m,n", CASES)
def test_generator_grad_vs_device_finite_differences(torch_mod, kind, m, n):
    """complex128: the autograd generator gradient against FP64 central
    differences (step 1e-6) of the loss sum |A(gen) x - y|^2 evaluated by two
    device forwards per coordinate (the loss is quadratic in gen, so central
    differences are exact up to rounding)."""
    torch = torch_mod
    import fastmat_amd.autograd as fa
    rng = np.random.default_rng(m + 3 * n)
    d = n if kind == "circulant" else m + n - 1
    g0 = crandn(rng, d, dtype=np.complex128)
    x = crandn(rng, n, 2, dtype=np.complex128)
    y = crandn(rng, m, 2, dtype=np.complex128)
    A = make_op(kind, m, g0)
    gen = torch.from_numpy(g0.copy()).cuda().requires_grad_(True)
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    yt = torch.from_numpy(y).cuda()
    say("device FD", kind, m, n)
    loss = ((fa.conv_matvec(A, gen, xt) - yt).abs() ** 2).sum()
    loss.backward()
    gg = gen.grad.cpu().numpy()

    def dev_loss(g):
        with torch.no_grad():
            return float(((fa.conv_matvec(A, g, xt) - yt).abs() ** 2).sum())

    h = 1e-6
    idx = sorted(set([0, 1, d // 2, d - 1] + list(rng.integers(0, d, 4))))
    fd = []
    base = gen.detach()
    for q in idx:
        vals = []
        for unit in (1.0, 1j):
            e = torch.zeros_like(base)
            e[q] = h * unit
            vals.append((dev_loss(base + e) - dev_loss(base - e)) / (2 * h))
        fd.append(vals[0] + 1j * vals[1])
    fd = np.array(fd)
    err = np.max(np.abs(gg[idx] - fd)) / np.max(np.abs(fd))
    say("  generator grad vs device FD", err)
    assert err < 1e-6
    # and the whole gradient against the FP64 closed form
    r = oracle_apply(kind, m, g0, x) - y
    ref = fast64.conv_generator_grad(kind, m, n, x, 2 * r)
    assert accept.max_rel_error(gg, ref) < 1e-12
    assert accept.max_rel_error(xt.grad.cpu().numpy(),
                                2 * oracle_apply(kind, m, g0, r, adjoint=True)) < 1e-12


def test_backward_after_an_intervening_forward(torch_mod):
    """Two forwards of one matrix with different generators, then one
    backward: each output's gradient uses its own generator."""
    torch = torch_mod
    import fastmat_amd.autograd as fa
    rng = np.random.default_rng(31)
    n = 700
    ga, gb = crandn(rng, n), crandn(rng, n)
    x = crandn(rng, n, 2)
    A = make_op("circulant", n, ga)
    ta = torch.from_numpy(ga).cuda().requires_grad_(True)
    tb = torch.from_numpy(gb).cuda().requires_grad_(True)
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    say("two forwards, one backward")
    ya = fa.conv_matvec(A, ta, xt)
    yb = fa.conv_matvec(A, tb, xt)
    (ya.abs().pow(2).sum() + 2 * yb.abs().pow(2).sum()).backward()
    x64 = x.astype(np.complex128)
    ra = oracle_apply("circulant", n, ga, x64)
    rb = oracle_apply("circulant", n, gb, x64)
    ref_x = 2 * oracle_apply("circulant", n, ga, ra, True) + \
        4 * oracle_apply("circulant", n, gb, rb, True)
    assert accept.max_rel_error(xt.grad.cpu().numpy(), ref_x) < 1e-5
    assert accept.max_rel_error(ta.grad.cpu().numpy(),
                                fast64.conv_generator_grad("circulant", n, n, x, 2 * ra)) < 1e-5
    assert accept.max_rel_error(tb.grad.cpu().numpy(),
                                fast64.conv_generator_grad("circulant", n, n, x, 4 * rb)) < 1e-5


def test_packed_batch_follows_generators_and_fd_complex128(torch_mod):
    """The packed batch Function in complex128 on a mixed batch (one- and
    multi-pass Circulant / Toeplitz): outputs of the generators' current
    values after an in-place update, every generator gradient against the
    FP64 closed form, and sampled coordinates against central differences of
    the device loss."""
    torch = torch_mod
    import fastmat_amd.autograd as fa
    from fastmat_amd.batch import ConvBatch
    rng = np.random.default_rng(41)
    spec = [("circulant", 300, 300, 2), ("toeplitz", 200, 350, 1),
            ("circulant", 5000, 5000, 2), ("toeplitz", 3000, 5000, 3),
            ("circulant", 64, 64, 1)]
    gens0 = [crandn(rng, n if k_ == "circulant" else m + n - 1, dtype=np.complex128)
             for k_, m, n, _ in spec]
    ops = [make_op(k_, m, g) for (k_, m, n, _), g in zip(spec, gens0)]
    ks = [e[3] for e in spec]
    batch = ConvBatch(ops, ks, dtype=np.complex128)
    n_in, n_out = batch.packed_sizes()
    gens = [torch.from_numpy(g.copy()).cuda().requires_grad_(True) for g in gens0]
    xp = torch.from_numpy(crandn(rng, n_in, dtype=np.complex128)).cuda().requires_grad_(True)
    yp = torch.from_numpy(crandn(rng, n_out, dtype=np.complex128)).cuda()
    deltas = [crandn(rng, g.size, dtype=np.complex128) for g in gens0]
    with torch.no_grad():
        for g, dl in zip(gens, deltas):
            g.add_(torch.from_numpy(dl).cuda())
    cur = [g0 + dl for g0, dl in zip(gens0, deltas)]
    say("packed batch complex128, updated generators")
    out = fa.batch_conv_matvec_packed(batch, gens, xp)
    ((out - yp).abs() ** 2).sum().backward()
    xv = batch.packed_views(xp.detach(), inputs=True)
    yv = batch.packed_views(yp, inputs=False)
    ov = batch.packed_views(out.detach(), inputs=False)
    for i, ((k_, m, n, k), g) in enumerate(zip(spec, cur)):
        x = xv[i].cpu().numpy().T
        y = yv[i].cpu().numpy().T
        ref = oracle_apply(k_, m, g, x)
        assert accept.max_rel_error(ov[i].cpu().numpy().T, ref) < 1e-12, i
        gref = fast64.conv_generator_grad(k_, m, n, x, 2 * (ref - y))
        assert accept.max_rel_error(gens[i].grad.cpu().numpy(), gref) < 1e-12, i
    # central differences (step 1e-6) of operator i's own loss term
    # |A_i x_i - y_i|^2, each evaluated by a device forward of the batch; the
    # loss is quadratic in the generator, so the differences carry rounding
    # only, about eps * loss / h, which sets the bound
    base = [g.detach().clone() for g in gens]

    def dev_loss(gl, i):
        with torch.no_grad():
            o = fa.batch_conv_matvec_packed(batch, gl, xp.detach())
            oi = batch.packed_views(o, inputs=False)[i]
            return float(((oi - yv[i]).abs() ** 2).sum())

    h = 1e-6
    for i in (0, 3):
        q = int(rng.integers(0, base[i].numel()))
        fd = 0j
        for unit in (1.0, 1j):
            e = torch.zeros_like(base[i])
            e[q] = h * unit
            gp = list(base)
            gm = list(base)
            gp[i] = base[i] + e
            gm[i] = base[i] - e
            fd += unit * (dev_loss(gp, i) - dev_loss(gm, i)) / (2 * h)
        got = complex(gens[i].grad[q])
        bound = 1e-7 * abs(got) + 50 * 2.2e-16 * dev_loss(base, i) / h
        say("  batch FD op", i, "coord", q, "abs err", abs(got - fd), "bound", bound)
        assert abs(got - fd) < bound


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_least_squares_function_matches_composition(torch_mod, dtype):
    """batch_least_squares_packed (residual and loss formed inside the
    forward apply) against the generic composition
    ((batch_conv_matvec_packed(...) - y).abs() ** 2).sum() and the FP64
    oracle: loss, x gradient and every generator gradient, on a mixed batch of
    one- and multi-pass operators; twice (bindings reused), the second time
    with an upstream gradient of 2.5 (the adjoint writes the x gradient,
    scaled by it, straight into a new tensor)."""
    torch = torch_mod
    import fastmat_amd.autograd as fa
    from fastmat_amd.batch import ConvBatch
    rng = np.random.default_rng(51)
    spec = [("circulant", 300, 300, 2), ("toeplitz", 200, 350, 5),
            ("circulant", 5000, 5000, 6), ("toeplitz", 3000, 9000, 3),
            ("circulant", 64, 64, 1), ("toeplitz", 9000, 7000, 2)]
    gens0 = [crandn(rng, n if k_ == "circulant" else m + n - 1, dtype=dtype)
             for k_, m, n, _ in spec]
    ops = [make_op(k_, m, g) for (k_, m, n, _), g in zip(spec, gens0)]
    batch = ConvBatch(ops, [e[3] for e in spec], dtype=dtype)
    n_in, n_out = batch.packed_sizes()
    tol = 1e-5 if dtype == np.complex64 else 1e-11
    for rep in range(2):
        gens = [torch.from_numpy(g.copy()).cuda().requires_grad_(True) for g in gens0]
        xp = torch.from_numpy(crandn(rng, n_in, dtype=dtype)).cuda().requires_grad_(True)
        yp = torch.from_numpy(crandn(rng, n_out, dtype=dtype)).cuda()
        w = 1.0 if rep == 0 else 2.5
        say("least squares", np.dtype(dtype).name, "rep", rep, "upstream", w)
        loss = fa.batch_least_squares_packed(batch, gens, xp, yp)
        (loss * w).backward()
        gens2 = [g.detach().clone().requires_grad_(True) for g in gens]
        xp2 = xp.detach().clone().requires_grad_(True)
        out = fa.batch_conv_matvec_packed(batch, gens2, xp2)
        loss2 = ((out - yp).abs() ** 2).sum()
        (loss2 * w).backward()
        assert abs(float(loss.detach()) - float(loss2.detach())) <= tol * abs(float(loss2.detach()))
        assert accept.max_rel_error(xp.grad.cpu().numpy(), xp2.grad.cpu().numpy()) < tol
        for g, g2 in zip(gens, gens2):
            assert accept.max_rel_error(g.grad.cpu().numpy(), g2.grad.cpu().numpy()) < tol
        # FP64 oracle
        xv = batch.packed_views(xp.detach(), inputs=True)
        yv = batch.packed_views(yp, inputs=False)
        gxv = batch.packed_views(xp.grad, inputs=True)
        tot = 0.0
        for i, ((k_, m, n, k), g0) in enumerate(zip(spec, gens0)):
            x = xv[i].cpu().numpy().T.astype(np.complex128)
            y = yv[i].cpu().numpy().T.astype(np.complex128)
            r = oracle_apply(k_, m, g0.astype(np.complex128), x) - y
            tot += float(np.sum(np.abs(r) ** 2))
            gx = 2 * w * oracle_apply(k_, m, g0.astype(np.complex128), r, adjoint=True)
            assert accept.max_rel_error(gxv[i].cpu().numpy().T, gx) < 10 * tol, i
            assert accept.max_rel_error(gens[i].grad.cpu().numpy(),
                                        fast64.conv_generator_grad(k_, m, n, x, 2 * w * r)) \
                < 10 * tol
        assert abs(float(loss.detach()) - tot) <= 10 * tol * tot


@pytest.mark.parametrize("want", ["x", "gen"])
def test_least_squares_one_kind_of_gradient(torch_mod, want):
    """Only the x gradient (generators without requires_grad: the adjoint
    alone, redirected) or only the generator gradients (x without
    requires_grad: the correlations alone) against the composition through
    the generic matvec Function."""
    torch = torch_mod
    import fastmat_amd.autograd as fa
    from fastmat_amd.batch import ConvBatch
    rng = np.random.default_rng(57)
    spec = [("circulant", 300, 300, 2), ("toeplitz", 3000, 9000, 3),
            ("circulant", 5000, 5000, 4), ("toeplitz", 200, 350, 1)]
    gens0 = [crandn(rng, n if k_ == "circulant" else m + n - 1) for k_, m, n, _ in spec]
    ops = [make_op(k_, m, g) for (k_, m, n, _), g in zip(spec, gens0)]
    batch = ConvBatch(ops, [e[3] for e in spec])
    n_in, n_out = batch.packed_sizes()
    gens = [torch.from_numpy(g).cuda().requires_grad_(want == "gen") for g in gens0]
    xp = torch.from_numpy(crandn(rng, n_in)).cuda().requires_grad_(want == "x")
    yp = torch.from_numpy(crandn(rng, n_out)).cuda()
    say("least squares, only", want)
    (fa.batch_least_squares_packed(batch, gens, xp, yp) * 1.5).backward()
    gens2 = [g.detach().clone().requires_grad_(want == "gen") for g in gens]
    xp2 = xp.detach().clone().requires_grad_(want == "x")
    out = fa.batch_conv_matvec_packed(ConvBatch(ops, [e[3] for e in spec]), gens2, xp2)
    (((out - yp).abs() ** 2).sum() * 1.5).backward()
    if want == "x":
        assert all(g.grad is None for g in gens)
        assert accept.max_rel_error(xp.grad.cpu().numpy(), xp2.grad.cpu().numpy()) < 1e-5
    else:
        assert xp.grad is None
        for g, g2 in zip(gens, gens2):
            assert accept.max_rel_error(g.grad.cpu().numpy(), g2.grad.cpu().numpy()) < 1e-5


def test_least_squares_backward_after_another_forward(torch_mod):
    """Two least-squares forwards on one batch (different x and y), then
    the first one's backward, then the second's: each backward must give its
    own step's gradients (the forward's last pass holds the adjoint's column
    transforms, so a stale one is run again), against the FP64 oracle."""
    torch = torch_mod
    import fastmat_amd.autograd as fa
    from fastmat_amd.batch import ConvBatch
    rng = np.random.default_rng(53)
    spec = [("circulant", 300, 300, 2), ("toeplitz", 3000, 9000, 3),
            ("circulant", 5000, 5000, 4)]
    gens0 = [crandn(rng, n if k_ == "circulant" else m + n - 1) for k_, m, n, _ in spec]
    ops = [make_op(k_, m, g) for (k_, m, n, _), g in zip(spec, gens0)]
    batch = ConvBatch(ops, [e[3] for e in spec])
    n_in, n_out = batch.packed_sizes()
    steps = []
    for _ in range(2):
        gens = [torch.from_numpy(g.copy()).cuda().requires_grad_(True) for g in gens0]
        xp = torch.from_numpy(crandn(rng, n_in)).cuda().requires_grad_(True)
        yp = torch.from_numpy(crandn(rng, n_out)).cuda()
        steps.append((gens, xp, yp, fa.batch_least_squares_packed(batch, gens, xp, yp)))
    say("least squares backward after another forward")
    for gens, xp, yp, loss in steps:
        loss.backward()
        xv = batch.packed_views(xp.detach(), inputs=True)
        yv = batch.packed_views(yp, inputs=False)
        gxv = batch.packed_views(xp.grad, inputs=True)
        for i, ((k_, m, n, k), g0) in enumerate(zip(spec, gens0)):
            x = xv[i].cpu().numpy().T.astype(np.complex128)
            y = yv[i].cpu().numpy().T.astype(np.complex128)
            r = oracle_apply(k_, m, g0.astype(np.complex128), x) - y
            gx = 2 * oracle_apply(k_, m, g0.astype(np.complex128), r, adjoint=True)
            assert accept.max_rel_error(gxv[i].cpu().numpy().T, gx) < 1e-5, i
            assert accept.max_rel_error(gens[i].grad.cpu().numpy(),
                                        fast64.conv_generator_grad(k_, m, n, x, 2 * r)) < 1e-5


def test_least_squares_step_replays_in_a_graph(torch_mod):
    """The fused least-squares step (generators to spectra, forward with
    residual, loss, adjoint, generator gradient) captured in one HIP graph:
    every replay reproduces the eager step, with eager work between."""
    torch = torch_mod
    import fastmat_amd.autograd as fa
    from fastmat_amd.batch import ConvBatch
    rng = np.random.default_rng(52)
    spec = [("circulant", 300, 300, 2), ("toeplitz", 3000, 9000, 3),
            ("circulant", 5000, 5000, 4)]
    gens0 = [crandn(rng, n if k_ == "circulant" else m + n - 1) for k_, m, n, _ in spec]
    ops = [make_op(k_, m, g) for (k_, m, n, _), g in zip(spec, gens0)]
    batch = ConvBatch(ops, [e[3] for e in spec])
    n_in, n_out = batch.packed_sizes()
    gens = [torch.from_numpy(g).cuda().requires_grad_(True) for g in gens0]
    xp = torch.from_numpy(crandn(rng, n_in)).cuda().requires_grad_(True)
    yp = torch.from_numpy(crandn(rng, n_out)).cuda()

    def eager():
        for t in gens + [xp]:
            t.grad = None
        loss = fa.batch_least_squares_packed(batch, gens, xp, yp)
        loss.backward()
        return loss

    say("least squares graph capture")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            ref_loss = float(eager())
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    ref_x = xp.grad.clone()
    ref_g = [g.grad.clone() for g in gens]
    graph = torch.cuda.CUDAGraph()
    for t in gens + [xp]:
        t.grad = None
    with torch.cuda.graph(graph):
        loss = fa.batch_least_squares_packed(batch, gens, xp, yp)
        loss.backward()
    gx, gg = xp.grad, [g.grad for g in gens]
    for rep in range(3):
        gx.zero_()
        graph.replay()
        torch.cuda.synchronize()
        assert float(loss) == ref_loss, rep
        assert torch.equal(gx, ref_x), rep
        for a, b in zip(gg, ref_g):
            assert torch.equal(a, b), rep
        _ = (yp * 2).sum().item()  # eager work between replays

"""GPU tests of the uneven batch (persistent launches + HIP graph + autotuner)
and of the autograd Functions (adjoint backward, generator gradients).

Oracle: FP64 evaluations (oracle.fast64) per operator; gradients against
FP64 central finite differences (step 1e-6) of the FP64 loss and against the
closed forms  dL/dx = 2 A^H (A x - y)  (the reference's backward()).
complex64 results are held to 1e-5 max-norm relative error.
"""
import sys

import numpy as np
import pytest

from oracle import accept, fast64

pytestmark = pytest.mark.gpu

TOL32 = 1e-5


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


def small_spec(seed, count=24, lo=16, hi=12000, kmax=5):
    """Uneven mix including multi-pass lengths (L > 8192)."""
    rng = np.random.default_rng(seed)
    spec = []
    for i in range(count):
        k = int(rng.integers(1, kmax + 1))
        if i % 2 == 0:
            n = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
            spec.append(("circulant", n, n, k, crandn(rng, n)))
        else:
            m = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
            n = int(min(max(round(m / rng.uniform(0.5, 2.0)), lo), hi))
            spec.append(("toeplitz", m, n, k, crandn(rng, m + n - 1)))
    # make sure both a 3-pass Circulant and a 3-pass Toeplitz are present
    spec.append(("circulant", 9000, 9000, 2, crandn(rng, 9000)))
    spec.append(("toeplitz", 7000, 9000, 3, crandn(rng, 15999)))
    return spec


def build(spec):
    import fastmat_amd as fm
    ops, ks = [], []
    for kind, m, n, k, gen in spec:
        ops.append(fm.Circulant(gen) if kind == "circulant" else fm.Toeplitz(gen[:m], gen[m:]))
        ks.append(k)
    return ops, ks


def oracle_apply(entry, x, adjoint):
    kind, m, n, k, gen = entry
    if kind == "circulant":
        return fast64.circulant(gen, x, adjoint=adjoint)
    return fast64.toeplitz(gen, m, x, adjoint=adjoint)


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a.T)).cuda()


def test_batch_matches_oracle(torch_mod):
    torch = torch_mod
    from fastmat_amd.batch import ConvBatch
    spec = small_spec(1)
    ops, ks = build(spec)
    assert any(op.fftLength > 8192 for op in ops)
    batch = ConvBatch(ops, ks)
    rng = np.random.default_rng(2)
    xs_h = [crandn(rng, op.numCols, k) for op, k in zip(ops, ks)]
    gs_h = [crandn(rng, op.numRows, k) for op, k in zip(ops, ks)]
    xs = [dev(torch, x) for x in xs_h]
    gs = [dev(torch, g) for g in gs_h]
    say("batch forward/adjoint", len(ops))
    ys = batch.forward(xs)
    zs = batch.backward(gs)
    torch.cuda.synchronize()
    for e, x, g, y, z in zip(spec, xs_h, gs_h, ys, zs):
        assert accept.max_rel_error(y.cpu().numpy().T, oracle_apply(e, x, False)) < TOL32, e[:4]
        assert accept.max_rel_error(z.cpu().numpy().T, oracle_apply(e, g, True)) < TOL32, e[:4]
    # same results as the per-operator path
    for op, x, y in zip(ops, xs, ys):
        np.testing.assert_allclose(op._apply(x, False).cpu().numpy(), y.cpu().numpy(),
                                   rtol=0, atol=1e-5 * float(y.abs().max()))


@pytest.mark.parametrize("adjoint", [False, True])
def test_redirected_apply_writes_scaled_outputs_elsewhere(torch_mod, adjoint):
    """fm_batch_apply_redirected: the bound direction's outputs land at the
    same places of another buffer (a byte shift), multiplied by a device
    scalar, and the bound output buffer is not written; the plain apply
    afterwards writes the bound buffer again."""
    torch = torch_mod
    from fastmat_amd.batch import ConvBatch
    spec = small_spec(3)
    ops, ks = build(spec)
    assert any(op.fftLength > 8192 for op in ops)
    batch = ConvBatch(ops, ks)
    n_in, n_out = batch.packed_sizes()
    if adjoint:
        n_in, n_out = n_out, n_in
    rng = np.random.default_rng(4)
    xin = torch.from_numpy(crandn(rng, n_in)).cuda()
    out = torch.empty(n_out, dtype=torch.complex64, device="cuda")
    xs = batch.packed_views(xin, adjoint=adjoint, inputs=True)
    ys = batch.packed_views(out, adjoint=adjoint, inputs=False)
    batch.bind(xs, ys, adjoint)
    say("redirected apply", "adjoint" if adjoint else "forward", len(ops))
    batch.launch(adjoint)
    ref = out.clone()
    out.fill_(7.0)
    other = torch.full((n_out,), 3.0, dtype=torch.complex64, device="cuda")
    s = torch.tensor([2.5], dtype=torch.float32, device="cuda")
    batch.launch(adjoint, out_shift=other.data_ptr() - out.data_ptr(), out_scale=s)
    torch.cuda.synchronize()
    assert bool((out == 7.0).all())
    assert accept.max_rel_error(other.cpu().numpy(), 2.5 * ref.cpu().numpy()) < 1e-6
    batch.launch(adjoint)
    torch.cuda.synchronize()
    assert bool((out == ref).all())
    # against the oracle, per operator
    for e, x, y in zip(spec, xs, batch.packed_views(other, adjoint=adjoint, inputs=False)):
        xr = x.cpu().numpy().T.astype(np.complex128)
        assert accept.max_rel_error(y.cpu().numpy().T, 2.5 * oracle_apply(e, xr, adjoint)) \
            < TOL32, e[:4]


def test_batch_graph_replays_and_tuner(torch_mod, tmp_path):
    torch = torch_mod
    from fastmat_amd.batch import Autotuner, ConvBatch
    spec = small_spec(3, count=10)
    ops, ks = build(spec)
    batch = ConvBatch(ops, ks)
    rng = np.random.default_rng(4)
    xs = [dev(torch, crandn(rng, op.numCols, k)) for op, k in zip(ops, ks)]
    ys = batch.new_outputs(False)
    say("batch autotune")
    tuner = Autotuner(str(tmp_path / "tune.json"))
    best = batch.tune(xs, ys, False, tuner=tuner)
    assert best in (1, 2, 3, 4)
    assert (tmp_path / "tune.json").exists()
    batch.launch(False)
    torch.cuda.synchronize()
    eager = [y.clone() for y in ys]
    say("batch graph capture")
    g = batch.graph(xs, ys, False)
    for rep in range(3):
        for y in ys:
            y.zero_()
        # unrelated eager work between replays
        junk = torch.randn(1 << 16, device="cuda").sum()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(eager, ys):
            assert torch.equal(a, b), "replay %d differs" % rep
        del junk


def test_full_uneven_batch_config(torch_mod):
    """The 256-operator uneven batch at full size: sampled operators against
    the oracle, and adjointness <A x, g> = <x, A^H g> for all operators."""
    torch = torch_mod
    import bench
    from fastmat_amd.batch import ConvBatch
    spec = bench.make_uneven_batch()
    ops, ks = bench.build_batch_ops(spec)
    batch = ConvBatch(ops, ks)
    rng = np.random.default_rng(5)
    xs = [torch.complex(torch.randn(k, op.numCols, device="cuda"),
                        torch.randn(k, op.numCols, device="cuda")) for op, k in zip(ops, ks)]
    gs = [torch.complex(torch.randn(k, op.numRows, device="cuda"),
                        torch.randn(k, op.numRows, device="cuda")) for op, k in zip(ops, ks)]
    say("full batch", len(ops), "ops", sum(ks), "cols")
    ys = batch.forward(xs)
    zs = batch.backward(gs)
    torch.cuda.synchronize()
    for i in range(0, len(ops), 29):
        x = xs[i].cpu().numpy().T
        assert accept.max_rel_error(ys[i].cpu().numpy().T, oracle_apply(spec[i], x, False)) < TOL32
    for i in range(len(ops)):
        lhs = torch.vdot(ys[i].flatten().to(torch.complex128), gs[i].flatten().to(torch.complex128))
        rhs = torch.vdot(xs[i].flatten().to(torch.complex128), zs[i].flatten().to(torch.complex128))
        scale = float(ys[i].abs().max() * gs[i].abs().sum() + 1.0)
        assert abs(complex(lhs) - complex(rhs)) < 1e-5 * scale, i


def _loss64(entry, gen, x, y):
    kind, m, n, k, _ = entry
    e2 = (kind, m, n, k, gen)
    r = oracle_apply(e2, x, False) - y
    return float(np.sum(np.abs(r) ** 2))


def test_autograd_conv_generator_grad_vs_finite_differences(torch_mod):
    torch = torch_mod
    import fastmat_amd.autograd as fa
    rng = np.random.default_rng(6)
    for entry in [("circulant", 300, 300, 3, crandn(rng, 300)),
                  ("circulant", 1024, 1024, 2, crandn(rng, 1024)),
                  ("toeplitz", 200, 350, 2, crandn(rng, 549)),
                  ("toeplitz", 9000, 6000, 1, crandn(rng, 14999))]:
        kind, m, n, k, gen = entry
        say("autograd", kind, m, n, k)
        x = crandn(rng, n, k)
        y = crandn(rng, m, k)
        gen_t = torch.from_numpy(gen).cuda().requires_grad_(True)
        x_t = torch.from_numpy(x).cuda().requires_grad_(True)
        y_t = torch.from_numpy(y).cuda()
        A = fa.circulant(gen_t) if kind == "circulant" else fa.toeplitz(gen_t, m)
        out = fa.conv_matvec(A, gen_t, x_t)
        loss = ((out - y_t).abs() ** 2).sum()
        loss.backward()
        gx = x_t.grad.cpu().numpy()
        gg = gen_t.grad.cpu().numpy()
        # dL/dx = 2 A^H (A x - y)  (the reference's backward applied to the residual)
        g64 = gen.astype(np.complex128)
        e64 = (kind, m, n, k, g64)
        resid = oracle_apply(e64, x.astype(np.complex128), False) - y
        gx_ref = 2 * oracle_apply(e64, resid, True)
        assert accept.max_rel_error(gx, gx_ref) < 1e-4
        # generator gradient vs FP64 central differences on sampled coordinates
        x64, y64 = x.astype(np.complex128), y.astype(np.complex128)
        idx = sorted(set([0, 1, len(gen) // 2, len(gen) - 1] +
                         list(rng.integers(0, len(gen), 4))))
        h = 1e-6
        fd = []
        for d in idx:
            vals = []
            for unit in (1.0, 1j):
                gp = g64.copy(); gp[d] += h * unit
                gm = g64.copy(); gm[d] -= h * unit
                vals.append((_loss64(entry, gp, x64, y64) - _loss64(entry, gm, x64, y64)) / (2 * h))
            fd.append(vals[0] + 1j * vals[1])
        fd = np.array(fd)
        err = np.max(np.abs(gg[idx] - fd)) / np.max(np.abs(fd))
        say("  generator grad rel err vs FD", err)
        assert err < 1e-4


def test_autograd_batch_and_matvec(torch_mod):
    torch = torch_mod
    import fastmat_amd as fm
    import fastmat_amd.autograd as fa
    from fastmat_amd.batch import ConvBatch
    rng = np.random.default_rng(7)
    spec = small_spec(8, count=6, hi=3000, kmax=3)
    ops, ks = build(spec)
    batch = ConvBatch(ops, ks)
    gens = [torch.from_numpy(e[4]).cuda().requires_grad_(True) for e in spec]
    xs = [torch.from_numpy(crandn(rng, op.numCols, k)).cuda().requires_grad_(True)
          for op, k in zip(ops, ks)]
    ys = [torch.from_numpy(crandn(rng, op.numRows, k)).cuda() for op, k in zip(ops, ks)]
    say("autograd batch")
    outs = fa.batch_conv_matvec(batch, gens, xs)
    loss = sum(((o - y).abs() ** 2).sum() for o, y in zip(outs, ys))
    loss.backward()
    for op, g, x, y in zip(ops, gens, xs, ys):
        # compare with the single-operator Function
        g2 = g.detach().clone().requires_grad_(True)
        x2 = x.detach().clone().requires_grad_(True)
        o2 = fa.conv_matvec(op, g2, x2)
        ((o2 - y).abs() ** 2).sum().backward()
        np.testing.assert_allclose(x.grad.cpu().numpy(), x2.grad.cpu().numpy(), rtol=0,
                                   atol=1e-5 * float(x2.grad.abs().max()))
        np.testing.assert_allclose(g.grad.cpu().numpy(), g2.grad.cpu().numpy(), rtol=0,
                                   atol=1e-5 * float(g2.grad.abs().max()))
    # generic matvec Function on a composed operator
    F = fm.Fourier(64)
    P = F * fm.Diag(crandn(rng, 64))
    x = torch.from_numpy(crandn(rng, 64, 2)).cuda().requires_grad_(True)
    say("autograd matvec")
    fa.matvec(P, x).abs().pow(2).sum().backward()
    Pd = np.fft.fft(np.eye(64), axis=0) @ np.diag(P.content[1].vecD)
    ref = 2 * Pd.conj().T @ (Pd @ x.detach().cpu().numpy())
    assert accept.max_rel_error(x.grad.cpu().numpy(), ref) < 1e-5


def test_batched_generator_gradients_match_per_operator(torch_mod):
    """ConvBatch.generator_grads (one batched program) against each
    operator's own generator_grad, on a mixed batch of one- and multi-pass
    Circulant / Toeplitz operators, called twice (the binding is reused)."""
    torch = torch_mod
    from fastmat_amd.batch import ConvBatch
    rng = np.random.default_rng(12)
    spec = small_spec(12, count=10, hi=20000, kmax=4)
    ops, ks = build(spec)
    batch = ConvBatch(ops, ks)
    say("batched generator gradients", len(ops), "ops")
    for rep in range(2):
        xs = [torch.complex(torch.randn(k, op.numCols, device="cuda"),
                            torch.randn(k, op.numCols, device="cuda")) for op, k in zip(ops, ks)]
        gs = [torch.complex(torch.randn(k, op.numRows, device="cuda"),
                            torch.randn(k, op.numRows, device="cuda")) for op, k in zip(ops, ks)]
        got = batch.generator_grads(xs, gs)
        for op, x, g, r in zip(ops, xs, gs, got):
            ref = op.generator_grad(x, g)
            assert accept.max_rel_error(r.cpu().numpy(), ref.cpu().numpy()) < 1e-5, (rep, op)


def test_spectra_overlapped_with_forward_match_sequential(torch_mod):
    """fm_batch_gen_forward_apply (the forward's column transforms issued on
    a side stream while the spectra are formed) gives what load_generators
    followed by the forward apply gives, bit for bit, on a mixed batch with
    new generator values each time, and matches the oracle."""
    torch = torch_mod
    from fastmat_amd.batch import ConvBatch
    rng = np.random.default_rng(14)
    spec = small_spec(14, count=8, hi=20000, kmax=4)
    ops, ks = build(spec)
    batch = ConvBatch(ops, ks)
    xs = [torch.complex(torch.randn(k, op.numCols, device="cuda"),
                        torch.randn(k, op.numCols, device="cuda")) for op, k in zip(ops, ks)]
    ys = batch.new_outputs(False)
    batch.bind(xs, ys, False)
    gens = [torch.empty(e[4].size, dtype=torch.complex64, device="cuda") for e in spec]
    for rep in range(2):
        say("spectra overlapped with the forward, rep", rep)
        new = [crandn(rng, e[4].size) for e in spec]
        for g, h in zip(gens, new):
            g.copy_(torch.from_numpy(h))
        batch.load_generators(gens)
        batch.launch(False)
        ref = [y.clone() for y in ys]
        for y in ys:
            y.zero_()
        batch.launch_with_generators(gens)
        torch.cuda.synchronize()
        for i, (y, r) in enumerate(zip(ys, ref)):
            assert torch.equal(y, r), (rep, i)
        for i, e in enumerate(spec):
            kind, m, n, k, _ = e
            entry = (kind, m, n, k, new[i])
            want = oracle_apply(entry, xs[i].cpu().numpy().T.astype(np.complex128), False)
            assert accept.max_rel_error(ys[i].cpu().numpy().T, want) < TOL32, (rep, i)


def test_autograd_packed_batch_matches_per_tensor(torch_mod):
    """batch_conv_matvec_packed (all blocks in two flat tensors) gives the
    outputs and gradients of batch_conv_matvec, twice in a row."""
    torch = torch_mod
    import fastmat_amd.autograd as fa
    from fastmat_amd.batch import ConvBatch
    rng = np.random.default_rng(13)
    spec = small_spec(13, count=8, hi=15000, kmax=3)
    ops, ks = build(spec)
    batch = ConvBatch(ops, ks)
    n_in, n_out = batch.packed_sizes()
    say("packed autograd batch", len(ops))
    for rep in range(2):
        gens = [torch.from_numpy(e[4]).cuda().requires_grad_(True) for e in spec]
        xp = torch.from_numpy(crandn(rng, n_in)).cuda().requires_grad_(True)
        yp = torch.from_numpy(crandn(rng, n_out)).cuda()
        out = fa.batch_conv_matvec_packed(batch, gens, xp)
        ((out - yp).abs() ** 2).sum().backward()
        # the per-tensor Function on the same data
        gens2 = [g.detach().clone().requires_grad_(True) for g in gens]
        xb = [v.detach().clone().t().requires_grad_(True)
              for v in batch.packed_views(xp.detach(), inputs=True)]
        yb = [v.t() for v in batch.packed_views(yp, inputs=False)]
        outs2 = fa.batch_conv_matvec(batch, gens2, xb)
        sum(((o - y).abs() ** 2).sum() for o, y in zip(outs2, yb)).backward()
        ref_out = torch.cat([o.t().reshape(-1) for o in outs2])
        assert accept.max_rel_error(out.detach().cpu().numpy(), ref_out.detach().cpu().numpy()) < 1e-6
        ref_gx = torch.cat([x.grad.t().reshape(-1) for x in xb])
        assert accept.max_rel_error(xp.grad.cpu().numpy(), ref_gx.cpu().numpy()) < 1e-6
        for g, g2 in zip(gens, gens2):
            assert accept.max_rel_error(g.grad.cpu().numpy(), g2.grad.cpu().numpy()) < 1e-6


def test_packed_autograd_step_replays_in_a_graph(torch_mod):
    """The packed batch autograd step (forward, loss, backward) captured in
    one HIP graph: every replay reproduces the eager step's gradients, with
    eager work and synchronisations between the replays."""
    torch = torch_mod
    import fastmat_amd.autograd as fa
    from fastmat_amd.batch import ConvBatch
    rng = np.random.default_rng(14)
    spec = small_spec(14, count=6, hi=12000, kmax=3)
    ops, ks = build(spec)
    batch = ConvBatch(ops, ks)
    n_in, n_out = batch.packed_sizes()
    gens = [torch.from_numpy(e[4]).cuda().requires_grad_(True) for e in spec]
    xp = torch.from_numpy(crandn(rng, n_in)).cuda().requires_grad_(True)
    yp = torch.from_numpy(crandn(rng, n_out)).cuda()

    def eager():
        for t in gens + [xp]:
            t.grad = None
        out = fa.batch_conv_matvec_packed(batch, gens, xp)
        ((out - yp).abs() ** 2).sum().backward()

    say("packed autograd graph capture")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            eager()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    ref_x = xp.grad.clone()
    ref_g = [g.grad.clone() for g in gens]
    graph = torch.cuda.CUDAGraph()
    for t in gens + [xp]:
        t.grad = None
    with torch.cuda.graph(graph):
        out = fa.batch_conv_matvec_packed(batch, gens, xp)
        ((out - yp).abs() ** 2).sum().backward()
    gx, gg = xp.grad, [g.grad for g in gens]
    for rep in range(3):
        gx.zero_()
        graph.replay()
        torch.cuda.synchronize()
        assert accept.max_rel_error(gx.cpu().numpy(), ref_x.cpu().numpy()) < 1e-6, rep
        for a, b in zip(gg, ref_g):
            assert accept.max_rel_error(a.cpu().numpy(), b.cpu().numpy()) < 1e-6, rep
        _ = (yp * 2).sum().item()  # eager work between replays

"""torch.autograd integration (differentiable structured matvecs).

``matvec(A, x)``: y = A x for any fastmat_amd matrix; the backward pass is
the adjoint matvec A^H g on the device (Matrix.backward, Matrix.pyx:1937-2007).

``conv_matvec(A, gen, x)``: the same for a Circulant / Toeplitz A whose
generator ``gen`` is itself a tensor: the backward pass also returns the
gradient with respect to the generator, computed on the device as an FFT
correlation  grad_t'[d] = sum_cols sum_i g[i] conj(x'[(i - d) mod L])  folded
back onto c (Circulant) or [vecC, vecR] (Toeplitz); see
fm_conv_generator_grad in include/fastmat_amd.h.

``batch_conv_matvec(batch, gens, xs)``: the uneven batch (ConvBatch) with
the generator gradients of all operators in one batched program.

``batch_conv_matvec_packed(batch, gens, xp)``: the same on packed data: the
input blocks of all operators (operator i's (k_i, n_i) block, i.e. its
column-major (n_i, k_i) data) one after another in one flat tensor ``xp``,
the result packed the same way (``batch.packed_sizes()``), so that a loss
over the whole batch is a handful of elementwise operations on two flat
tensors instead of several per operator.

Tensors are user-facing (n,) or (n, k) device tensors; complex gradients
follow PyTorch's convention (grad = dL/dRe + i dL/dIm).
"""
import torch


def _to_block(x):
    was1d = x.dim() == 1
    X = (x.reshape(-1, 1) if was1d else x).t().contiguous()
    return X, was1d


def _from_block(Y, was1d):
    out = Y.t()
    return out[:, 0] if was1d else out


class _Matvec(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, A):
        X, was1d = _to_block(x)
        ctx.A = A
        ctx.was1d = was1d
        return _from_block(A._apply(X, False), was1d)

    @staticmethod
    def backward(ctx, g):
        G, _ = _to_block(g)
        return _from_block(ctx.A._apply(G, True), ctx.was1d), None


def matvec(A, x):
    """Differentiable y = A x (x: device tensor (n,) or (n, k))."""
    return _Matvec.apply(x, A)


class _ConvMatvec(torch.autograd.Function):
    @staticmethod
    def forward(ctx, gen, x, A):
        X, was1d = _to_block(x)
        ctx.A = A
        ctx.gen_complex = gen.is_complex()
        ctx.gen_dtype = gen.dtype
        ctx.was1d = was1d
        ctx.save_for_backward(X)
        return _from_block(A._apply(X, False), was1d)

    @staticmethod
    def backward(ctx, g):
        (X,) = ctx.saved_tensors
        G, _ = _to_block(g)
        A = ctx.A
        gx = _from_block(A._apply(G, True), ctx.was1d) if ctx.needs_input_grad[1] else None
        ggen = None
        if ctx.needs_input_grad[0]:
            ggen = A.generator_grad(X, G)
            if not ctx.gen_complex:
                ggen = ggen.real
            ggen = ggen.to(ctx.gen_dtype)
        return ggen, gx, None


def conv_matvec(A, gen, x):
    """Differentiable y = A(gen) x for a Circulant/Toeplitz ``A`` built from
    the values of ``gen`` (c, or the Toeplitz defining vector [vecC, vecR])."""
    return _ConvMatvec.apply(gen, x, A)


def circulant(gen, **options):
    """Build a Circulant from a (possibly grad-requiring) torch generator."""
    from .circulant import Circulant
    return Circulant(gen.detach().cpu().numpy(), **options)


def toeplitz(gen, rows, **options):
    """Build a Toeplitz from a torch defining vector [vecC (rows), vecR]."""
    from .circulant import Toeplitz
    t = gen.detach().cpu().numpy()
    return Toeplitz(t[:rows], t[rows:], **options)


class _BatchConvMatvec(torch.autograd.Function):
    @staticmethod
    def forward(ctx, batch, *args):
        n = len(batch.ops)
        gens, xs = args[:n], args[n:]
        ctx.gen_complex = [g.is_complex() for g in gens]
        ctx.gen_dtype = [g.dtype for g in gens]
        X = [x.t().contiguous() for x in xs]
        Y = batch.forward(X)
        ctx.batch = batch
        ctx.save_for_backward(*X)
        return tuple(y.t() for y in Y)

    @staticmethod
    def backward(ctx, *gs):
        batch = ctx.batch
        n = len(batch.ops)
        X = ctx.saved_tensors
        G = [g.t().contiguous() for g in gs]
        need_gen = any(ctx.needs_input_grad[1:1 + n])
        need_x = any(ctx.needs_input_grad[1 + n:])
        gx = [None] * n
        if need_x:
            gx = [t.t() for t in batch.backward(G)]
        ggen = [None] * n
        if need_gen:
            raw = batch.generator_grads(list(X), G)
            ggen = [r.to(dt) if cplx else r.real.to(dt)
                    for r, cplx, dt in zip(raw, ctx.gen_complex, ctx.gen_dtype)]
        return (None, *ggen, *gx)


def batch_conv_matvec(batch, gens, xs):
    """Differentiable uneven batch: returns [A_i(gen_i) x_i]."""
    return list(_BatchConvMatvec.apply(batch, *gens, *xs))


def _packed_buffers(batch, dtype, device):
    """Persistent packed buffers of a batch (forward output, output gradient,
    input gradient): fixed addresses, so the batch's bindings are reused from
    call to call; callers receive copies."""
    bufs = getattr(batch, "_packed_bufs", None)
    if bufs is None or bufs[0].dtype != dtype:
        n_in, n_out = batch.packed_sizes()
        bufs = (torch.empty(n_out, dtype=dtype, device=device),
                torch.empty(n_out, dtype=dtype, device=device),
                torch.empty(n_in, dtype=dtype, device=device))
        batch._packed_bufs = bufs
    return bufs


class _BatchConvPacked(torch.autograd.Function):
    @staticmethod
    def forward(ctx, batch, xp, *gens):
        y_buf, _, _ = _packed_buffers(batch, xp.dtype, xp.device)
        batch.apply(batch.packed_views(xp, adjoint=False, inputs=True),
                    batch.packed_views(y_buf, adjoint=False, inputs=False), adjoint=False)
        ctx.batch = batch
        ctx.gen_complex = [g.is_complex() for g in gens]
        ctx.gen_dtype = [g.dtype for g in gens]
        ctx.save_for_backward(xp)
        return y_buf.clone()

    @staticmethod
    def backward(ctx, gyp):
        batch = ctx.batch
        n = len(batch.ops)
        (xp,) = ctx.saved_tensors
        _, g_buf, gx_buf = _packed_buffers(batch, xp.dtype, xp.device)
        g_buf.copy_(gyp)
        G = batch.packed_views(g_buf, adjoint=False, inputs=False)
        gxp = None
        if ctx.needs_input_grad[1]:
            batch.apply(G, batch.packed_views(gx_buf, adjoint=False, inputs=True), adjoint=True)
            gxp = gx_buf.clone()
        ggen = [None] * n
        if any(ctx.needs_input_grad[2:]):
            X = batch.packed_views(xp, adjoint=False, inputs=True)
            raw = batch.generator_grads(X, G)
            ggen = [r.to(dt) if cplx else r.real.to(dt)
                    for r, cplx, dt in zip(raw, ctx.gen_complex, ctx.gen_dtype)]
        return (None, gxp, *ggen)


def batch_conv_matvec_packed(batch, gens, xp):
    """Differentiable uneven batch on packed data: y_p = [A_i(gen_i) x_i]
    packed; xp and the result are flat device tensors of the batch dtype."""
    if xp.dim() != 1 or xp.numel() != batch.packed_sizes()[0]:
        raise ValueError("packed input must be a flat tensor of %d elements"
                         % batch.packed_sizes()[0])
    return _BatchConvPacked.apply(batch, xp, *gens)

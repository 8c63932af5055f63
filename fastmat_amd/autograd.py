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

from . import _device as D


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
        cdt = A.compute_dtype(X.dtype)
        Xc = X.to(D.torch_dtype(cdt))
        # the forward is the operator of gen's current values: the spectrum
        # is recomputed from gen on the device (a private plan of A's)
        h, tok = A.load_generator(gen, cdt)
        ctx.A, ctx.cdt, ctx.tok = A, cdt, tok
        ctx.gen_complex = gen.is_complex()
        ctx.gen_dtype = gen.dtype
        ctx.x_dtype = x.dtype
        ctx.was1d = was1d
        ctx.save_for_backward(Xc, gen)
        return _from_block(A._run_plan(h, Xc, False), was1d)

    @staticmethod
    def backward(ctx, g):
        Xc, gen = ctx.saved_tensors
        A, cdt = ctx.A, ctx.cdt
        G, _ = _to_block(g)
        G = G.to(D.torch_dtype(cdt)).contiguous()
        gx = None
        if ctx.needs_input_grad[1]:
            h = A._gen_plan(cdt)
            if A.generator_token(cdt) != ctx.tok:
                # another forward loaded a different generator since
                h, _ = A.load_generator(gen, cdt)
            gx = A._run_plan(h, G, True)
            if not ctx.x_dtype.is_complex:
                gx = gx.real
            gx = _from_block(gx.to(ctx.x_dtype), ctx.was1d)
        ggen = None
        if ctx.needs_input_grad[0]:
            ggen = A.generator_grad(Xc, G)
            if not ctx.gen_complex:
                ggen = ggen.real
            ggen = ggen.to(ctx.gen_dtype)
        return ggen, gx, None


def conv_matvec(A, gen, x):
    """Differentiable y = A(gen) x: the Circulant/Toeplitz of the structure of
    ``A`` (sizes, kind) whose generator is the device tensor ``gen`` (c, or the
    Toeplitz defining vector [vecC, vecR]).  Every call computes the operator
    from gen's current values on the device, so updates of gen (an optimizer
    step) take effect; ``A`` itself keeps its constructor's generator."""
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
        ctx.tok = batch.load_generators(gens)
        X = [x.t().contiguous() for x in xs]
        Y = batch.forward(X)
        ctx.batch = batch
        ctx.n = n
        ctx.save_for_backward(*X, *gens)
        return tuple(y.t() for y in Y)

    @staticmethod
    def backward(ctx, *gs):
        batch = ctx.batch
        n = ctx.n
        saved = ctx.saved_tensors
        X, gens = saved[:n], saved[n:]
        G = [g.t().contiguous() for g in gs]
        need_gen = any(ctx.needs_input_grad[1:1 + n])
        need_x = any(ctx.needs_input_grad[1 + n:])
        gx = [None] * n
        if need_x:
            if batch.generator_token != ctx.tok:
                batch.load_generators(gens)
            gx = [t.t() for t in batch.backward(G)]
        ggen = [None] * n
        if need_gen:
            raw = batch.generator_grads(list(X), G)
            ggen = [r.to(dt) if cplx else r.real.to(dt)
                    for r, cplx, dt in zip(raw, ctx.gen_complex, ctx.gen_dtype)]
        return (None, *ggen, *gx)


def batch_conv_matvec(batch, gens, xs):
    """Differentiable uneven batch: returns [A_i(gen_i) x_i], every operator
    computed from its generator's current values on the device."""
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
        ctx.tok = batch.load_generators(gens)
        batch.apply(batch.packed_views(xp, adjoint=False, inputs=True),
                    batch.packed_views(y_buf, adjoint=False, inputs=False), adjoint=False)
        ctx.fwd_count = batch.apply_count[0]
        ctx.batch = batch
        ctx.gen_complex = [g.is_complex() for g in gens]
        ctx.gen_dtype = [g.dtype for g in gens]
        ctx.save_for_backward(xp, *gens)
        return y_buf.clone()

    @staticmethod
    def backward(ctx, gyp):
        batch = ctx.batch
        n = len(batch.ops)
        saved = ctx.saved_tensors
        xp, gens = saved[0], saved[1:]
        _, g_buf, gx_buf = _packed_buffers(batch, xp.dtype, xp.device)
        g_buf.copy_(gyp)
        G = batch.packed_views(g_buf, adjoint=False, inputs=False)
        gxp = None
        if ctx.needs_input_grad[1]:
            if batch.generator_token != ctx.tok:
                batch.load_generators(gens)
            batch.apply(G, batch.packed_views(gx_buf, adjoint=False, inputs=True), adjoint=True)
            gxp = gx_buf.clone()
        ggen = [None] * n
        if any(ctx.needs_input_grad[2:]):
            X = batch.packed_views(xp, adjoint=False, inputs=True)
            # the forward of this step left x's column transforms in the
            # batch workspace and the adjoint apply above g's: reuse them
            # unless another forward ran in between
            reuse = (ctx.needs_input_grad[1] and
                     batch.apply_count[0] == ctx.fwd_count)
            raw = batch.generator_grads(X, G, reuse=reuse)
            ggen = [(r if r.dtype == dt else r.to(dt)) if cplx else r.real.to(dt)
                    for r, cplx, dt in zip(raw, ctx.gen_complex, ctx.gen_dtype)]
        return (None, gxp, *ggen)


def batch_conv_matvec_packed(batch, gens, xp):
    """Differentiable uneven batch on packed data: y_p = [A_i(gen_i) x_i]
    packed; xp and the result are flat device tensors of the batch dtype."""
    if xp.dim() != 1 or xp.numel() != batch.packed_sizes()[0]:
        raise ValueError("packed input must be a flat tensor of %d elements"
                         % batch.packed_sizes()[0])
    return _BatchConvPacked.apply(batch, xp, *gens)


class _BatchLeastSquares(torch.autograd.Function):
    """loss = sum_i ||A_i(gen_i) x_i - y_i||^2 over a packed batch in one
    least-squares binding of the batch (fm_batch_bind_lsq): the forward forms
    the residual 2 (A x - y) (dLoss/d(A x)) and the loss inside its last pass,
    and each multi-pass operator's last pass also runs the adjoint's column
    transform of that residual; the backward is the rest of the adjoint
    apply and the batched generator gradient, reusing the column transforms
    of x and of the residual that the forward left in the workspace."""

    @staticmethod
    def _bind(batch, xp, yp):
        r_buf, _, gx_buf = _packed_buffers(batch, xp.dtype, xp.device)
        X = batch.packed_views(xp, adjoint=False, inputs=True)
        Y = batch.packed_views(yp, adjoint=False, inputs=False)
        R = batch.packed_views(r_buf, adjoint=False, inputs=False)
        GX = batch.packed_views(gx_buf, adjoint=False, inputs=True)
        batch.bind_lsq(X, Y, R, GX, scale=2.0)
        return X, R, gx_buf

    @staticmethod
    def forward(ctx, batch, xp, yp, *gens):
        _BatchLeastSquares._bind(batch, xp, yp)
        # spectra from the generators, then the forward (its column
        # transforms of x overlap the spectrum computation)
        ctx.tok = batch.launch_with_generators(gens)
        ctx.fwd_count = batch.apply_count[0]
        loss = getattr(batch, "_lsq_loss", None)
        if loss is None or loss.device != xp.device:
            loss = batch._lsq_loss = torch.empty(1, dtype=torch.float64, device=xp.device)
        batch.residual_loss(loss)
        ctx.batch = batch
        ctx.gen_complex = [g.is_complex() for g in gens]
        ctx.gen_dtype = [g.dtype for g in gens]
        ctx.save_for_backward(xp, yp, *gens)
        real_t = torch.float32 if xp.dtype == torch.complex64 else torch.float64
        return loss[0].to(real_t)

    @staticmethod
    def backward(ctx, gl):
        batch = ctx.batch
        n = len(batch.ops)
        saved = ctx.saved_tensors
        xp, yp, gens = saved[0], saved[1], saved[2:]
        need_x, need_g = ctx.needs_input_grad[1], any(ctx.needs_input_grad[3:])
        ggen = [None] * n
        if not (need_x or need_g):
            return (None, None, None, *ggen)
        X, R, gx_buf = _BatchLeastSquares._bind(batch, xp, yp)
        if batch.generator_token != ctx.tok or batch.apply_count[0] != ctx.fwd_count:
            # another forward ran on this batch since: run this step's
            # forward again (its last pass holds the adjoint's column
            # transforms of the residual)
            batch.launch_with_generators(gens)
        gxp = adj = None
        if need_x:
            # the adjoint writes gl * A^H r straight into a new tensor (the
            # bound gradient buffer stays as it is)
            real_t = torch.float32 if xp.dtype == torch.complex64 else torch.float64
            glr = gl.detach().to(real_t).reshape(1).contiguous()
            gxp = torch.empty_like(gx_buf)
            adj = (gxp.data_ptr() - gx_buf.data_ptr(), glr)
            if not need_g:
                batch.launch(True, out_shift=adj[0], out_scale=adj[1])
        if need_g:
            # with the x gradient: the adjoint and the generator gradient's
            # correlations in one backward (fm_batch_lsq_backward), both
            # scaled by gl on the device; without it, the gradient alone
            scale = gl.to(xp.dtype) if adj is None else None
            raw = batch.generator_grads(X, R, reuse=True, scale=scale, lsq_adjoint=adj)
            ggen = [(r if r.dtype == dt else r.to(dt)) if cplx else r.real.to(dt)
                    for r, cplx, dt in zip(raw, ctx.gen_complex, ctx.gen_dtype)]
        return (None, gxp, None, *ggen)


def batch_least_squares_packed(batch, gens, xp, yp):
    """Differentiable loss sum_i ||A_i(gen_i) x_i - y_i||^2 of a batch on
    packed data (xp the inputs, yp the targets, as batch_conv_matvec_packed
    packs them), with gradients for xp and every generator.  Equals
    ((batch_conv_matvec_packed(batch, gens, xp) - yp).abs() ** 2).sum(), with
    the residual and its sum formed inside the batch's forward apply."""
    n_in, n_out = batch.packed_sizes()
    if xp.dim() != 1 or xp.numel() != n_in or yp.dim() != 1 or yp.numel() != n_out:
        raise ValueError("packed inputs of %d and targets of %d elements expected"
                         % (n_in, n_out))
    if yp.dtype != xp.dtype:
        raise TypeError("targets must have the inputs' dtype")
    return _BatchLeastSquares.apply(batch, xp, yp.detach().contiguous(), *gens)

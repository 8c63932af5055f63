"""Uneven batches of Circulant / Toeplitz operators in a few persistent launches.

A fastmat user applies many operators by calling ``forward``/``backward`` per
operator (Matrix.pyx:1844-2007 once per matrix).  ``ConvBatch`` applies all
of them together: the library groups every operator's passes by phase and
transform length, and runs one persistent kernel per group over a device work
table (see include/fastmat_amd.h, fm_batch_*).  Buffers are bound once per
set of device pointers; the apply itself only launches kernels, so it can be
captured in a HIP graph (``ConvBatch.graph``) and replayed.

The launch shape (workgroups per CU of the persistent grid) is chosen by
``Autotuner``, which times the candidates on the actual batch with HIP events
and caches the winner per batch signature (in memory, and in a JSON file when
FASTMAT_AMD_TUNE_CACHE names one).
"""
import ctypes
import hashlib
import json
import os

import numpy as np
import torch

from . import _device as D
from . import _lib
from .circulant import _ConvBase
from .fourier import precision_of


class Autotuner:
    """Times candidate launch configurations and caches the fastest."""

    def __init__(self, path=None):
        self.path = path if path is not None else os.environ.get("FASTMAT_AMD_TUNE_CACHE")
        self.cache = {}
        if self.path and os.path.exists(self.path):
            try:
                with open(self.path) as f:
                    self.cache = json.load(f)
            except (OSError, ValueError):
                self.cache = {}

    def best(self, key, candidates, run, reps=5):
        """Return the candidate with the lowest mean time of run(candidate)."""
        if key in self.cache:
            cached = self.cache[key]["best"]
            cached = tuple(cached) if isinstance(cached, list) else cached
            if cached in candidates:
                return cached
        times = {}
        for c in candidates:
            run(c)  # warm-up / first-touch
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                run(c)
            e1.record()
            e1.synchronize()
            times[c] = e0.elapsed_time(e1) / reps
        best = min(times, key=times.get)
        self.cache[key] = {"best": best, "ms": {str(k): v for k, v in times.items()}}
        if self.path:
            try:
                os.makedirs(os.path.dirname(os.path.abspath(self.path)), exist_ok=True)
                with open(self.path, "w") as f:
                    json.dump(self.cache, f, indent=1)
            except OSError:
                pass
        return best


_default_tuner = None


def default_tuner():
    global _default_tuner
    if _default_tuner is None:
        _default_tuner = Autotuner()
    return _default_tuner


class ConvBatch:
    """Apply a list of Circulant/Toeplitz operators with per-operator column
    counts ``ks`` in one batched call.

    ``dtype`` (complex64 default) is the compute and data type of every
    operator in the batch.
    """

    def __init__(self, ops, ks, dtype=np.complex64):
        if len(ops) != len(ks) or not ops:
            raise ValueError("ConvBatch: need one column count per operator")
        for op in ops:
            if not isinstance(op, _ConvBase):
                raise TypeError("ConvBatch: operators must be Circulant or Toeplitz")
        self.ops = list(ops)
        self.ks = [int(k) for k in ks]
        self.dtype = np.dtype(dtype)
        prec = precision_of(self.dtype)
        # private copies of the operators' plans: the batch's spectra can be
        # rewritten from device generators (load_generators) while every
        # operator keeps its own
        self._own = []
        for op in ops:
            h = ctypes.c_void_p()
            _lib.call("fm_conv_plan_clone", op.plan_handle(self.dtype), D.stream(),
                      ctypes.byref(h))
            self._own.append(h)
        arr = (ctypes.c_void_p * len(ops))(*[h.value for h in self._own])
        karr = (ctypes.c_int64 * len(ops))(*self.ks)
        self._h = ctypes.c_void_p()
        _lib.call("fm_batch_plan_create", len(ops), arr, karr, ctypes.byref(self._h))
        need = ctypes.c_int64()
        _lib.call("fm_batch_workspace_size", self._h, ctypes.byref(need))
        self._ws = D.workspace(need.value)
        self._ws_bytes = need.value
        self._bound = [None, None]
        self.grid_per_cu = 2
        self.group_mb = 96
        self.lanes = 2
        self._prec = prec

    def __del__(self):
        try:
            if _lib._lib is not None:
                if self._h:
                    _lib._lib.fm_batch_plan_destroy(self._h)
                for h in getattr(self, "_own", []):
                    _lib._lib.fm_conv_plan_destroy(h)
        except Exception:
            pass

    # ------------------------------------------------- device generators
    @property
    def generator_token(self):
        """Changes with every load_generators call."""
        return getattr(self, "_spec_tok", 0)

    def load_generators(self, gens):
        """Set every operator's spectrum from its device generator gens[i]
        (c, or [vecC, vecR]; one real or complex float type for all): the
        batch then applies the operators of the generators' current values
        (Circulant.pyx:130, Toeplitz.pyx:253-255), as one embedding launch and
        one batched FFT program (fm_batch_gen_*).  Returns a token that
        changes with every load."""
        self._bind_generators(gens)
        _lib.call("fm_batch_gen_apply", self._h, D.stream())
        self._spec_tok = self.generator_token + 1
        return self._spec_tok

    def _bind_generators(self, gens):
        if len(gens) != len(self.ops):
            raise ValueError("ConvBatch: need one generator per operator")
        gts = []
        for op, g in zip(self.ops, gens):
            if g.dim() != 1 or g.numel() != op._gen128.size:
                raise ValueError("ConvBatch: generator of %d elements expected, got shape %s"
                                 % (op._gen128.size, tuple(g.shape)))
            gts.append(g.detach())
        dts = {D.np_dtype(g.dtype) for g in gts}
        ok = (np.dtype(np.float32), np.dtype(np.float64), np.dtype(np.complex64),
              np.dtype(np.complex128))
        if len(dts) != 1 or next(iter(dts)) not in ok:
            gts = [g.to(D.torch_dtype(self.dtype)) for g in gts]
        gts = [g.contiguous() for g in gts]
        gdt = D.fm_type(gts[0].dtype)
        key = (gdt,) + tuple(g.data_ptr() for g in gts)
        if getattr(self, "_sbound", None) != key:
            if getattr(self, "_sws", None) is None:
                need = ctypes.c_int64()
                _lib.call("fm_batch_gen_workspace_size", self._h, ctypes.byref(need))
                self._sws = D.workspace(need.value)
                self._sws_bytes = need.value
            n = len(gts)
            pg = (ctypes.c_void_p * n)(*[g.data_ptr() for g in gts])
            _lib.call("fm_batch_gen_bind", self._h, gdt, pg, _lib.ptr(self._sws), self._sws_bytes,
                      D.stream())
            self._sbound = key
        self._sgens = gts  # the bound generators stay alive with the binding

    def launch_with_generators(self, gens):
        """load_generators(gens) then launch(False) (the bound forward), with
        the forward's column transforms of its inputs overlapping the
        spectrum computation (fm_batch_gen_forward_apply).  Returns the
        generator token."""
        self._bind_generators(gens)
        _lib.call("fm_batch_gen_forward_apply", self._h, D.stream())
        self._spec_tok = self.generator_token + 1
        self.apply_count[0] += 1
        return self._spec_tok

    @property
    def signature(self):
        desc = ",".join("%s%dx%d:%d" % (type(op).__name__[0], op.numRows, op.numCols, k)
                        for op, k in zip(self.ops, self.ks))
        return "batch-%s-%s" % (self.dtype.name, hashlib.sha1(desc.encode()).hexdigest()[:16])

    def algorithmic_bytes(self, adjoint=False):
        """Bytes one apply must move: read every input block and generator,
        write every output block."""
        esz = self.dtype.itemsize
        tot = 0
        for op, k in zip(self.ops, self.ks):
            tot += ((op.numRows + op.numCols) * k + op._gen128.size) * esz
        return tot

    # ----------------------------------------------------------- buffers
    def packed_sizes(self):
        """(elements of all forward inputs, elements of all forward outputs)
        when the (k_i, n_i) blocks are stored one after another."""
        return (int(sum(op.numCols * k for op, k in zip(self.ops, self.ks))),
                int(sum(op.numRows * k for op, k in zip(self.ops, self.ks))))

    def packed_views(self, flat, adjoint=False, inputs=True):
        """(k_i, len_i) block views into a flat tensor holding every
        operator's block in turn: the forward inputs (len = cols) or outputs
        (len = rows); with adjoint=True the roles swap."""
        use_cols = inputs != adjoint
        out, off = [], 0
        for op, k in zip(self.ops, self.ks):
            w = op.numCols if use_cols else op.numRows
            out.append(flat[off:off + k * w].view(k, w))
            off += k * w
        if off != flat.numel():
            raise ValueError("ConvBatch: packed tensor has %d elements, the blocks %d"
                             % (flat.numel(), off))
        return out

    def new_inputs(self, adjoint=False, fill=None):
        """Allocate input blocks (k_i, n_i) for a direction."""
        out = []
        for op, k in zip(self.ops, self.ks):
            n = op.numRows if adjoint else op.numCols
            t = D.empty(k, n, self.dtype)
            if fill is not None:
                t.copy_(fill(k, n))
            out.append(t)
        return out

    def new_outputs(self, adjoint=False):
        return [D.empty(k, (op.numCols if adjoint else op.numRows), self.dtype)
                for op, k in zip(self.ops, self.ks)]

    def bind(self, ins, outs, adjoint=False):
        key = tuple(t.data_ptr() for t in ins) + tuple(t.data_ptr() for t in outs)
        if self._bound[int(adjoint)] == key:
            return
        for i, (op, k) in enumerate(zip(self.ops, self.ks)):
            nin = op.numRows if adjoint else op.numCols
            mout = op.numCols if adjoint else op.numRows
            if tuple(ins[i].shape) != (k, nin) or tuple(outs[i].shape) != (k, mout):
                raise ValueError("ConvBatch: buffer %d has the wrong shape" % i)
            if not (ins[i].is_contiguous() and outs[i].is_contiguous()):
                raise ValueError("ConvBatch: buffers must be contiguous blocks")
            if D.np_dtype(ins[i].dtype) != self.dtype or D.np_dtype(outs[i].dtype) != self.dtype:
                raise TypeError("ConvBatch: buffers must be %s" % self.dtype)
        n = len(self.ops)
        pin = (ctypes.c_void_p * n)(*[t.data_ptr() for t in ins])
        pout = (ctypes.c_void_p * n)(*[t.data_ptr() for t in outs])
        ldin = (ctypes.c_int64 * n)(*[t.shape[1] for t in ins])
        ldout = (ctypes.c_int64 * n)(*[t.shape[1] for t in outs])
        ws = None if self._ws is None else _lib.ptr(self._ws)
        _lib.call("fm_batch_bind", self._h, int(adjoint), pin, ldin, pout, ldout, ws,
                  self._ws_bytes, D.stream())
        self._bound[int(adjoint)] = key
        self._keep = getattr(self, "_keep", {})
        self._keep[int(adjoint)] = (list(ins), list(outs))

    def bind_residual(self, ins, targets, outs, scale=2.0):
        """Bind the forward direction in residual mode: outs[i] receives
        scale * (A_i ins[i] - targets[i]) and ``residual_loss`` the sum of
        |A_i ins[i] - targets[i]|^2 (fm_batch_bind_residual)."""
        key = ("res", float(scale)) + tuple(t.data_ptr() for t in ins) + \
            tuple(t.data_ptr() for t in targets) + tuple(t.data_ptr() for t in outs)
        if self._bound[0] == key:
            return
        for i, (op, k) in enumerate(zip(self.ops, self.ks)):
            if (tuple(ins[i].shape) != (k, op.numCols) or tuple(outs[i].shape) != (k, op.numRows)
                    or tuple(targets[i].shape) != (k, op.numRows)):
                raise ValueError("ConvBatch: buffer %d has the wrong shape" % i)
            for t in (ins[i], targets[i], outs[i]):
                if not t.is_contiguous():
                    raise ValueError("ConvBatch: buffers must be contiguous blocks")
                if D.np_dtype(t.dtype) != self.dtype:
                    raise TypeError("ConvBatch: buffers must be %s" % self.dtype)
        n = len(self.ops)
        pin = (ctypes.c_void_p * n)(*[t.data_ptr() for t in ins])
        ptg = (ctypes.c_void_p * n)(*[t.data_ptr() for t in targets])
        pout = (ctypes.c_void_p * n)(*[t.data_ptr() for t in outs])
        ldin = (ctypes.c_int64 * n)(*[t.shape[1] for t in ins])
        ldt = (ctypes.c_int64 * n)(*[t.shape[1] for t in targets])
        ldout = (ctypes.c_int64 * n)(*[t.shape[1] for t in outs])
        ws = None if self._ws is None else _lib.ptr(self._ws)
        _lib.call("fm_batch_bind_residual", self._h, pin, ldin, ptg, ldt, pout, ldout,
                  float(scale), ws, self._ws_bytes, D.stream())
        self._bound[0] = key
        self._keep = getattr(self, "_keep", {})
        self._keep[0] = (list(ins), list(targets), list(outs))

    def bind_lsq(self, ins, targets, residuals, grads_x, scale=2.0):
        """Bind both directions for a least-squares step
        (fm_batch_bind_lsq): the forward in residual mode (ins, targets,
        residuals), the adjoint on the residuals into grads_x, with each
        multi-pass operator's forward column inverse and adjoint column
        transform fused (its residual is formed on the chip and never
        written).  launch(False) then launch(True) run the step's applies;
        generator_grads(ins, residuals, reuse=True) its gradients."""
        key = ("lsq", float(scale)) + tuple(t.data_ptr() for t in ins) + \
            tuple(t.data_ptr() for t in targets) + tuple(t.data_ptr() for t in residuals) + \
            tuple(t.data_ptr() for t in grads_x)
        if self._bound[0] == key and self._bound[1] == key:
            return
        for i, (op, k) in enumerate(zip(self.ops, self.ks)):
            shapes = ((ins[i], op.numCols), (targets[i], op.numRows), (residuals[i], op.numRows),
                      (grads_x[i], op.numCols))
            for t, w in shapes:
                if tuple(t.shape) != (k, w):
                    raise ValueError("ConvBatch: buffer %d has the wrong shape" % i)
                if not t.is_contiguous():
                    raise ValueError("ConvBatch: buffers must be contiguous blocks")
                if D.np_dtype(t.dtype) != self.dtype:
                    raise TypeError("ConvBatch: buffers must be %s" % self.dtype)
        n = len(self.ops)

        def ptrs(ts):
            return (ctypes.c_void_p * n)(*[t.data_ptr() for t in ts])

        def lds(ts):
            return (ctypes.c_int64 * n)(*[t.shape[1] for t in ts])
        ws = None if self._ws is None else _lib.ptr(self._ws)
        _lib.call("fm_batch_bind_lsq", self._h, ptrs(ins), lds(ins), ptrs(targets), lds(targets),
                  ptrs(residuals), lds(residuals), ptrs(grads_x), lds(grads_x), float(scale), ws,
                  self._ws_bytes, D.stream())
        self._bound[0] = self._bound[1] = key
        self._keep = getattr(self, "_keep", {})
        self._keep[0] = (list(ins), list(targets), list(residuals), list(grads_x))
        self._keep[1] = self._keep[0]

    def residual_loss(self, out):
        """Write the residual-mode loss of the last forward apply to the
        device float64 tensor ``out`` (one element)."""
        if out.dtype != torch.float64 or out.numel() < 1 or not out.is_cuda:
            raise TypeError("residual_loss needs a float64 device tensor")
        _lib.call("fm_batch_residual_loss", self._h, _lib.ptr(out), D.stream())
        return out

    def bind_views(self, ins, outs, adjoint=False):
        """Bind column slices of wider blocks (unit element stride, row
        stride >= width) without copying them; the rows' stride is passed as
        each operator's leading dimension.  Used by BlockDiag."""
        key = tuple((t.data_ptr(), D.ld(t)) for t in ins) + \
            tuple((t.data_ptr(), D.ld(t)) for t in outs)
        if self._bound[int(adjoint)] == key:
            return
        for i, (op, k) in enumerate(zip(self.ops, self.ks)):
            nin = op.numRows if adjoint else op.numCols
            mout = op.numCols if adjoint else op.numRows
            if tuple(ins[i].shape) != (k, nin) or tuple(outs[i].shape) != (k, mout):
                raise ValueError("ConvBatch: view %d has the wrong shape" % i)
            for t, w in ((ins[i], nin), (outs[i], mout)):
                if t.stride(1) != 1 or D.ld(t) < w:
                    raise ValueError("ConvBatch: view %d is not a column slice" % i)
                if D.np_dtype(t.dtype) != self.dtype:
                    raise TypeError("ConvBatch: buffers must be %s" % self.dtype)
        n = len(self.ops)
        pin = (ctypes.c_void_p * n)(*[t.data_ptr() for t in ins])
        pout = (ctypes.c_void_p * n)(*[t.data_ptr() for t in outs])
        ldin = (ctypes.c_int64 * n)(*[D.ld(t) for t in ins])
        ldout = (ctypes.c_int64 * n)(*[D.ld(t) for t in outs])
        ws = None if self._ws is None else _lib.ptr(self._ws)
        _lib.call("fm_batch_bind", self._h, int(adjoint), pin, ldin, pout, ldout, ws,
                  self._ws_bytes, D.stream())
        # no reference to the views is kept: the binding is reused only
        # while the same addresses are handed in again
        self._bound[int(adjoint)] = key

    def generator_grads(self, xs, gs, reuse=False, scale=None, lsq_adjoint=None):
        """Gradients of a real loss with respect to every operator's
        generator, given the forward inputs xs[i] (k_i, cols_i) and the
        output gradients gs[i] (k_i, rows_i) as contiguous blocks: what
        ``op.generator_grad(x, g)`` returns per operator, for the whole batch
        in one batched program (fm_batch_grad_*).  reuse=True reads the
        column transforms the last forward apply (of xs) and adjoint apply
        (of gs) left in the workspace instead of recomputing them.  ``scale``
        (a device scalar tensor) multiplies all gradients in one pass.
        lsq_adjoint=(out_shift, out_scale), on a least-squares binding
        (bind_lsq, with xs its inputs and gs its residuals), also runs the
        adjoint apply as launch(True, out_shift, out_scale) does, its
        launches sharing the lanes with the gradient's
        (fm_batch_lsq_backward), and scales the gradients by out_scale (a
        real device scalar; ``scale`` is not used then)."""
        n = len(self.ops)
        for i, (op, k) in enumerate(zip(self.ops, self.ks)):
            if tuple(xs[i].shape) != (k, op.numCols) or tuple(gs[i].shape) != (k, op.numRows):
                raise ValueError("ConvBatch: gradient block %d has the wrong shape" % i)
            if not (xs[i].is_contiguous() and gs[i].is_contiguous()):
                raise ValueError("ConvBatch: gradient blocks must be contiguous")
            if D.np_dtype(xs[i].dtype) != self.dtype or D.np_dtype(gs[i].dtype) != self.dtype:
                raise TypeError("ConvBatch: gradient blocks must be %s" % self.dtype)
        # the program writes into one persistent buffer (stable addresses, so
        # the binding is reused); the caller gets views of one copy of it
        if getattr(self, "_gout", None) is None:
            lens = [int(op._gen128.size) for op in self.ops]
            self._goff = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
            self._gout = D.empty(1, int(self._goff[-1]), self.dtype)[0]
        off = self._goff
        grads = [self._gout[int(off[i]):int(off[i + 1])] for i in range(n)]
        if getattr(self, "_gws", None) is None:
            need = ctypes.c_int64()
            _lib.call("fm_batch_grad_workspace_size", self._h, ctypes.byref(need))
            self._gws = D.workspace(need.value)
            self._gws_bytes = need.value
        # the gradient reads and writes the directions' pass-A intermediates:
        # both directions need a binding (any buffers)
        if self._bound[0] is None:
            self.bind(xs, self.new_outputs(False), False)
        if self._bound[1] is None:
            self.bind(gs, self.new_outputs(True), True)
        key = tuple(t.data_ptr() for t in xs) + tuple(t.data_ptr() for t in gs) + \
            tuple(t.data_ptr() for t in grads)
        if getattr(self, "_gbound", None) != key:
            px = (ctypes.c_void_p * n)(*[t.data_ptr() for t in xs])
            pg = (ctypes.c_void_p * n)(*[t.data_ptr() for t in gs])
            po = (ctypes.c_void_p * n)(*[t.data_ptr() for t in grads])
            lx = (ctypes.c_int64 * n)(*[t.shape[1] for t in xs])
            lg = (ctypes.c_int64 * n)(*[t.shape[1] for t in gs])
            _lib.call("fm_batch_grad_bind", self._h, px, lx, pg, lg, po, _lib.ptr(self._gws),
                      self._gws_bytes, D.stream())
            self._gbound = key
        if lsq_adjoint is not None:
            # the fold writes the gradients, scaled by the same device
            # scalar as the adjoint's output, straight into a new tensor
            shift, sc = lsq_adjoint
            out = torch.empty_like(self._gout)
            _lib.call("fm_batch_lsq_backward", self._h, int(shift), self._scale_ptr(sc),
                      out.data_ptr() - self._gout.data_ptr(), D.stream())
            self.apply_count[1] += 1
            return [out[int(off[i]):int(off[i + 1])] for i in range(n)]
        _lib.call("fm_batch_grad_apply_reuse" if reuse else "fm_batch_grad_apply", self._h,
                  D.stream())
        out = self._gout.clone() if scale is None else self._gout * scale
        return [out[int(off[i]):int(off[i + 1])] for i in range(n)]

    def set_grid(self, workgroups_per_cu):
        _lib.call("fm_batch_set_grid", self._h, int(workgroups_per_cu))
        self.grid_per_cu = int(workgroups_per_cu)

    @property
    def schedule(self):
        """(workgroups per CU, group budget in MB, lanes)."""
        return (self.grid_per_cu, self.group_mb, self.lanes)

    def set_schedule(self, workgroups_per_cu, group_mb, lanes):
        """Persistent grid, budget of one group's four-step intermediates
        (0: all multi-pass operators in one group) and lanes (streams) the
        groups alternate between.  A new grouping rebinds on the next apply."""
        _lib.call("fm_batch_set_schedule", self._h, int(workgroups_per_cu), int(group_mb),
                  int(lanes))
        if (int(group_mb), int(lanes)) != (self.group_mb, self.lanes):
            self._bound = [None, None]
            self._gbound = None
        self.grid_per_cu, self.group_mb, self.lanes = (int(workgroups_per_cu), int(group_mb),
                                                       int(lanes))

    def _scale_ptr(self, out_scale):
        """Device address of a one-element real scale (float32 for complex64
        batches, float64 for complex128), or None."""
        if out_scale is None:
            return None
        want = torch.float32 if self.dtype == np.complex64 else torch.float64
        if (out_scale.dtype != want or out_scale.numel() != 1 or
                not out_scale.is_contiguous() or out_scale.device.type != "cuda"):
            raise TypeError("out_scale must be one contiguous %s element on the device" % want)
        return out_scale.data_ptr()

    @property
    def apply_count(self):
        """Launches issued per direction (forward, adjoint)."""
        if not hasattr(self, "_napply"):
            self._napply = [0, 0]
        return self._napply

    def launch(self, adjoint=False, out_shift=0, out_scale=None):
        """Launch the bound direction (kernels only; graph-capturable).
        out_shift (bytes, a whole number of elements) moves every output
        block the direction writes to the same place in another buffer, and
        out_scale (a one-element real device tensor: float32 for complex64,
        float64 for complex128) multiplies what is written
        (fm_batch_apply_redirected)."""
        if out_shift or out_scale is not None:
            _lib.call("fm_batch_apply_redirected", self._h, int(adjoint), int(out_shift),
                      self._scale_ptr(out_scale), D.stream())
        else:
            _lib.call("fm_batch_apply", self._h, int(adjoint), D.stream())
        self.apply_count[int(adjoint)] += 1

    def apply(self, ins, outs=None, adjoint=False):
        if outs is None:
            outs = self.new_outputs(adjoint)
        self.bind(ins, outs, adjoint)
        self.launch(adjoint)
        return outs

    def forward(self, ins, outs=None):
        return self.apply(ins, outs, False)

    def backward(self, ins, outs=None):
        return self.apply(ins, outs, True)

    # ---------------------------------------------------------- tuning
    # (workgroups per CU, group budget MB, lanes): ungrouped schedules (the
    # multi-pass operators' passes as one launch per length class), groups
    # sized for the memory-side cache on one or two lanes, and budgets of
    # 0.5-1.5 GB that split the longest operators from the rest so that two
    # or three groups' pass chains run side by side on as many lanes
    DEFAULT_SCHEDULES = ((2, 0, 2), (3, 0, 2), (4, 0, 2), (2, 48, 2), (3, 48, 2),
                         (2, 96, 2), (3, 96, 2), (4, 96, 2), (2, 128, 1), (3, 192, 1),
                         (4, 1024, 2), (4, 1536, 2), (3, 1024, 2), (6, 1536, 2),
                         (4, 512, 3), (4, 768, 3), (4, 1024, 3), (6, 768, 3), (6, 512, 3),
                         (8, 512, 3), (8, 768, 3))

    def tune(self, ins, outs, adjoint=False, tuner=None, candidates=(1, 2, 3, 4)):
        """Pick the persistent grid (workgroups per CU) of one direction."""
        tuner = tuner or default_tuner()
        self.bind(ins, outs, adjoint)

        def run(c):
            self.set_grid(c)
            self.launch(adjoint)

        best = tuner.best(self.signature + ("-adj" if adjoint else "-fwd"), list(candidates), run)
        self.set_grid(best)
        return best

    def tune_schedule(self, xs, ys, gs, zs, tuner=None, candidates=None):
        """Pick the schedule (grid, group budget, lanes) that minimises one
        forward plus one adjoint apply on the given buffers, timed on the
        device; cached per batch signature."""
        tuner = tuner or default_tuner()
        cands = [tuple(c) for c in (candidates or self.DEFAULT_SCHEDULES)]

        def run(c):
            self.set_schedule(*c)
            self.bind(xs, ys, False)
            self.bind(gs, zs, True)
            self.launch(False)
            self.launch(True)

        best = tuner.best(self.signature + "-schedule", cands, run)
        best = tuple(best)
        self.set_schedule(*best)
        self.bind(xs, ys, False)
        self.bind(gs, zs, True)
        return best

    # ----------------------------------------------------------- graphs
    def graph(self, ins, outs, adjoint=False):
        """Capture the bound apply into a HIP graph; returns the graph (call
        ``.replay()``)."""
        self.bind(ins, outs, adjoint)
        self.launch(adjoint)  # first launch outside capture
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.launch(adjoint)
        torch.cuda.current_stream().wait_stream(s)
        return g


def plain_torch_path(ops, xs, adjoint=False):
    """The plain per-operator GPU path a user would write with torch.fft:
    for every operator, zero-pad, FFT, multiply by the spectrum, inverse FFT
    and crop (the Product(F.H, Diag, F) structure, Circulant.pyx:127-137),
    one operator after another.  Used as the comparison baseline."""
    outs = []
    for op, X in zip(ops, xs):
        L = op.fftLength
        spec = getattr(op, "_torch_spec", None)
        if spec is None or spec.dtype != X.dtype:
            t = _embedded_generator(op)
            spec = torch.from_numpy((np.fft.fft(t) / L).astype(D.np_dtype(X.dtype))).to(X.device)
            op._torch_spec = spec
        if adjoint:
            spec = spec.conj()
        Y = torch.fft.ifft(torch.fft.fft(X, n=L, dim=1) * spec, dim=1, norm="forward")
        outs.append(Y[:, :(op.numCols if adjoint else op.numRows)])
    return outs


def _embedded_generator(op):
    from .circulant import Circulant
    L = op.fftLength
    if isinstance(op, Circulant):
        c = op._gen128
        n = c.size
        if L == n:
            return c
        return np.concatenate([c, np.zeros(L - 2 * n + 1), c[1:]])
    t = op._gen128
    m = op.numRows
    out = np.zeros(L, dtype=np.complex128)
    out[:m] = t[:m]
    if t.size > m:
        out[L - (t.size - m):] = t[m:]
    return out

"""Sparse-recovery solvers ISTA, FISTA, STELA and OMP with device-resident
iterations.

Mirrors fastmat.algorithms (algorithms/Algorithm.pyx, ISTA.py, FISTA.py,
STELA.py):
same class names, parameters (numLambda, numMaxSteps), callbacks (cbStep,
cbTrace, cbResult), ``process`` / ``_process`` and ``softThreshold``.  The
reference runs every step as numpy arithmetic around the operator's forward
and backward (ISTA.py:152-168, FISTA.py:154-178).  Here the iterate, the
measurements and all intermediates stay on the device: each step is the
operator's forward apply, the residual (fm_axpby), the adjoint apply and one
fused kernel (fm_shrink_step) that forms the gradient step, the soft
threshold and, for FISTA, the momentum update.  Without per-step callbacks
the first step runs eagerly (plans, workspaces) and one step is captured in a
HIP graph that is replayed for the remaining steps; FISTA's momentum per step
comes from a device table indexed by a device step counter the graph
advances (fm_counter_advance).  STELA's step (per-column exact line search)
is four fused kernels around the forward and adjoint applies, captured the
same way.

Results: numpy in -> numpy out, device tensor in -> device tensor out; the
result is 2-D (N, k) also for 1-D input, like the reference.
"""
import copy
import ctypes

import numpy as np
import torch

from . import _device as D
from . import _lib
from .matrix import Matrix


class Algorithm:
    """Algorithm.pyx:28-176: parameters, callbacks and the trace."""

    def __init__(self):
        if type(self) is Algorithm:
            raise NotImplementedError("Algorithm baseclass cannot be instantiated.")

    @property
    def cbTrace(self):
        return getattr(self, "_cbTrace", None)

    @cbTrace.setter
    def cbTrace(self, value):
        self._cbTrace = value

    @property
    def cbResult(self):
        return getattr(self, "_cbResult", None)

    @cbResult.setter
    def cbResult(self, value):
        self._cbResult = value

    @property
    def trace(self):
        if getattr(self, "_trace", None) is None:
            self._trace = []
        return self._trace

    @trace.setter
    def trace(self, value):
        if not isinstance(value, list):
            raise TypeError("Algorithm trace must be a list")
        self._trace = value

    def updateParameters(self, **kwargs):
        """Algorithm.pyx:96-121: unknown parameters raise AttributeError."""
        if getattr(self, "_attributes", None) is None:
            self._attributes = kwargs.copy()
        for key, value in kwargs.items():
            if not hasattr(self, key) and key not in self._attributes:
                raise AttributeError("Attribute '%s' not defined in %s" % (
                    key, self.__class__.__name__))
            setattr(self, key, value)

    def process(self, arrB, **kwargs):
        """Algorithm.pyx:123-140."""
        self.updateParameters(**kwargs)
        arrResult = self._process(arrB)
        self.handleCallback(self.cbResult)
        return arrResult

    def _process(self, arrB):
        raise NotImplementedError("Algorithm is not implemented yet.")

    def snapshot(self):
        """Algorithm.pyx:151-163: append a copy of the state to the trace."""
        trace, self._trace = getattr(self, "_trace", None), []
        if trace is None:
            trace = []
        trace.append(copy.copy(self))
        self._trace = trace

    def handleCallback(self, callback):
        if callback is not None:
            return callback(self)
        return None

    @property
    def nbytes(self):
        """Bytes of the arrays and device buffers the instance holds, its
        matrix included (Algorithm.pyx's getMemoryFootprint)."""
        from .matrix import _footprint
        seen = set()
        return sum(_footprint(v, seen) for v in vars(self).values()
                   if not callable(v))


class _Shrinkage(Algorithm):
    """Common device loop of ISTA (momentum 0) and FISTA."""

    _name = "ISTA"
    _momentum = False

    def __init__(self, fmatA, **kwargs):
        if not isinstance(fmatA, Matrix):
            raise TypeError("fmatA must be a fastmat matrix")
        self.fmatA = fmatA
        self.numLambda = 0.1
        self.numMaxSteps = 100
        self.cbStep = None
        self.updateParameters(**kwargs)

    def softThreshold(self, arrX, numAlpha):
        """ISTA.py:113-123 (host numpy, as in the reference)."""
        self.arrM = np.maximum(np.abs(arrX) - numAlpha, 0)
        return np.multiply((self.arrM / (self.arrM + numAlpha)), arrX)

    # ------------------------------------------------------------ helpers
    def _work_dtype(self, bdt):
        """The iterate's dtype: promote(float32, b) widened by whatever the
        forward / residual / adjoint products return (the reference's numpy
        promotion of ``x - L * A^H (A x - b)``), found by applying the
        operator to one zero column until the dtype is stable."""
        A = self.fmatA
        w = np.promote_types(np.float32, bdt)
        for _ in range(4):
            f = A._apply(D.empty(1, A.numCols, w).zero_(), False)
            r = np.promote_types(D.np_dtype(f.dtype), bdt)
            g = A._apply(D.empty(1, A.numRows, r).zero_(), True)
            w2 = np.promote_types(w, D.np_dtype(g.dtype))
            if w2 == w:
                break
            w = w2
        if w not in (np.dtype(np.float32), np.dtype(np.float64), np.dtype(np.complex64),
                     np.dtype(np.complex128)):
            raise TypeError("%s needs a floating-point problem, got %s" % (self._name, w))
        return w

    def _step(self, B, X, Y, Step, L, alpha, betas, counter, wdt):
        """One iteration on the device: Y is the point the gradient is taken
        at (X itself for ISTA)."""
        A = self.fmatA
        k = X.shape[0]
        F = A._apply(Y, False)
        if D.np_dtype(F.dtype) != wdt:
            F = F.to(D.torch_dtype(wdt))
        # F <- F - B  (alpha x + beta y with x = B, y = F)
        _lib.call("fm_axpby", _lib.NP_TO_FM[np.dtype(wdt)], A.numRows, k, _lib.dvec(-1.0),
                  _lib.ptr(B), A.numRows, 0, _lib.dvec(1.0), _lib.ptr(F), A.numRows, D.stream())
        G = A._apply(F, True)
        if D.np_dtype(G.dtype) != wdt:
            G = G.to(D.torch_dtype(wdt))
        G = G.contiguous()
        _lib.call("fm_shrink_step", _lib.NP_TO_FM[np.dtype(wdt)], X.numel(), _lib.ptr(Y),
                  _lib.ptr(G), float(L), float(alpha), _lib.ptr(Step), _lib.ptr(X),
                  _lib.ptr(betas) if betas is not None else None,
                  _lib.ptr(counter) if counter is not None else None, 0.0,
                  _lib.ptr(Y) if self._momentum else None, D.stream())
        if counter is not None:
            _lib.call("fm_counter_advance", _lib.ptr(counter), D.stream())

    # ------------------------------------------------------------ solver
    def _process(self, arrB):
        A = self.fmatA
        on_device = torch.is_tensor(arrB)
        b = arrB
        if b.ndim > 2 or b.ndim < 1:
            raise ValueError("Only n x m arrays are supported for %s" % self._name)
        if self.numMaxSteps <= 0:
            raise ValueError("%s would like to do at least one step for you" % self._name)
        if b.ndim == 1:
            b = b.reshape((-1, 1))
        self.arrB = b
        self.numL = 1.0 / (A.largestSingularValue ** 2)
        alpha = self.numL * self.numLambda * 0.5
        bdt = D.np_dtype(b.dtype)
        wdt = self._work_dtype(bdt)
        Bd, _ = D.to_block(b, A.numRows)
        Bd = Bd.to(D.torch_dtype(wdt)).contiguous()
        k = Bd.shape[0]
        X = D.empty(k, A.numCols, wdt).zero_()
        Y = X.clone() if self._momentum else X
        Step = torch.zeros_like(X)
        betas = counter = None
        if self._momentum:
            t, bl = 1.0, []
            for _ in range(int(self.numMaxSteps)):
                t_old = t
                t = (1 + np.sqrt(1 + 4 * t ** 2)) / 2
                bl.append((t_old - 1) / t)
            betas = torch.tensor(bl, dtype=torch.float64, device=X.device)
            counter = torch.zeros(1, dtype=torch.int64, device=X.device)
        callbacks = self.cbStep is not None or self.cbTrace is not None
        steps = int(self.numMaxSteps)
        self.numStep = 0
        self._step(Bd, X, Y, Step, self.numL, alpha, betas, counter, wdt)
        if callbacks:
            self._expose(X, Step, on_device)
            self.handleCallback(self.cbStep)
            self.handleCallback(self.cbTrace)
            for self.numStep in range(1, steps):
                self._step(Bd, X, Y, Step, self.numL, alpha, betas, counter, wdt)
                self._expose(X, Step, on_device)
                self.handleCallback(self.cbStep)
                self.handleCallback(self.cbTrace)
        elif steps > 1:
            graph = torch.cuda.CUDAGraph()
            cur = torch.cuda.current_stream()
            side = torch.cuda.Stream()
            side.wait_stream(cur)
            with torch.cuda.stream(side):
                with torch.cuda.graph(graph, stream=side):
                    self._step(Bd, X, Y, Step, self.numL, alpha, betas, counter, wdt)
            cur.wait_stream(side)
            for self.numStep in range(1, steps):
                graph.replay()
            del graph
        res = torch.where(X != 0, Step, X)
        self._expose(X, Step, on_device)
        if on_device:
            self.arrResult = res.t()
        else:
            torch.cuda.synchronize()
            self.arrResult = np.ascontiguousarray(res.cpu().numpy().T)
        return self.arrResult

    def _expose(self, X, Step, on_device):
        """The reference's per-step attributes (arrX, arrStep), for callbacks
        and after the run: device views for device input, numpy otherwise."""
        if on_device:
            self.arrX, self.arrStep = X.t(), Step.t()
        else:
            self.arrX = np.ascontiguousarray(X.cpu().numpy().T)
            self.arrStep = np.ascontiguousarray(Step.cpu().numpy().T)


class ISTA(_Shrinkage):
    """Iterative soft thresholding (ISTA.py:24-168):
    min ||A x - b||_2^2 + lambda ||x||_1, step 1 / sigma_max(A)^2."""
    _name = "ISTA"
    _momentum = False


class FISTA(_Shrinkage):
    """Fast iterative soft thresholding (FISTA.py:24-178): ISTA with the
    momentum y = x + (t_k - 1) / t_{k+1} (x - x_old)."""
    _name = "FISTA"
    _momentum = True


class STELA(Algorithm):
    """Soft thresholding with simplified exact line search (STELA.py:24-267).

    Per column of b: bx = soft(d x - z, lambda) / d with d = 1 / colNorms^2
    and z = A^H (A x - b), then the exact line-search step gamma in [0, 1]
    along bx - x; columns whose stopping value drops to numMaxError freeze.
    Works in promote(float64, b) like the reference.  Each step is four fused
    kernels (fm_stela_prox, fm_stela_line, fm_stela_update, fm_stela_select)
    around one forward and one adjoint apply; without callbacks the first
    step runs eagerly and one captured step is replayed for the rest (once
    every column is frozen the replays change nothing, which is the
    reference's early return).  With callbacks each step runs eagerly and the
    loop returns as soon as no column is active, as the reference does."""

    _name = "STELA"

    def __init__(self, fmatA, **kwargs):
        if not isinstance(fmatA, Matrix):
            raise TypeError("fmatA must be a fastmat matrix")
        self.fmatA = fmatA
        self.numLambda = 0.1
        self.numMaxSteps = 100
        self.numMaxError = 1e-6
        self.cbStep = None
        self.updateParameters(**kwargs)

    def softThreshold(self, arrX, numAlpha):
        """STELA.py:117-126 (host numpy, as in the reference)."""
        arrM = np.maximum(np.abs(arrX) - numAlpha, 0)
        return np.multiply((arrM / (arrM + numAlpha)), arrX)

    def _step(self, s):
        A = self.fmatA
        n, m, k = A.numCols, A.numRows, s["k"]
        xt, zt = s["xt"], s["zt"]
        lam = float(self.numLambda)
        _lib.call("fm_stela_prox", xt, zt, n, k, _lib.ptr(s["X"]), _lib.ptr(s["Z"]),
                  _lib.ptr(s["D"]), lam, _lib.ptr(s["Bx"]), _lib.ptr(s["Dlt"]),
                  _lib.ptr(s["part"]), s["parts"], D.stream())
        F = A._apply(s["Dlt"], False)
        if F.is_complex() and not s["X"].is_complex():
            F = F.real  # the reference assigns into its real array
        F = F.to(s["X"].dtype).contiguous()
        _lib.call("fm_stela_line", xt, m, k, _lib.ptr(s["Res"]), _lib.ptr(F),
                  _lib.ptr(s["part"]), s["parts"], lam, float(self.numMaxError),
                  _lib.ptr(s["gamma"]), _lib.ptr(s["active"]), D.stream())
        _lib.call("fm_stela_update", xt, n, m, k, _lib.ptr(s["Bx"]), _lib.ptr(s["X"]),
                  _lib.ptr(F), _lib.ptr(s["Res"]), _lib.ptr(s["gamma"]), _lib.ptr(s["active"]),
                  D.stream())
        Zn = A._apply(s["Res"], True)
        if Zn.dtype != s["Z"].dtype:
            Zn = Zn.to(s["Z"].dtype)
        Zn = Zn.contiguous()
        _lib.call("fm_stela_select", zt, n, k, _lib.ptr(s["Z"]), _lib.ptr(Zn),
                  _lib.ptr(s["active"]), D.stream())

    def _process(self, arrB):
        A = self.fmatA
        on_device = torch.is_tensor(arrB)
        if arrB.ndim > 2 or arrB.ndim < 1:
            raise ValueError("Only n x m arrays are supported for STELA")
        b = arrB.reshape((-1, 1)) if arrB.ndim == 1 else arrB
        self.arrB = b
        if self.numMaxSteps <= 0:
            raise ValueError("STELA would like to do at least one step for you")
        wdt = np.promote_types(np.float64, D.np_dtype(b.dtype))
        if wdt not in (np.dtype(np.float64), np.dtype(np.complex128)):
            raise TypeError("STELA needs a floating-point problem, got %s" % wdt)
        n, m = A.numCols, A.numRows
        Bd, _ = D.to_block(b, m)
        k = Bd.shape[0]
        if n == 0 or m == 0 or k == 0:
            self.arrX = np.zeros((n, k), dtype=wdt)
            return self.arrX if not on_device else torch.zeros(k, n, dtype=D.torch_dtype(wdt),
                                                                 device=D.device()).t()
        tw = D.torch_dtype(wdt)
        Res = (-Bd.to(tw)).contiguous()
        Z = A._apply(Res, True)
        zdt = D.np_dtype(Z.dtype)
        if zdt not in (np.dtype(np.float64), np.dtype(np.complex128)):
            zdt = np.promote_types(zdt, np.float64)
            Z = Z.to(D.torch_dtype(zdt))
        X = D.empty(k, n, wdt).zero_()
        norms = np.asarray(A._getColNorms(), dtype=np.float64)
        parts = int(min(256, max(1, -(-max(n, m) // 4096))))
        s = dict(k=k, xt=_lib.NP_TO_FM[np.dtype(wdt)], zt=_lib.NP_TO_FM[np.dtype(zdt)],
                 X=X, Z=Z.contiguous(), Res=Res, Bx=torch.zeros_like(X),
                 Dlt=torch.zeros_like(X), parts=parts,
                 D=torch.tensor(1.0 / norms ** 2, dtype=torch.float64, device=X.device),
                 part=torch.zeros(4 * k * parts, dtype=torch.float64, device=X.device),
                 gamma=torch.zeros(k, dtype=torch.float64, device=X.device),
                 active=torch.ones(k, dtype=torch.int32, device=X.device))
        callbacks = self.cbStep is not None or self.cbTrace is not None
        steps = int(self.numMaxSteps)
        self.numStep = 0
        if callbacks:
            for self.numStep in range(steps):
                self._step(s)
                active = s["active"].cpu().numpy() != 0
                self.arrActive = active
                if not active.any():
                    break
                self._expose(s, on_device)
                self.handleCallback(self.cbStep)
                self.handleCallback(self.cbTrace)
        else:
            self._step(s)
            if steps > 1:
                graph = torch.cuda.CUDAGraph()
                cur = torch.cuda.current_stream()
                side = torch.cuda.Stream()
                side.wait_stream(cur)
                with torch.cuda.stream(side):
                    with torch.cuda.graph(graph, stream=side):
                        self._step(s)
                cur.wait_stream(side)
                for self.numStep in range(1, steps):
                    graph.replay()
                del graph
            self.arrActive = s["active"].cpu().numpy() != 0
        self._expose(s, on_device)
        return self.arrX

    def _expose(self, s, on_device):
        if on_device:
            self.arrX = s["X"].t()
            self.arrGamma = s["gamma"]
        else:
            self.arrX = np.ascontiguousarray(s["X"].cpu().numpy().T)
            self.arrGamma = s["gamma"].cpu().numpy()


class OMP(Algorithm):
    """Orthogonal matching pursuit (OMP.pyx:24-240).

    Each step correlates the residual with the normalised columns through the
    operator's adjoint, picks the largest correlation per problem, forms the
    new column with one forward apply of a one-hot block and updates the
    pseudo-inverse of the selected columns by the reference's rank-1 rule;
    the selection, the update and the result scatter are device kernels
    (fm_omp_step).  Works in promote(A.dtype, b.dtype, float64) like the
    reference; numpy in -> numpy (M, L) out, device tensor in -> device."""

    _name = "OMP"

    def __init__(self, fmatA, **kwargs):
        if not isinstance(fmatA, Matrix):
            raise TypeError("fmatA must be a fastmat matrix")
        self.fmatA = fmatA
        self.numMaxSteps = 0
        self.cbStep = None
        self.updateParameters(**kwargs)

    def _process(self, arrB):
        A = self.fmatA
        on_device = torch.is_tensor(arrB)
        if arrB.ndim > 2 or arrB.ndim < 1:
            raise ValueError("Only n x m arrays are supported for OMP")
        b = arrB.reshape((-1, 1)) if arrB.ndim == 1 else arrB
        self.arrB = b
        if self.numMaxSteps <= 0:
            raise ValueError("OMP would like to do at least one step for you")
        steps = int(self.numMaxSteps)
        self.numN, self.numM = A.numRows, A.numCols
        n, m = self.numN, self.numM
        norms = np.asarray(A.colNorms, dtype=np.float64)
        if np.any(norms == 0):
            raise ValueError("Normalization: Matrix has zero-norm column.")
        self.returnType = np.promote_types(
            np.promote_types(np.promote_types(A.dtype, np.float64), D.np_dtype(b.dtype)),
            np.float64)
        vdt = self.returnType
        if vdt not in (np.dtype(np.float64), np.dtype(np.complex128)):
            raise TypeError("OMP needs a numeric problem, got %s" % vdt)
        tv = D.torch_dtype(vdt)
        Bd, _ = D.to_block(b, n)
        Bd = Bd.to(tv).contiguous()
        L = Bd.shape[0]
        dev = Bd.device
        R = Bd.clone()
        P = torch.zeros(L, steps, n, dtype=tv, device=dev)
        Ac = torch.zeros(L, steps, n, dtype=tv, device=dev)
        Xt = torch.zeros(L, steps, dtype=tv, device=dev)
        E = torch.empty(L, m, dtype=tv, device=dev)
        idx = torch.zeros(L, dtype=torch.int64, device=dev)
        sup = torch.zeros(L, steps, dtype=torch.int64, device=dev)
        inv = torch.tensor(1.0 / norms, dtype=torch.float64, device=dev)
        fm_t = _lib.NP_TO_FM[np.dtype(vdt)]
        need = ctypes.c_int64()
        _lib.call("fm_omp_workspace_size", fm_t, L, n, steps, ctypes.byref(need))
        work = D.workspace(need.value)
        Xf = torch.empty(L, m, dtype=tv, device=dev)
        st = D.stream()
        for self.numStep in range(steps):
            ii = self.numStep
            G = A._apply(R, True)
            G = (G if G.dtype == tv else G.to(tv)).contiguous()
            _lib.call("fm_omp_step", fm_t, 0, L, m, n, steps, ii, _lib.ptr(G), _lib.ptr(inv),
                      _lib.ptr(idx), _lib.ptr(sup), _lib.ptr(E), None, None, None, None, None,
                      _lib.ptr(work), need.value, None, None, st)
            C = A._apply(E, False)
            C = (C if C.dtype == tv else C.to(tv)).contiguous()
            _lib.call("fm_omp_step", fm_t, 1, L, m, n, steps, ii, None, None, None, None, None,
                      _lib.ptr(C), _lib.ptr(Bd), _lib.ptr(P), _lib.ptr(Ac), _lib.ptr(R),
                      _lib.ptr(work), need.value, _lib.ptr(Xt), None, st)
            if self.cbStep is not None or self.cbTrace is not None:
                self.arrSupport = sup.cpu().numpy().T
                self.handleCallback(self.cbStep)
                self.handleCallback(self.cbTrace)
        _lib.call("fm_omp_step", fm_t, 2, L, m, n, steps, 0, None, None, None, _lib.ptr(sup),
                  None, None, None, None, None, None, _lib.ptr(work), need.value, _lib.ptr(Xt),
                  _lib.ptr(Xf), st)
        if on_device:
            self.arrSupport = sup.t()
            self.arrX = Xf.t()
        else:
            self.arrSupport = sup.cpu().numpy().T
            self.arrX = np.ascontiguousarray(Xf.cpu().numpy().T)
        return self.arrX

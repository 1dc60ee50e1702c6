"""Batched form: many alignments of uneven L and N optimised together.

The reference optimises one alignment per ``ccmpred`` process
(ccmpred.c:537-648); here a :class:`Batch` holds one BF16 evaluator per
alignment and runs every step of the evaluation and of the conjugate
gradient as one grouped launch over all of them, with the line search on
the device and CG rounds replayed from a captured hipGraph.  An alignment
follows the trajectory :meth:`ccmpred_amd.pll.PLL.conjugrad` takes on the
same BF16 instance.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .pll import PRECISIONS, nvar_padded

STATUS = {0: "success", 1: "already minimized", 2: "maximum iterations", -1: "line search failed",
          -2: "device error"}


class Batch:
    """Alignments ``msas`` (uint8 N x L index matrices), each with the
    reference's defaults (theta, lambda_single, lambda_pair = factor (L-1))."""

    def __init__(self, msas, theta=0.8, lambda_single=0.01, lambda_pair_factor=0.2,
                 precision="bf16", device=0, sparse=True):
        if precision not in PRECISIONS or PRECISIONS[precision][0] != 1:
            raise ValueError("the batched form runs a BF16 precision form (bf16, bf16f, bf16x2)")
        self.msas = [np.ascontiguousarray(m, dtype=np.uint8) for m in msas]
        n = len(self.msas)
        if n == 0:
            raise ValueError("empty batch")
        self.n = n
        self.precision = precision
        self.weights = [np.zeros(m.shape[0], dtype=np.float32) for m in self.msas]
        self.ncol = [int(m.shape[1]) for m in self.msas]
        self.nrow = [int(m.shape[0]) for m in self.msas]
        self.nvar = [nvar_padded(L) for L in self.ncol]
        self.lambda_single = float(np.float32(lambda_single))
        self.lambda_pair = [float(np.float32(lambda_pair_factor * (L - 1))) for L in self.ncol]
        self._uds = (_lib.UserData * n)()
        for i, m in enumerate(self.msas):
            ud = self._uds[i]
            ud.msa = m.ctypes.data_as(C.POINTER(C.c_ubyte))
            ud.weights = self.weights[i].ctypes.data_as(C.POINTER(C.c_float))
            ud.ncol = self.ncol[i]
            ud.nrow = self.nrow[i]
            ud.nsingle = self.ncol[i] * 20
            ud.nvar = ud.nsingle + self.ncol[i] ** 2 * 21 * 21
            ud.lambda_single = self.lambda_single
            ud.lambda_pair = self.lambda_pair[i]
            ud.reweighting_threshold = float(theta)
        cfg = _lib.HipConfig()
        _lib.lib().ccm_hip_default_config(C.byref(cfg))
        cfg.device = int(device)
        cfg.precision, cfg.bf16_passes, cfg.bf16_fwd_passes = PRECISIONS[precision]
        cfg.verbose = 0
        cfg.sparse_fwd = 1 if sparse else 0
        h = C.c_void_p()
        if _lib.lib().ccm_batch_create(C.addressof(self._uds), n, C.byref(cfg), C.byref(h)) != 0:
            raise RuntimeError("ccm_batch_create failed")
        self._h = h

    def _p(self, p):
        if not 0 <= p < self.n:
            raise IndexError(p)
        return int(p)

    def set_x(self, p, x):
        """Host float32 x (nvar_padded entries) of alignment p."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.size != self.nvar[self._p(p)]:
            raise ValueError("x must have nvar_padded entries")
        if _lib.lib().ccm_batch_write_x(self._h, p, x.ctypes.data) != 0:
            raise RuntimeError("ccm_batch_write_x failed")

    def get_x(self, p):
        out = np.empty(self.nvar[self._p(p)], dtype=np.float32)
        if _lib.lib().ccm_batch_read(self._h, p, out.ctypes.data, None) != 0:
            raise RuntimeError("ccm_batch_read failed")
        return out

    def contact_map(self, p, apc=True, normalize=False):
        """The L x L contact map of alignment p's current x, computed on the
        device (util.c:16-139; equal bit for bit to pll.contact_map on get_x(p))."""
        L = self.ncol[self._p(p)]
        out = np.empty((L, L), dtype=np.float32)
        if _lib.lib().ccm_batch_contact_map(self._h, p, int(apc), int(normalize), out.ctypes.data) != 0:
            raise RuntimeError("ccm_batch_contact_map failed")
        return out

    def get_g(self, p):
        out = np.empty(self.nvar[self._p(p)], dtype=np.float32)
        if _lib.lib().ccm_batch_read(self._h, p, None, out.ctypes.data) != 0:
            raise RuntimeError("ccm_batch_read failed")
        return out

    def evaluate(self):
        """One grouped evaluation at every alignment's x: FP64 objectives
        (gradients stay on the device, see :meth:`get_g`)."""
        f = np.zeros(self.n, dtype=np.float64)
        if _lib.lib().ccm_batch_evaluate(self._h, f.ctypes.data) != 0:
            raise RuntimeError("ccm_batch_evaluate failed")
        return f

    def variants(self):
        fwd = np.zeros(self.n, dtype=np.int32)
        grad = np.zeros(self.n, dtype=np.int32)
        _lib.lib().ccm_batch_get_variants(self._h, fwd.ctypes.data, grad.ctypes.data)
        return fwd, grad

    def set_variants(self, fwd=None, grad=None):
        f = None if fwd is None else np.ascontiguousarray(fwd, dtype=np.int32)
        g = None if grad is None else np.ascontiguousarray(grad, dtype=np.int32)
        rc = _lib.lib().ccm_batch_set_variants(self._h, None if f is None else f.ctypes.data,
                                               None if g is None else g.ctypes.data)
        if rc != 0:
            raise RuntimeError("ccm_batch_set_variants failed")

    def autotune(self, reps=3):
        """Pick each alignment's GEMM tiles by timing the grouped launches;
        returns the tuned forward and gradient milliseconds."""
        ms = np.zeros(2, dtype=np.float64)
        if _lib.lib().ccm_batch_autotune(self._h, int(reps), ms.ctypes.data) != 0:
            raise RuntimeError("ccm_batch_autotune failed")
        return float(ms[0]), float(ms[1])

    def solve(self, max_iterations=50, epsilon=0.0, k=5, max_linesearch=5, alpha_mul=0.5, ftol=1e-4,
              wolfe=0.2, rounds_per_graph=4):
        """Device CG on every alignment from its x; returns the rounds run."""
        p = _lib.CGParam()
        _lib.lib().ccm_conjugrad_init(C.byref(p))
        p.max_iterations, p.epsilon, p.k = int(max_iterations), float(epsilon), int(k)
        p.max_linesearch, p.alpha_mul, p.ftol, p.wolfe = int(max_linesearch), alpha_mul, ftol, wolfe
        self._max_it = int(max_iterations)
        rounds = C.c_int()
        if _lib.lib().ccm_batch_solve(self._h, C.byref(p), int(rounds_per_graph), C.byref(rounds)) != 0:
            raise RuntimeError("ccm_batch_solve failed")
        return rounds.value

    def result(self, p):
        """status, f, iterations, evaluations and the per-iteration log
        (columns f, step, line-search trials, <g,g>) of alignment p."""
        st, it, ne = C.c_int(), C.c_int(), C.c_int()
        fx = C.c_double()
        log = np.zeros((max(1, getattr(self, "_max_it", 1)), 4), dtype=np.float64)
        if _lib.lib().ccm_batch_result(self._h, self._p(p), C.byref(st), C.byref(fx), C.byref(it),
                                       C.byref(ne), log.ctypes.data) != 0:
            raise RuntimeError("ccm_batch_result failed")
        return dict(ret=st.value, fx=fx.value, iterations=it.value, evaluations=ne.value,
                    log=log[: it.value].copy())

    @property
    def device_bytes(self) -> int:
        return int(_lib.lib().ccm_batch_device_bytes(self._h))

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().ccm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

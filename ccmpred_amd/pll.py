"""Python front end of the HIP PLL backend.

Mirrors the reference's backend lifecycle (ccmpred.c:537-558 and
evaluate_cuda.c): :class:`PLL` builds a ``userdata`` instance, ``init``
computes the sequence weights on the device, :meth:`PLL.evaluate` is one call
of the ``evaluate(instance, x, g, nvar)`` callback on device tensors, and
:meth:`PLL.conjugrad` runs the device CG.  torch only provides device memory.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

# precision form -> (precision enum, gradient-operand passes, forward-operand passes)
PRECISIONS = {"fp32": (0, 1, 1), "bf16": (1, 1, 1), "bf16x2": (1, 2, 2), "bf16f": (1, 1, 2)}


def nsingle_padded(ncol: int) -> int:
    return int(_lib.lib().ccm_nsingle_padded(int(ncol)))


def nvar_padded(ncol: int) -> int:
    return int(_lib.lib().ccm_nvar_padded(int(ncol)))


def init_bias(msa: np.ndarray) -> np.ndarray:
    """Host init_bias (ccmpred.c:90-131) from the library."""
    msa = np.ascontiguousarray(msa, dtype=np.uint8)
    nrow, ncol = msa.shape
    x = np.zeros(nvar_padded(ncol), dtype=np.float32)
    _lib.lib().ccm_init_bias(x.ctypes.data, msa.ctypes.data, ncol, nrow)
    return x


def contact_map(x: np.ndarray, ncol: int, apc: bool = True, normalize: bool = False) -> np.ndarray:
    """sum_submatrices -> apc -> normalize (util.c:16-139) on a host x."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros((ncol, ncol), dtype=np.float32)
    L = _lib.lib()
    L.ccm_sum_submatrices(x.ctypes.data, out.ctypes.data, int(ncol))
    if apc:
        L.ccm_apc(out.ctypes.data, int(ncol))
    if normalize:
        L.ccm_normalize(out.ctypes.data, int(ncol))
    return out


def contact_map_device(x, ncol: int, apc: bool = True, normalize: bool = False):
    """contact_map on the device: x is a CUDA float32 tensor in the padded
    layout; returns the L x L map as a CUDA tensor (equal bit for bit to
    contact_map on the same values)."""
    import torch
    if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float32):
        raise TypeError("x must be a CUDA float32 tensor")
    if x.numel() != nvar_padded(ncol) or not x.is_contiguous():
        raise ValueError("x must be a contiguous tensor of nvar_padded(ncol) floats")
    L = _lib.lib()
    out = torch.empty((ncol, ncol), dtype=torch.float32, device=x.device)
    work = torch.empty(max(1, L.ccm_contact_map_work_len(int(ncol))), dtype=torch.float32, device=x.device)
    stream = torch.cuda.current_stream(x.device).cuda_stream
    rc = L.ccm_launch_contact_map(x.data_ptr(), int(ncol), int(apc), int(normalize), out.data_ptr(),
                                  work.data_ptr(), stream)
    if rc != 0:
        raise RuntimeError(f"ccm_launch_contact_map failed ({rc})")
    return out


def top_pairs(mat: np.ndarray, num: int, min_sep: int = 6):
    """Highest-scoring (i, j), j - i >= min_sep, ties by (i, j) ascending
    (top_couplings.py:28-39 selection rule)."""
    L = mat.shape[0]
    ii, jj = np.triu_indices(L, k=min_sep)
    vals = mat[ii, jj]
    order = np.lexsort((jj, ii, -vals))[:num]
    return [(int(ii[o]), int(jj[o])) for o in order]


class PLL:
    """One MSA on one GPU: weights, f+grad evaluation and the device CG."""

    def __init__(self, msa, theta=0.8, lambda_single=0.01, lambda_pair=None,
                 lambda_pair_factor=0.2, precision="fp32", device=0, verbose=False, chunk_rows=0,
                 gemm_variant=-1, sparse=True):
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {list(PRECISIONS)}")
        self.msa = np.ascontiguousarray(msa, dtype=np.uint8)
        self.nrow, self.ncol = self.msa.shape
        if lambda_pair is None:
            lambda_pair = lambda_pair_factor * (self.ncol - 1)
        self.lambda_single = float(np.float32(lambda_single))
        self.lambda_pair = float(np.float32(lambda_pair))
        self.precision = precision
        self.weights = np.zeros(self.nrow, dtype=np.float32)
        self.nvar = nvar_padded(self.ncol)
        self.nsingle_padded = nsingle_padded(self.ncol)
        ud = _lib.UserData()
        ud.msa = self.msa.ctypes.data_as(C.POINTER(C.c_ubyte))
        ud.weights = self.weights.ctypes.data_as(C.POINTER(C.c_float))
        ud.ncol = self.ncol
        ud.nrow = self.nrow
        ud.nsingle = self.ncol * 20
        ud.nvar = ud.nsingle + self.ncol * self.ncol * 21 * 21
        ud.lambda_single = self.lambda_single
        ud.lambda_pair = self.lambda_pair
        ud.reweighting_threshold = float(theta)
        self._ud = ud
        self._uptr = C.addressof(ud)
        cfg = _lib.HipConfig()
        _lib.lib().ccm_hip_default_config(C.byref(cfg))
        cfg.device = int(device)
        cfg.precision, cfg.bf16_passes, cfg.bf16_fwd_passes = PRECISIONS[precision]
        cfg.verbose = 1 if verbose else 0
        cfg.chunk_rows = int(chunk_rows)
        cfg.gemm_variant = int(gemm_variant)
        cfg.sparse_fwd = 1 if sparse else 0
        rc = _lib.lib().init_hip_ex(self._uptr, C.byref(cfg))
        if rc != 0:
            raise RuntimeError("init_hip_ex failed")
        self._open = True

    # -- instance pointer for C callbacks
    @property
    def instance(self):
        return C.cast(self._uptr, C.c_void_p)

    @property
    def stream_ptr(self) -> int:
        return int(_lib.lib().ccm_hip_stream(self._uptr) or 0)

    def evaluate(self, x, g=None):
        """f (FP64) and the tied gradient for device tensor x (nvar float32)."""
        import torch
        if not (x.is_cuda and x.dtype == torch.float32 and x.numel() == self.nvar and x.is_contiguous()):
            raise ValueError("x must be a contiguous CUDA float32 tensor of nvar_padded entries")
        if g is None:
            g = torch.empty_like(x)
        torch.cuda.current_stream(x.device).synchronize()
        f = C.c_double()
        rc = _lib.lib().ccm_hip_evaluate(self._uptr, C.c_void_p(x.data_ptr()),
                                         C.c_void_p(g.data_ptr()), self.nvar, C.byref(f))
        if rc != 0:
            raise RuntimeError("ccm_hip_evaluate failed")
        return f.value, g

    def evaluate_ptr(self, x_ptr: int, g_ptr: int) -> float:
        """Raw device-pointer evaluate (no stream synchronisation before the call)."""
        f = C.c_double()
        rc = _lib.lib().ccm_hip_evaluate(self._uptr, C.c_void_p(x_ptr), C.c_void_p(g_ptr),
                                         self.nvar, C.byref(f))
        if rc != 0:
            raise RuntimeError("ccm_hip_evaluate failed")
        return f.value

    def _check(self, x):
        import torch
        if not (x.is_cuda and x.dtype == torch.float32 and x.numel() == self.nvar and x.is_contiguous()):
            raise ValueError("x must be a contiguous CUDA float32 tensor of nvar_padded entries")
        torch.cuda.current_stream(x.device).synchronize()

    def objective(self, x) -> float:
        """f alone (FP64); per-sequence probabilities are not kept."""
        self._check(x)
        f = C.c_double()
        if _lib.lib().ccm_hip_objective(self._uptr, C.c_void_p(x.data_ptr()), self.nvar, C.byref(f)):
            raise RuntimeError("ccm_hip_objective failed")
        return f.value

    def forward(self, x) -> float:
        """f (FP64), keeping the probability operand on the device for :meth:`backward`."""
        self._check(x)
        f = C.c_double()
        if _lib.lib().ccm_hip_forward(self._uptr, C.c_void_p(x.data_ptr()), self.nvar, C.byref(f)):
            raise RuntimeError("ccm_hip_forward failed")
        return f.value

    def backward(self, x, g=None):
        """Tied gradient from the probabilities the last :meth:`forward` kept."""
        import torch
        self._check(x)
        if g is None:
            g = torch.empty_like(x)
        if _lib.lib().ccm_hip_backward(self._uptr, C.c_void_p(x.data_ptr()), C.c_void_p(g.data_ptr()),
                                       self.nvar):
            raise RuntimeError("ccm_hip_backward failed")
        return g

    @property
    def device_bytes(self) -> int:
        return int(_lib.lib().ccm_hip_device_bytes(self._uptr))

    def set_timing(self, on=True):
        _lib.lib().ccm_hip_set_timing(self._uptr, 1 if on else 0)

    def timing(self):
        lib = _lib.lib()
        names = (C.c_char_p * 5)()
        cnt = lib.ccm_hip_step_names(self._uptr, names, 5)
        out = (C.c_double * 5)()
        n = lib.ccm_hip_get_timing(self._uptr, out, 5)
        return n, {names[i].decode(): out[i] for i in range(cnt)}

    def conjugrad(self, x, max_iterations=50, epsilon=0.0, k=5, max_linesearch=5, alpha_mul=0.5,
                  ftol=1e-4, wolfe=0.2, line_search="device"):
        """Device CG from x (modified in place); returns status and counters.
        line_search "device": decisions on the device where the instance
        allows it (CCM_LINE_SEARCH_DEVICE), "host": after every evaluation
        on the host; both give the same results bit for bit."""
        import torch
        if line_search not in ("device", "host"):
            raise ValueError("line_search must be 'device' or 'host'")
        p = _lib.CGParam()
        _lib.lib().ccm_conjugrad_init(C.byref(p))
        p.max_iterations, p.epsilon, p.k = int(max_iterations), float(epsilon), int(k)
        p.max_linesearch, p.alpha_mul, p.ftol, p.wolfe = int(max_linesearch), alpha_mul, ftol, wolfe
        p.line_search = 0 if line_search == "device" else 1
        torch.cuda.current_stream(x.device).synchronize()
        fx = C.c_double()
        it = C.c_int()
        ne = C.c_int()
        hist = (C.c_double * max(1, int(max_iterations)))()
        ret = _lib.lib().ccm_conjugrad_gpu_ex(self.nvar, C.c_void_p(x.data_ptr()), C.byref(fx),
                                              _lib.evaluate_fnptr(), None, self._uptr,
                                              C.byref(p), C.byref(it), C.byref(ne), hist)
        return dict(ret=ret, fx=fx.value, iterations=it.value, evaluations=ne.value,
                    fhist=np.array(hist[: it.value]))

    def close(self):
        if getattr(self, "_open", False):
            _lib.lib().destroy_hip(self._uptr)
            self._open = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

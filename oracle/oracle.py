"""ctypes wrapper of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg.
The product package (ccmpred_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

NA = 21
NAP = 32


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class CGParam(C.Structure):
    _fields_ = [("max_iterations", C.c_int), ("epsilon", C.c_double), ("k", C.c_int),
                ("max_linesearch", C.c_int), ("alpha_mul", C.c_double), ("ftol", C.c_double),
                ("wolfe", C.c_double), ("fres", C.c_double)]


def _declare(L):
    L.ora_identity_threshold.argtypes = [C.c_int, C.c_float]
    L.ora_identity_threshold.restype = C.c_int
    L.ora_identity_counts.argtypes = [_u8p, C.c_int, C.c_int, C.c_int, _i32p]
    L.ora_calculate_weights.argtypes = [_u8p, C.c_int, C.c_int, C.c_float, _f32p]
    L.ora_uniform_weights.argtypes = [C.c_int, _f32p]
    L.ora_init_bias.argtypes = [_u8p, C.c_int, C.c_int, _f32p]
    L.ora_evaluate_f32.argtypes = [_u8p, _f32p, C.c_int, C.c_int, C.c_float, C.c_float, _f32p, _f32p]
    L.ora_evaluate_f32.restype = C.c_float
    L.ora_evaluate_f64.argtypes = [_u8p, _f64p, C.c_int, C.c_int, C.c_double, C.c_double, _f64p, _f64p]
    L.ora_evaluate_f64.restype = C.c_double
    L.ora_evaluate_f64_par.argtypes = [_u8p, _f64p, C.c_int, C.c_int, C.c_double, C.c_double, _f64p,
                                       _f64p]
    L.ora_evaluate_f64_par.restype = C.c_double
    L.ora_forward_lse_f32.argtypes = [_u8p, C.c_int, C.c_int, _f32p, _f32p]
    L.ora_forward_lse_f64.argtypes = [_u8p, C.c_int, C.c_int, _f64p, _f64p]
    L.ora_omp_create.argtypes = [_u8p, _f32p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int]
    L.ora_omp_create.restype = C.c_void_p
    L.ora_evaluate_omp_f32.argtypes = [C.c_void_p, _f32p, _f32p]
    L.ora_evaluate_omp_f32.restype = C.c_float
    L.ora_omp_destroy.argtypes = [C.c_void_p]
    L.ora_omp_max_threads.restype = C.c_int
    L.ora_sum_submatrices.argtypes = [_f32p, C.c_int, _f32p]
    L.ora_apc.argtypes = [_f32p, C.c_int]
    L.ora_normalize.argtypes = [_f32p, C.c_int]
    L.ora_cg_pll_f32.argtypes = [_u8p, _f32p, C.c_int, C.c_int, C.c_float, C.c_float, _f32p,
                                 C.POINTER(CGParam), C.POINTER(C.c_double), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int), _f64p]
    L.ora_cg_pll_f32.restype = C.c_int
    L.ora_cg_pll_f64.argtypes = L.ora_cg_pll_f32.argtypes
    L.ora_cg_pll_f64.restype = C.c_int
    L.ora_cg_pll_f64par.argtypes = L.ora_cg_pll_f32.argtypes
    L.ora_cg_pll_f64par.restype = C.c_int


def nsingle_padded(ncol: int) -> int:
    ns = ncol * (NA - 1)
    return ns + NAP - ns % NAP


def nvar_padded(ncol: int) -> int:
    return nsingle_padded(ncol) + ncol * ncol * NA * NAP


def _msa(msa):
    return np.ascontiguousarray(msa, dtype=np.uint8)


def identity_threshold(ncol, theta):
    return lib().ora_identity_threshold(int(ncol), float(theta))


def identity_counts(msa, idthres):
    msa = _msa(msa)
    out = np.zeros(msa.shape[0], dtype=np.int32)
    lib().ora_identity_counts(msa, msa.shape[1], msa.shape[0], int(idthres), out)
    return out


def calculate_weights(msa, theta=0.8):
    msa = _msa(msa)
    w = np.zeros(msa.shape[0], dtype=np.float32)
    if np.float32(theta) == np.float32(1.0):
        lib().ora_uniform_weights(msa.shape[0], w)
    else:
        lib().ora_calculate_weights(msa, msa.shape[1], msa.shape[0], float(theta), w)
    return w


def init_bias(msa):
    msa = _msa(msa)
    x = np.zeros(nvar_padded(msa.shape[1]), dtype=np.float32)
    lib().ora_init_bias(msa, msa.shape[1], msa.shape[0], x)
    return x


def evaluate_f32(msa, w, x, lambda_single, lambda_pair):
    msa = _msa(msa)
    x = np.ascontiguousarray(x, dtype=np.float32)
    g = np.zeros_like(x)
    f = lib().ora_evaluate_f32(msa, np.ascontiguousarray(w, dtype=np.float32), msa.shape[1],
                               msa.shape[0], float(lambda_single), float(lambda_pair), x, g)
    return float(f), g


def evaluate_f64(msa, w, x, lambda_single, lambda_pair):
    msa = _msa(msa)
    x = np.ascontiguousarray(x, dtype=np.float64)
    g = np.zeros_like(x)
    f = lib().ora_evaluate_f64(msa, np.ascontiguousarray(w, dtype=np.float64), msa.shape[1],
                               msa.shape[0], float(lambda_single), float(lambda_pair), x, g)
    return float(f), g


def evaluate_f64_par(msa, w, x, lambda_single, lambda_pair):
    """FP64 evaluator parallel over sequence blocks (OpenMP): for large sizes."""
    msa = _msa(msa)
    x = np.ascontiguousarray(x, dtype=np.float64)
    g = np.zeros_like(x)
    f = lib().ora_evaluate_f64_par(msa, np.ascontiguousarray(w, dtype=np.float64), msa.shape[1],
                                   msa.shape[0], float(lambda_single), float(lambda_pair), x, g)
    return float(f), g


def forward_lse_f32(msa, x):
    msa = _msa(msa)
    out = np.zeros(msa.shape, dtype=np.float32)
    lib().ora_forward_lse_f32(msa, msa.shape[1], msa.shape[0],
                              np.ascontiguousarray(x, dtype=np.float32), out)
    return out


def forward_lse_f64(msa, x):
    msa = _msa(msa)
    out = np.zeros(msa.shape, dtype=np.float64)
    lib().ora_forward_lse_f64(msa, msa.shape[1], msa.shape[0],
                              np.ascontiguousarray(x, dtype=np.float64), out)
    return out


class OmpEvaluator:
    """The OpenMP CPU evaluator, as timed for the CPU baseline."""

    def __init__(self, msa, w, lambda_single, lambda_pair, nthreads=0):
        self.msa = _msa(msa)
        self.w = np.ascontiguousarray(w, dtype=np.float32)
        self.ncol = self.msa.shape[1]
        self._ctx = lib().ora_omp_create(self.msa, self.w, self.ncol, self.msa.shape[0],
                                         float(lambda_single), float(lambda_pair), int(nthreads))
        self.threads = lib().ora_omp_max_threads()

    def __call__(self, x, g=None):
        if g is None:
            g = np.zeros_like(x)
        return float(lib().ora_evaluate_omp_f32(self._ctx, x, g)), g

    def close(self):
        if self._ctx:
            lib().ora_omp_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        self.close()


def sum_submatrices(x, ncol):
    out = np.zeros((ncol, ncol), dtype=np.float32)
    lib().ora_sum_submatrices(np.ascontiguousarray(x, dtype=np.float32), int(ncol), out)
    return out


def apc(mat):
    m = np.array(mat, dtype=np.float32, order="C", copy=True)
    lib().ora_apc(m, m.shape[0])
    return m


def normalize(mat):
    m = np.array(mat, dtype=np.float32, order="C", copy=True)
    lib().ora_normalize(m, m.shape[0])
    return m


def cg_pll_f32(msa, w, x0, lambda_single, lambda_pair, max_iterations=50, epsilon=0.0, k=5,
               max_linesearch=5, alpha_mul=0.5, ftol=1e-4, wolfe=0.2, evaluator="f32", fres=1e-6):
    msa = _msa(msa)
    x = np.array(x0, dtype=np.float32, copy=True)
    p = CGParam(max_iterations, epsilon, k, max_linesearch, alpha_mul, ftol, wolfe, fres)
    fx = C.c_double()
    it = C.c_int()
    ne = C.c_int()
    hist = np.zeros(max_iterations + 1, dtype=np.float64)
    fn = {"f32": lib().ora_cg_pll_f32, "f64": lib().ora_cg_pll_f64,
          "f64par": lib().ora_cg_pll_f64par}[evaluator]
    ret = fn(msa, np.ascontiguousarray(w, dtype=np.float32), msa.shape[1],
                               msa.shape[0], float(lambda_single), float(lambda_pair), x,
                               C.byref(p), C.byref(fx), C.byref(it), C.byref(ne), hist)
    return dict(ret=ret, x=x, fx=fx.value, iterations=it.value, evaluations=ne.value,
                fhist=hist[: it.value])


_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def tied_fd_f64(msa, w, x, lambda_single, lambda_pair, idx_a, idx_b, h=1e-5):
    """FP64 central differences of f along e_a + e_b (e_a alone when a == b)."""
    L_ = lib()
    if not getattr(L_, "_fd_declared", False):
        L_.ora_tied_fd_f64.argtypes = [_u8p, _f64p, C.c_int, C.c_int, C.c_double, C.c_double, _f64p,
                                       _i64p, _i64p, C.c_int, C.c_double, _f64p]
        L_._fd_declared = True
    msa = _msa(msa)
    ia = np.ascontiguousarray(idx_a, dtype=np.int64)
    ib = np.ascontiguousarray(idx_b, dtype=np.int64)
    out = np.zeros(len(ia), dtype=np.float64)
    L_.ora_tied_fd_f64(msa, np.ascontiguousarray(w, dtype=np.float64), msa.shape[1], msa.shape[0],
                       float(lambda_single), float(lambda_pair), np.ascontiguousarray(x, dtype=np.float64),
                       ia, ib, len(ia), float(h), out)
    return out


def random_tied_coordinates(ncol, count, rng, singles_frac=0.2):
    """Random (idx_a, idx_b) pairs: non-self, non-padding couplings with their
    mirror, and single potentials (idx_b == idx_a)."""
    L = ncol
    nsp = nsingle_padded(L)
    ia, ib = [], []
    while len(ia) < count:
        if rng.random() < singles_frac:
            v = int(rng.integers(0, 20 * L))
            ia.append(v)
            ib.append(v)
            continue
        b, a = (int(t) for t in rng.integers(0, NA, 2))
        k, j = (int(t) for t in rng.integers(0, L, 2))
        if k == j:
            continue
        ia.append(nsp + ((b * L + k) * NAP + a) * L + j)
        ib.append(nsp + ((a * L + j) * NAP + b) * L + k)
    return np.array(ia, dtype=np.int64), np.array(ib, dtype=np.int64)

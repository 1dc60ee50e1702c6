"""Precision forms at L=250, N=5000 and at L=600, N=20000.

For each size: one evaluation of every precision form at the same x
(column log-frequencies + symmetric N(0, 0.01^2) couplings) against the FP64
oracle (the OpenMP FP64 restatement, checked against the serial one in
tests/test_oracle_pll.py).  The 50-iteration CG of every form at both sizes
is compared with the FP64-oracle CG in tests/test_gpu_cg_fp64.py.

Tolerances (gradient relative L2 against FP64): fp32 1e-4 (and f 1e-5);
bf16f and bf16 at L=250 1e-3; bf16x2 1e-4.  Single-pass bf16 at L=600,
N=20000 measures 1.3e-3 (its forward logits sum 600 BF16-rounded couplings;
the two-pass forward of bf16f brings the gradient to 3e-4), so it is held to
2e-3 there.
"""
import time

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as ora  # noqa: E402
from ccmpred_amd import pll as P  # noqa: E402
from ccmpred_amd.msa import synthetic_msa  # noqa: E402
from tests.pll_bruteforce import nsingle_padded, NA, NAP  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def start_x(msa, seed):
    L = msa.shape[1]
    x = ora.init_bias(msa).astype(np.float64)
    nsp = nsingle_padded(L)
    rng = np.random.default_rng(seed)
    W = rng.normal(0, 0.01, (NA, L, NA, L))
    W = W + W.transpose(2, 3, 0, 1)
    W[:, np.arange(L), :, np.arange(L)] = 0
    x[nsp:].reshape(NA, L, NAP, L)[:, :, :NA, :] += W
    return x.astype(np.float32)


def grad_tol(prec, L):
    return {"fp32": 1e-4, "bf16": 1e-3 if L <= 250 else 2e-3, "bf16f": 1e-3, "bf16x2": 1e-4}[prec]


@pytest.mark.parametrize("L,N,seed", [(250, 5000, 2), (600, 20000, 4)])
def test_precision_forms_against_fp64(L, N, seed):
    msa, _ = synthetic_msa(L, N, seed, 60)
    x = start_x(msa, seed + 50)
    say(f"precision report L={L} N={N}")
    with P.PLL(msa, precision="fp32") as p32:
        w = p32.weights.copy()
        lam1, lam2 = p32.lambda_single, p32.lambda_pair
    t0 = time.perf_counter()
    f64, g64 = ora.evaluate_f64_par(msa, w.astype(np.float64), x.astype(np.float64), lam1, lam2)
    say(f"  FP64 oracle f={f64:.10g} ({time.perf_counter() - t0:.1f} s)")
    nsp = nsingle_padded(L)
    xd = torch.from_numpy(x).cuda()
    for prec in ("fp32", "bf16", "bf16f", "bf16x2"):
        tol = grad_tol(prec, L)
        with P.PLL(msa, precision=prec) as p:
            assert np.array_equal(p.weights, w)
            f, g = p.evaluate(xd)
            g = g.cpu().numpy().astype(np.float64)
        ef = abs(f - f64) / abs(f64)
        eg = np.linalg.norm(g - g64) / np.linalg.norm(g64)
        eg1 = np.linalg.norm(g[:nsp] - g64[:nsp]) / np.linalg.norm(g64[:nsp])
        eg2 = np.linalg.norm(g[nsp:] - g64[nsp:]) / np.linalg.norm(g64[nsp:])
        say(f"  {prec:7s} f rel err {ef:.2e}  g rel L2 {eg:.2e} (singles {eg1:.2e}, pairs {eg2:.2e})")
        assert eg <= tol
        if prec == "fp32":
            assert ef <= 1e-5
        else:
            assert ef <= 1e-4

"""Pins the oracle's PLL evaluator (evaluate_cpu.c restatement).

Checks: brute-force FP64 on tiny alignments, x = 0 closed forms, tied-pair
central differences (numdif.c:29-45 style), FP32 vs FP64 agreement, the
gradient invariants, and the OpenMP restatement against the serial one.
"""
import numpy as np
import pytest

from oracle import oracle as ora
from ccmpred_amd.msa import synthetic_msa
from tests import pll_bruteforce as bf
from tests.pll_bruteforce import pll, random_symmetric_x, nsingle_padded, NA, NAP


def _small(L=5, N=9, seed=0):
    rng = np.random.default_rng(seed)
    msa = rng.integers(0, 21, size=(N, L)).astype(np.uint8)
    w = rng.uniform(0.2, 1.0, N).astype(np.float32)
    return rng, msa, w


@pytest.mark.parametrize("L,N,seed", [(3, 4, 0), (4, 8, 1), (6, 11, 2)])
def test_oracle_f64_matches_bruteforce(L, N, seed):
    rng, msa, w = _small(L, N, seed)
    x = random_symmetric_x(L, rng)
    f_ref, g_ref = pll(msa, w.astype(np.float64), x, 0.01, 0.2 * (L - 1))
    f, g = ora.evaluate_f64(msa, w, x, 0.01, 0.2 * (L - 1))
    assert f == pytest.approx(f_ref, rel=1e-12)
    np.testing.assert_allclose(g, g_ref, rtol=1e-10, atol=1e-12)


def test_oracle_f64_bruteforce_nonsymmetric_x():
    """Non-symmetric x, nonzero self-edges and padding: the reference formula
    still defines f and g (L2 over the padded block, self-edges get 2*lam2*x)."""
    rng, msa, w = _small(4, 7, 3)
    x = rng.normal(0, 0.1, ora.nvar_padded(4))
    x[20 * 4: nsingle_padded(4)] = 0.0   # singles padding is never read
    f_ref, g_ref = pll(msa, w, x, 0.03, 0.5)
    f, g = ora.evaluate_f64(msa, w, x, 0.03, 0.5)
    assert f == pytest.approx(f_ref, rel=1e-12)
    np.testing.assert_allclose(g, g_ref, rtol=1e-10, atol=1e-12)


def test_x_zero_closed_forms():
    msa, _ = synthetic_msa(12, 40, seed=7, n_pairs=2)
    N, L = msa.shape
    w = ora.calculate_weights(msa, 0.8).astype(np.float64)
    x = np.zeros(ora.nvar_padded(L))
    f, g = ora.evaluate_f64(msa, w, x, 0.01, 0.2 * (L - 1))
    W = w.sum()
    # all logits 0 -> P = 1/21 and lse = ln 21
    assert f == pytest.approx(W * L * np.log(21.0), rel=1e-13)
    onehot = np.eye(NA)[msa]                              # [n, k, b]
    C = np.einsum("n,nkb->kb", w, onehot)                 # C(k, b)
    H2 = np.einsum("n,nkb,nja->bkaj", w, onehot, onehot)  # H2(b,k,a,j)
    g1 = g[: 20 * L].reshape(20, L)
    np.testing.assert_allclose(g1, (W / 21.0 - C[:, :20]).T, atol=1e-12)
    G2 = g[nsingle_padded(L):].reshape(NA, L, NAP, L)
    expect = (C.T[:, :, None, None] + C.T[None, None, :, :]) / 21.0 - 2 * H2
    for k in range(L):
        expect[:, k, :, k] = 0.0
    np.testing.assert_allclose(G2[:, :, :NA, :], expect, atol=1e-12)
    assert np.all(G2[:, :, NA:, :] == 0.0)


def _tied_fd(msa, w, x, lam1, lam2, idx_pairs, h=1e-5):
    out = []
    for i_a, i_b in idx_pairs:
        xp = x.copy()
        xm = x.copy()
        xp[i_a] += h
        xm[i_a] -= h
        if i_b != i_a:
            xp[i_b] += h
            xm[i_b] -= h
        fp, _ = ora.evaluate_f64(msa, w, xp, lam1, lam2)
        fm, _ = ora.evaluate_f64(msa, w, xm, lam1, lam2)
        out.append((fp - fm) / (2 * h))
    return np.array(out)


def test_tied_central_differences():
    rng = np.random.default_rng(11)
    msa, _ = synthetic_msa(10, 30, seed=11, n_pairs=2)
    N, L = msa.shape
    w = ora.calculate_weights(msa, 0.8).astype(np.float64)
    x = random_symmetric_x(L, rng, scale=0.05)
    lam1, lam2 = 0.01, 0.2 * (L - 1)
    f, g = ora.evaluate_f64(msa, w, x, lam1, lam2)
    nsp = nsingle_padded(L)
    pairs = []
    while len(pairs) < 40:
        b, a = rng.integers(0, NA, 2)
        k, j = rng.integers(0, L, 2)
        if k == j:
            continue
        ia = nsp + ((b * L + k) * NAP + a) * L + j
        ib = nsp + ((a * L + j) * NAP + b) * L + k
        pairs.append((ia, ib))
    for _ in range(20):
        v = int(rng.integers(0, 20 * L))
        pairs.append((v, v))
    fd = _tied_fd(msa, w, x, lam1, lam2, pairs)
    an = np.array([g[i] for i, _ in pairs])
    np.testing.assert_allclose(an, fd, rtol=2e-6, atol=2e-6)


def test_f32_matches_f64():
    rng = np.random.default_rng(5)
    msa, _ = synthetic_msa(30, 200, seed=5, n_pairs=4)
    N, L = msa.shape
    w = ora.calculate_weights(msa, 0.8)
    x = ora.init_bias(msa).astype(np.float64)
    x += random_symmetric_x(L, rng, scale=0.01, single_scale=0.0)
    f64, g64 = ora.evaluate_f64(msa, w.astype(np.float64), x, 0.01, 0.2 * (L - 1))
    f32, g32 = ora.evaluate_f32(msa, w, x.astype(np.float32), 0.01, 0.2 * (L - 1))
    assert f32 == pytest.approx(f64, rel=1e-5)
    assert np.linalg.norm(g32 - g64) / np.linalg.norm(g64) < 1e-5


def test_gradient_invariants():
    """g symmetric, zero padding, and sum_a of the likelihood part of
    g2(b,k,a,j) equals the node-gradient sum_n w_n (P_n(b,k) - [x_nk=b])."""
    rng = np.random.default_rng(9)
    msa, _ = synthetic_msa(9, 25, seed=9, n_pairs=1)
    N, L = msa.shape
    w = ora.calculate_weights(msa, 0.8).astype(np.float64)
    x = random_symmetric_x(L, rng, scale=0.05)
    lam2 = 0.2 * (L - 1)
    f, g = ora.evaluate_f64(msa, w, x, 0.0, lam2)
    nsp = nsingle_padded(L)
    G2 = g[nsp:].reshape(NA, L, NAP, L)
    X2 = x[nsp:].reshape(NA, L, NAP, L)
    T = G2[:, :, :NA, :]
    np.testing.assert_allclose(T, T.transpose(2, 3, 0, 1), atol=1e-13)
    assert np.all(G2[:, :, NA:, :] == 0.0)
    lik = (G2 - 2 * lam2 * X2)[:, :, :NA, :]
    g1 = g[: 20 * L].reshape(20, L)
    for k in range(L):
        for j in range(L):
            if j == k:
                continue
            s = lik[:, k, :, j].sum(axis=1)       # sum over a -> per b
            np.testing.assert_allclose(s[:20], g1[:, k], atol=1e-12)


def test_omp_matches_serial():
    msa, _ = synthetic_msa(20, 64, seed=3, n_pairs=3)
    N, L = msa.shape
    w = ora.calculate_weights(msa, 0.8)
    rng = np.random.default_rng(3)
    x = (ora.init_bias(msa) + random_symmetric_x(L, rng, 0.02, 0.0)).astype(np.float32)
    f0, g0 = ora.evaluate_f32(msa, w, x, 0.01, 0.2 * (L - 1))
    ev = ora.OmpEvaluator(msa, w, 0.01, 0.2 * (L - 1), nthreads=4)
    f1, g1 = ev(x)
    ev.close()
    assert f1 == pytest.approx(f0, rel=1e-6)
    np.testing.assert_allclose(g1, g0, rtol=1e-5, atol=1e-6)


def test_fast_tied_fd_matches_full_reevaluation():
    """ora_tied_fd_f64 (incremental) against central differences by full
    re-evaluation, and against the oracle's tied gradient."""
    rng = np.random.default_rng(21)
    msa, _ = synthetic_msa(11, 40, seed=21, n_pairs=2)
    N, L = msa.shape
    w = ora.calculate_weights(msa, 0.8).astype(np.float64)
    x = random_symmetric_x(L, rng, scale=0.05)
    lam1, lam2 = 0.01, 0.2 * (L - 1)
    ia, ib = ora.random_tied_coordinates(L, 40, rng)
    # include a self-edge coupling and a padding entry: no likelihood term
    nsp = nsingle_padded(L)
    ia = np.append(ia, [nsp + ((3 * L + 2) * NAP + 5) * L + 2, nsp + ((1 * L + 0) * NAP + 25) * L + 4])
    ib = np.append(ib, [nsp + ((5 * L + 2) * NAP + 3) * L + 2, nsp + ((1 * L + 0) * NAP + 25) * L + 4])
    fast = ora.tied_fd_f64(msa, w, x, lam1, lam2, ia, ib)
    slow = _tied_fd(msa, w, x, lam1, lam2, list(zip(ia.tolist(), ib.tolist())))
    np.testing.assert_allclose(fast, slow, rtol=1e-6, atol=1e-6)
    f, g = ora.evaluate_f64(msa, w, x, lam1, lam2)
    np.testing.assert_allclose(fast[:40], g[ia[:40]], rtol=2e-6, atol=2e-6)


def test_parallel_f64_matches_serial():
    """The OpenMP FP64 evaluator (blocked over sequences, used for large-size
    error reports) agrees with the serial FP64 restatement to rounding."""
    msa, _ = synthetic_msa(23, 600, 31, 3)
    w = ora.calculate_weights(msa, 0.8).astype(np.float64)
    rng = np.random.default_rng(3)
    x = ora.init_bias(msa).astype(np.float64)
    L = msa.shape[1]
    nsp = nsingle_padded(L)
    W = rng.normal(0, 0.05, (NA, L, NA, L))
    W = W + W.transpose(2, 3, 0, 1)
    W[:, np.arange(L), :, np.arange(L)] = 0
    x[nsp:].reshape(NA, L, NAP, L)[:, :, :NA, :] += W
    f1, g1 = ora.evaluate_f64(msa, w, x, 0.01, 0.2 * (L - 1))
    f2, g2 = ora.evaluate_f64_par(msa, w, x, 0.01, 0.2 * (L - 1))
    assert abs(f1 - f2) <= 1e-12 * abs(f1)
    assert np.max(np.abs(g1 - g2)) <= 1e-12 * np.max(np.abs(g1))


@pytest.mark.parametrize("L,N,seed", [(1, 3, 31), (3, 7, 32), (5, 12, 33)])
def test_forward_lse_matches_bruteforce(L, N, seed):
    """ora_forward_lse_f32/_f64 (the per-(sequence, column) log-partition of
    evaluate_cpu.c:55-87) against the brute-force model's max-shifted
    log-sum-exp, including non-symmetric couplings and the gap state."""
    rng = np.random.default_rng(seed)
    msa = rng.integers(0, 21, (N, L), dtype=np.uint8)
    x = bf.random_symmetric_x(L, rng, scale=0.3, single_scale=0.7)
    X2 = x[nsingle_padded(L):].reshape(NA, L, NAP, L)
    X2[:, :, :NA, :] += rng.normal(0, 0.05, (NA, L, NA, L))  # not symmetric any more
    ref = bf.lse(msa, x)
    got64 = ora.forward_lse_f64(msa, x)
    got32 = ora.forward_lse_f32(msa, x.astype(np.float32))
    assert got64.shape == (N, L)
    np.testing.assert_allclose(got64, ref, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(got32, ref, rtol=2e-6, atol=2e-6)


def test_forward_lse_x_zero():
    """x = 0: every log-partition is ln 21."""
    msa = np.random.default_rng(5).integers(0, 21, (9, 6), dtype=np.uint8)
    x = np.zeros(ora.nvar_padded(6))
    np.testing.assert_allclose(ora.forward_lse_f64(msa, x), np.log(21.0), rtol=1e-15)
    np.testing.assert_allclose(ora.forward_lse_f32(msa, x.astype(np.float32)), np.log(21.0), rtol=1e-6)

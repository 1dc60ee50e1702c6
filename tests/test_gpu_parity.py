"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Tolerances: FP32 path f within 1e-5 relative and gradient within 1e-4
relative L2 of the FP64 oracle; BF16 path gradient within 1e-3 relative L2;
reweighting bit-exact; after 50 CG iterations the APC maps agree with
Spearman >= 0.999 and identical top-L/5 pairs.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as ora  # noqa: E402
from ccmpred_amd import _lib, pll as P  # noqa: E402
from ccmpred_amd.msa import read_psicov, synthetic_msa  # noqa: E402
from tests.pll_bruteforce import nsingle_padded, NA, NAP  # noqa: E402

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def sym_couplings(L, rng, scale=0.01):
    nsp = nsingle_padded(L)
    W = rng.normal(0, scale, (NA, L, NA, L))
    W = W + W.transpose(2, 3, 0, 1)
    for k in range(L):
        W[:, k, :, k] = 0.0
    x = np.zeros(nsp + NA * L * NAP * L)
    X2 = x[nsp:].reshape(NA, L, NAP, L)
    X2[:, :, :NA, :] = W
    return x


def make_x(msa, seed, scale=0.01):
    L = msa.shape[1]
    x = ora.init_bias(msa).astype(np.float64)
    if scale:
        x += sym_couplings(L, np.random.default_rng(seed), scale)
    return x.astype(np.float32)


# ---- reweighting ---------------------------------------------------------

@pytest.mark.parametrize("L,N,seed,theta", [(75, 1000, 1, 0.8), (75, 1000, 1, 0.3),
                                            (40, 300, 7, 0.8), (33, 257, 9, 0.6),
                                            (75, 400, 2, 1.0), (20, 50, 3, 1.5),
                                            (250, 5000, 2, 0.8)])
def test_weights_bit_exact(L, N, seed, theta):
    say(f"weights L={L} N={N} seed={seed} theta={theta}")
    msa, _ = synthetic_msa(L, N, seed)
    msa[N // 3: N // 3 + N // 10] = msa[0]          # force identity clusters
    msa[N // 2: N // 2 + 5, : L // 2] = msa[1, : L // 2]
    w_ref = ora.calculate_weights(msa, theta)
    with P.PLL(msa, theta=theta) as p:
        assert p.weights.tobytes() == w_ref.tobytes()


# ---- FP32 plain path -----------------------------------------------------

@pytest.mark.parametrize("L,N,seed,scale", [(75, 1000, 1, 0.0), (75, 1000, 1, 0.01),
                                            (120, 2000, 5, 0.01), (13, 37, 4, 0.1),
                                            (64, 129, 6, 0.05)])
def test_fp32_evaluate_parity(L, N, seed, scale):
    say(f"fp32 evaluate L={L} N={N} seed={seed} scale={scale}")
    msa, _ = synthetic_msa(L, N, seed)
    x = make_x(msa, seed + 100, scale)
    with P.PLL(msa, precision="fp32") as p:
        lam1, lam2 = p.lambda_single, p.lambda_pair
        f64, g64 = ora.evaluate_f64(msa, p.weights.astype(np.float64), x.astype(np.float64), lam1, lam2)
        f32, g32 = ora.evaluate_f32(msa, p.weights, x, lam1, lam2)
        xd = torch.from_numpy(x).cuda()
        f, g = p.evaluate(xd)
        g = g.cpu().numpy()
    say(f"  f={f:.10g} f64={f64:.10g} f32={f32:.10g} rel={abs(f - f64) / abs(f64):.2e} "
        f"g_rel_l2={rel_l2(g, g64):.2e} oracle32_g_rel={rel_l2(g32, g64):.2e}")
    assert abs(f - f64) / abs(f64) <= 1e-5
    assert rel_l2(g, g64) <= 1e-4
    assert rel_l2(g, g32.astype(np.float64)) <= 1e-4


def test_fp32_deterministic_and_structured():
    """Full-size workload L=250, N=5000: two evaluations are bitwise equal, the
    pair gradient is symmetric, padding and self-edge gradients are 2*lam2*x."""
    L, N = 250, 5000
    say(f"fp32 structure L={L} N={N}")
    msa, _ = synthetic_msa(L, N, 2, 60)
    x = make_x(msa, 7, 0.01)
    nsp = nsingle_padded(L)
    with P.PLL(msa, precision="fp32") as p:
        xd = torch.from_numpy(x).cuda()
        f1, g1 = p.evaluate(xd)
        f2, g2 = p.evaluate(xd)
        assert f1 == f2 and torch.equal(g1, g2)
        G = g1[nsp:].view(NA, L, NAP, L)
        T = G[:, :, :NA, :]
        lam2 = p.lambda_pair
        assert torch.equal(T, T.permute(2, 3, 0, 1))
        X = xd[nsp:].view(NA, L, NAP, L)
        assert torch.equal(G[:, :, NA:, :], (2.0 * np.float32(lam2)) * X[:, :, NA:, :])
        idx = torch.arange(L, device=G.device)
        assert torch.all(G[:, idx, :, idx] == 0)
        assert torch.all(g1[20 * L: nsp] == 0)


@pytest.mark.parametrize("prec", ["bf16", "bf16f"])
def test_bf16_deterministic_and_structured(prec):
    """The BF16 forms at L=250, N=5000: bitwise repeatable, pair gradient
    exactly symmetric (the constant count block is added so that both
    orientations sum the same operands), padding 2*lam2*x, self-edges 0."""
    L, N = 250, 5000
    say(f"{prec} structure L={L} N={N}")
    msa, _ = synthetic_msa(L, N, 2, 60)
    x = make_x(msa, 7, 0.01)
    nsp = nsingle_padded(L)
    # nonzero padding couplings: they enter only through 2*lam2*x
    x[nsp:].reshape(NA, L, NAP, L)[:, :, NA:, :] = np.random.default_rng(8).normal(
        0, 0.01, (NA, L, NAP - NA, L)).astype(np.float32)
    with P.PLL(msa, precision=prec) as p:
        xd = torch.from_numpy(x).cuda()
        f1, g1 = p.evaluate(xd)
        f2, g2 = p.evaluate(xd)
        assert f1 == f2 and torch.equal(g1, g2)
        G = g1[nsp:].view(NA, L, NAP, L)
        T = G[:, :, :NA, :]
        assert torch.equal(T, T.permute(2, 3, 0, 1))
        X = xd[nsp:].view(NA, L, NAP, L)
        assert torch.equal(G[:, :, NA:, :], (2.0 * np.float32(p.lambda_pair)) * X[:, :, NA:, :])
        idx = torch.arange(L, device=G.device)
        assert torch.all(T[:, idx, :, idx] == 0)
        assert torch.all(g1[20 * L: nsp] == 0)


@pytest.mark.parametrize("prec", ["bf16", "bf16x2"])
@pytest.mark.parametrize("L,N", [(75, 1000), (37, 300), (130, 700), (250, 700), (418, 300), (700, 200)])
def test_tile_variants_bit_identical(prec, L, N):
    """Every MFMA tile variant (register-staged or direct-to-LDS, 256x256,
    256x128, 128x128, the gradient's 192x256; the forward's softmax epilogue
    at 12 or 6 alignment columns per tile) sums in the same order: f and g
    equal the default launch's bit for bit.  L = 75, 37, 130 leave the last
    tile rows partly past L; N = 300, 700 leave partly filled column tiles.
    At L = 250, 418 and 700 the default gradient launch is banded (256-row
    tiles over the first rows, 192-row tiles over the rest; 250: rows
    0..3071 of 5376, 418: rows 0..6911 of 9408, 700: rows 0..12287 of
    14784, whose last column tile is 192 wide)."""
    say(f"tile variants {prec} L={L} N={N}")
    msa, _ = synthetic_msa(L, N, 41, 6)
    xd = torch.from_numpy(make_x(msa, 42, 0.02)).cuda()
    with P.PLL(msa, precision=prec) as p:
        f0, g0 = p.evaluate(xd)
    nv = _lib.lib().ccm_gemm_variants()
    assert nv == 7
    for v in range(nv):
        with P.PLL(msa, precision=prec, gemm_variant=v) as p:
            f, g = p.evaluate(xd)
        say(f"  variant {v}: f={f:.17g} (default {f0:.17g}) g equal={torch.equal(g, g0)}")
        assert f == f0 and torch.equal(g, g0)


def test_gradient_only_variant_with_dense_forward():
    """The gradient-only tile (variant 6) requested with the dense fused
    forward: the forward runs variant 3 in its place, the gradient variant
    6; f and g equal the default dense-forward launch bit for bit."""
    L, N = 130, 700
    say(f"variant 6 with the dense forward L={L} N={N}")
    msa, _ = synthetic_msa(L, N, 43, 6)
    xd = torch.from_numpy(make_x(msa, 44, 0.02)).cuda()
    with P.PLL(msa, precision="bf16", sparse=False) as p:
        f0, g0 = p.evaluate(xd)
    with P.PLL(msa, precision="bf16", sparse=False, gemm_variant=6) as p:
        f, g = p.evaluate(xd)
    assert f == f0 and torch.equal(g, g0)


@pytest.mark.parametrize("prec", ["bf16", "bf16x2"])
def test_chunked_banded_gradient(prec):
    """Chunks accumulate into the stored pair gradient (the launch's
    accumulate mode): at L = 250 the default gradient launch is banded, and
    its chunked f and g equal the single-shape (variant 3) chunked ones bit
    for bit; N = 1100 in chunks of 512 leaves a ragged last chunk."""
    L, N = 250, 1100
    say(f"chunked banded gradient {prec} L={L} N={N}")
    msa, _ = synthetic_msa(L, N, 33, 20)
    xd = torch.from_numpy(make_x(msa, 34, 0.02)).cuda()
    with P.PLL(msa, precision=prec, chunk_rows=512) as p:
        f0, g0 = p.evaluate(xd)
    with P.PLL(msa, precision=prec, chunk_rows=512, gemm_variant=3) as p:
        f, g = p.evaluate(xd)
    say(f"  f banded={f0:.17g} variant 3={f:.17g} g equal={torch.equal(g, g0)}")
    assert f == f0 and torch.equal(g, g0)


@pytest.mark.parametrize("prec", ["fp32", "bf16", "bf16f"])
@pytest.mark.parametrize("chunk", [512, 1000, 3000])
def test_chunked_evaluate_matches_unchunked(prec, chunk):
    """Sequences processed in chunks (rounded up to the softmax block: 256
    sequences FP32, 512 BF16; N = 2900 leaves a ragged last chunk) give the
    unchunked f and g up to the order of the pair-gradient accumulation
    over chunks: f within 1e-9 relative, g within 1e-6 relative L2."""
    L, N = 75, 2900
    say(f"chunked {prec} chunk_rows={chunk} L={L} N={N}")
    msa, _ = synthetic_msa(L, N, 31, 12)
    x = make_x(msa, 32, 0.02)
    xd = torch.from_numpy(x).cuda()
    with P.PLL(msa, precision=prec) as whole:
        f0, g0 = whole.evaluate(xd)
        g0 = g0.cpu().numpy().astype(np.float64)
    with P.PLL(msa, precision=prec, chunk_rows=chunk) as part:
        f1, g1 = part.evaluate(xd)
        g1 = g1.cpu().numpy().astype(np.float64)
    rel = np.linalg.norm(g1 - g0) / np.linalg.norm(g0)
    say(f"  f whole={f0:.12g} chunked={f1:.12g} g rel diff={rel:.2e}")
    assert abs(f1 - f0) <= 1e-9 * abs(f0)
    assert rel <= 1e-6


def test_x_zero_closed_form_full_size():
    """At x = 0: f = L ln21 sum(w), g1 = sum(w)/21 - C, g2 = (C_kb + C_ja)/21 - 2 H2
    with H2 = A^T diag(w) A formed independently from the one-hot MSA."""
    L, N = 250, 5000
    say(f"x=0 closed form L={L} N={N}")
    msa, _ = synthetic_msa(L, N, 2, 60)
    nsp = nsingle_padded(L)
    for prec in ("fp32", "bf16"):
        with P.PLL(msa, precision=prec) as p:
            w = p.weights.astype(np.float64)
            xd = torch.zeros(p.nvar, dtype=torch.float32, device="cuda")
            f, g = p.evaluate(xd)
            W = w.sum()
            assert abs(f - L * np.log(21.0) * W) / (L * np.log(21.0) * W) < 1e-6
            A = torch.nn.functional.one_hot(torch.from_numpy(msa.astype(np.int64)), NA).double()
            C = torch.einsum("n,nkb->kb", torch.from_numpy(w), A)     # (k, b)
            g1 = g[: 20 * L].view(20, L).double().cpu()
            assert torch.allclose(g1, (W / 21.0 - C[:, :20]).T, atol=1e-3, rtol=1e-5)
            Ad = A.cuda().reshape(N, L * NA).float()                     # column (k, b)
            H = (Ad * torch.from_numpy(w).float().cuda()[:, None]).T @ Ad   # (k,b) x (j,a)
            H = H.view(L, NA, L, NA).permute(1, 0, 3, 2).double()        # (b,k,a,j)
            Cd = C.cuda().T                                              # (b, k)
            expect = (Cd[:, :, None, None] + Cd[None, None, :, :]) / 21.0 - 2 * H
            idx = torch.arange(L, device="cuda")
            expect[:, idx, :, idx] = 0
            G = g[nsp:].view(NA, L, NAP, L)[:, :, :NA, :].double()
            err = (G - expect).norm() / expect.norm()
            say(f"  {prec}: f rel={abs(f - L * np.log(21.0) * W) / (L * np.log(21.0) * W):.2e} g2 rel={err:.2e}")
            assert err < (1e-5 if prec == "fp32" else 1e-3)


# ---- BF16 MFMA path ------------------------------------------------------

@pytest.mark.parametrize("prec,tol", [("bf16", 1e-3), ("bf16x2", 1e-4)])
@pytest.mark.parametrize("L,N,seed", [(75, 1000, 1), (120, 2000, 5), (37, 300, 3)])
def test_bf16_evaluate_error(prec, tol, L, N, seed):
    say(f"{prec} evaluate L={L} N={N} seed={seed}")
    msa, _ = synthetic_msa(L, N, seed)
    x = make_x(msa, seed + 100, 0.01)
    with P.PLL(msa, precision=prec) as p:
        f64, g64 = ora.evaluate_f64(msa, p.weights.astype(np.float64), x.astype(np.float64),
                                    p.lambda_single, p.lambda_pair)
        f, g = p.evaluate(torch.from_numpy(x).cuda())
        g = g.cpu().numpy()
    e = rel_l2(g, g64)
    e1 = rel_l2(g[:20 * L], g64[:20 * L])
    say(f"  f rel={abs(f - f64) / abs(f64):.2e} g rel_l2={e:.2e} singles rel_l2={e1:.2e}")
    assert abs(f - f64) / abs(f64) <= 1e-4
    assert e <= tol
    # the node gradient on its own (a small part of g): the single-pass BF16
    # logits leave ~1.5e-3 there; a broken reduction would give O(1) errors
    assert e1 <= 5 * tol


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-4), ("bf16x2", 1e-4), ("bf16f", 1e-3), ("bf16", 1e-3)])
@pytest.mark.parametrize("L,N", [(1, 1), (2, 3), (3, 130), (12, 128), (13, 129), (7, 513)])
def test_tiny_shapes(prec, tol, L, N):
    """Degenerate sizes against the FP64 oracle: one column (only self-edges),
    one sequence, a single tile row holding 1..13 columns, N at and just past
    the 128-sequence blocks and the 512-sequence chunk granule.  Random MSA
    (gaps included) and random symmetric couplings; the same tolerances as
    at full size."""
    say(f"tiny {prec} L={L} N={N}")
    rng = np.random.default_rng(L * 1000 + N)
    msa = rng.integers(0, 21, (N, L), dtype=np.uint8)
    x = make_x(msa, L + N, 0.05)
    with P.PLL(msa, precision=prec) as p:
        f64, g64 = ora.evaluate_f64(msa, p.weights.astype(np.float64), x.astype(np.float64),
                                    p.lambda_single, p.lambda_pair)
        f, g = p.evaluate(torch.from_numpy(x).cuda())
        g = g.cpu().numpy()
    e = rel_l2(g, g64)
    say(f"  f={f:.12g} f64={f64:.12g} rel={abs(f - f64) / abs(f64):.2e} g rel_l2={e:.2e}")
    assert np.all(np.isfinite(g))
    assert abs(f - f64) / abs(f64) <= (1e-5 if prec in ("fp32", "bf16x2") else 1e-4)
    assert e <= tol


@pytest.mark.parametrize("prec", ["fp32", "bf16", "bf16f"])
@pytest.mark.parametrize("L,N,chunk", [(37, 300, 0), (130, 700, 0), (61, 1100, 512), (250, 600, 0)])
def test_padding_regularisation(prec, L, N, chunk):
    """Nonzero a >= 21 padding couplings and nonzero singles against the FP64
    oracle: the padding enters f through 0.5 lam2 x^2 and g through
    2 lam2 x (evaluate_cpu.c:151-163).  The sparse BF16 forms write both in
    the coupling conversion and form f inside the assembly; the padding
    holds about 8 % of f here, so a missing or doubled term fails the f
    bound.  L = 37, 61, 130 make 21 L no multiple of four (the padding
    columns start unaligned); one case runs in sequence chunks."""
    say(f"padding reg {prec} L={L} N={N} chunk={chunk}")
    rng = np.random.default_rng(L + 7 * N)
    msa = rng.integers(0, 21, (N, L), dtype=np.uint8)
    x = make_x(msa, L + N, 0.02)
    nsp = nsingle_padded(L)
    x[nsp:].reshape(NA, L, NAP, L)[:, :, NA:, :] = rng.normal(0, 0.05, (NA, L, NAP - NA, L))
    with P.PLL(msa, precision=prec, chunk_rows=chunk) as p:
        f64, g64 = ora.evaluate_f64(msa, p.weights.astype(np.float64), x.astype(np.float64),
                                    p.lambda_single, p.lambda_pair)
        xd = torch.from_numpy(x).cuda()
        f, g = p.evaluate(xd)
        G = g[nsp:].view(NA, L, NAP, L)
        X = xd[nsp:].view(NA, L, NAP, L)
        assert torch.equal(G[:, :, NA:, :], 0.0 + (2.0 * np.float32(p.lambda_pair)) * X[:, :, NA:, :])
        g = g.cpu().numpy()
    e = rel_l2(g, g64)
    say(f"  f={f:.12g} f64={f64:.12g} rel={abs(f - f64) / abs(f64):.2e} g rel_l2={e:.2e}")
    assert abs(f - f64) / abs(f64) <= (1e-5 if prec == "fp32" else 1e-4)
    assert e <= (1e-4 if prec == "fp32" else 1e-3)


# ---- device CG end to end ------------------------------------------------

def _spearman(a, b):
    from scipy.stats import spearmanr
    iu = np.triu_indices(a.shape[0], 1)
    return spearmanr(a[iu], b[iu]).correlation


@pytest.mark.parametrize("prec", ["fp32", "bf16f", "bf16"])
def test_cg_50_iterations_contacts(prec):
    """Device CG vs the same CG on the host driven by the FP64 oracle: all 50
    iterations in every form (single-pass BF16 quantises f near the optimum;
    the approximate decrease test on the slope keeps its line search going)."""
    L, N = 75, 1000
    say(f"cg50 {prec} L={L} N={N}")
    msa, pairs = synthetic_msa(L, N, 1, 15)
    x0 = ora.init_bias(msa)
    with P.PLL(msa, precision=prec) as p:
        host = ora.cg_pll_f32(msa, p.weights, x0, p.lambda_single, p.lambda_pair, max_iterations=50,
                              epsilon=0.0, evaluator="f64")
        xd = torch.from_numpy(x0.copy()).cuda()
        dev = p.conjugrad(xd, max_iterations=50, epsilon=0.0)
        xg = xd.cpu().numpy()
    m_dev = P.contact_map(xg, L)
    m_host = ora.apc(ora.sum_submatrices(host["x"], L))
    rho = _spearman(m_dev, m_host)
    top_dev = set(P.top_pairs(m_dev, L // 5, 6))
    top_host = set(P.top_pairs(m_host, L // 5, 6))
    planted = set(map(tuple, pairs.tolist()))
    say(f"  host ret={host['ret']} it={host['iterations']} evals={host['evaluations']} f={host['fx']:.6f} | "
        f"dev ret={dev['ret']} it={dev['iterations']} evals={dev['evaluations']} f={dev['fx']:.6f}")
    say(f"  spearman={rho:.6f} top-L/5 equal={top_dev == top_host} planted recovered={len(top_dev & planted)}")
    assert host["iterations"] == 50
    assert dev["iterations"] == 50
    assert abs(dev["fx"] - host["fx"]) / host["fx"] < 5e-3
    assert rho >= 0.999
    assert top_dev == top_host


def test_cli_1atz(tmp_path):
    """The reference's only test: the contact matrix of example/1atzA.aln has
    75 lines.  Also: the matrix is the APC map of the couplings the -r raw
    file holds (bit for bit: both print floats with %.20e), the -b
    MessagePack file holds the same couplings, and a warm start from the raw
    file (-i) continues from the same objective."""
    import ctypes as C
    msgpack = pytest.importorskip("msgpack")
    out = tmp_path / "1atzA.mat"
    raw = tmp_path / "1atzA.raw"
    braw = tmp_path / "1atzA.braw"
    aln = os.path.join(ROOT, "tests", "golden", "1atzA.aln")
    exe = os.path.join(ROOT, "ccmpred_amd", "bin", "ccmpred")
    say("cli 1atzA")
    r = subprocess.run([exe, "-n", "20", "-r", str(raw), "-b", str(braw), aln, str(out)],
                       capture_output=True, text=True, timeout=300)
    sys.stdout.write(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr
    lines = out.read_text().splitlines()
    assert len(lines) == 75
    mat = np.array([[float(v) for v in ln.split()] for ln in lines], dtype=np.float32)
    assert mat.shape == (75, 75) and np.all(np.diag(mat) == 0)
    assert np.allclose(mat, mat.T, atol=1e-5)

    L = 75
    from ccmpred_amd import _lib
    lib = _lib.lib()
    x = np.zeros(P.nvar_padded(L), np.float32)
    assert lib.ccm_read_raw(str(raw).encode(), L, C.c_void_p(x.ctypes.data)) == 0
    assert np.array_equal(P.contact_map(x, L, apc=True), mat)
    doc = msgpack.unpackb(braw.read_bytes(), strict_map_key=False)
    assert doc["format"] == "ccm-1" and doc["ncol"] == L and len(doc["x_pair"]) == L * (L - 1) // 2
    X2 = x[P.nsingle_padded(L):].reshape(NA, L, NAP, L)
    for e in doc["x_pair"].values():
        blk = np.asarray(e["x"], np.float32).reshape(NA, NA)  # [a][b] = W(b,j,a,i)
        assert np.array_equal(blk.T, X2[:, e["j"], :NA, e["i"]])
    singles = np.asarray(doc["x_single"], np.float32).reshape(L, 20).T.ravel()
    assert np.allclose(singles, x[:20 * L], rtol=1e-5, atol=1e-6)  # the raw text prints singles with %g

    def final_fx(text):
        return float([ln for ln in text.splitlines() if ln.startswith("Final fx")][0].split("=")[1])

    say("cli 1atzA warm start")
    out2 = tmp_path / "warm.mat"
    r2 = subprocess.run([exe, "-n", "2", "-i", str(raw), aln, str(out2)], capture_output=True, text=True,
                        timeout=300)
    sys.stdout.write(r2.stdout[-1500:])
    assert r2.returncode == 0, r2.stderr
    f1, f2 = final_fx(r.stdout), final_fx(r2.stdout)
    say(f"  final fx {f1} -> warm start {f2}")
    assert f2 <= f1 * (1 + 1e-4)


@pytest.mark.parametrize("prec", ["fp32", "bf16f", "bf16"])
def test_cli_1atz_against_fp64(tmp_path, prec):
    """The reference's example run (test/test.sh:1-9, ccmpred example/1atzA.aln)
    through the command-line driver with 50 iterations and no early stop
    (-n 50 -e 0): its .mat (ccmpred.c:683-715: sum_submatrices, APC,
    write_matrix) against the APC map of the FP64-oracle CG on the same
    alignment (tests/golden/cg_maps/1atzA.npy), under the criteria of
    tests/cg_fixtures.py."""
    from tests import cg_fixtures as F
    meta, m64 = F.load("1atzA")
    aln = os.path.join(ROOT, "tests", "golden", "1atzA.aln")
    assert F.msa_sha256(read_psicov(aln)) == meta["msa_sha256"]
    exe = os.path.join(ROOT, "ccmpred_amd", "bin", "ccmpred")
    out = tmp_path / f"1atzA_{prec}.mat"
    say(f"cli 1atzA -n 50 -e 0 -P {prec}")
    r = subprocess.run([exe, "-n", "50", "-e", "0", "-P", prec, aln, str(out)], capture_output=True,
                       text=True, timeout=300)
    sys.stdout.write(r.stdout[-800:])
    assert r.returncode == 0, r.stderr
    iters = [ln.split()[0] for ln in r.stdout.splitlines() if ln[:1].isdigit()]
    assert iters and int(iters[-1]) == 50
    lines = out.read_text().splitlines()
    assert len(lines) == 75
    mat = np.array([[float(v) for v in ln.split()] for ln in lines], dtype=np.float32)
    c = F.compare(mat, m64)
    say(f"  spearman={c['rho']:.6f} top-L/5 identical={c['identical']} swapped={c['swapped']}")
    assert c["rho"] >= 0.999
    assert c["identical"] or c["swaps_are_near_ties"]

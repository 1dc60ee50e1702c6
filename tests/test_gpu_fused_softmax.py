"""The forward contraction with the column softmax as its epilogue
(ccm_launch_gemm_fwd_softmax) against the unfused pair it replaces
(ccm_launch_gemm_fwd, then ccm_launch_softmax_cols), through the C ABI on the
same device buffers.

Both form every probability with the same float operations, so Dt (BF16,
and its residual half for two-pass forms) must agree bit for bit.  The
partial sums come at different granularity (128- against 512-sequence
blocks, 16-lane against 64-lane sums) and the fused partials leave out the
constant data term -sum_n w_n [x_nj = a], added here from the MSA and the
weights: the objective then agrees within 1e-12 relative and the node
gradient within 1e-6 relative L2.  Shapes leave the last tile rows partly
past L (12 alignment columns per tile row) and the last column tile partly
past N; the two-pass forward operand and the two-pass Dt are both covered.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from ccmpred_amd import _lib  # noqa: E402
from ccmpred_amd.msa import synthetic_msa  # noqa: E402
from ccmpred_amd.pll import init_bias, nsingle_padded  # noqa: E402

pytestmark = pytest.mark.gpu

NA, NAP = 21, 32


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.fixture(scope="module")
def lib():
    L = _lib.lib()
    vp, ip, lp = C.c_void_p, C.c_int, C.c_long
    L.ccm_launch_convert_wt.argtypes = [vp, ip, ip, vp, lp, ip, vp]
    L.ccm_launch_gemm_fwd.argtypes = [vp, lp, ip, ip, vp, ip, vp, lp, ip, vp]
    L.ccm_launch_softmax_cols.argtypes = [vp, lp, vp, vp, lp, vp, ip, ip, ip, ip, ip, vp, lp, lp, ip, ip,
                                          vp, vp, vp]
    L.ccm_launch_gemm_fwd_softmax.argtypes = [vp, lp, ip, ip, vp, ip, vp, vp, lp, vp, ip, ip, vp, lp, lp,
                                              ip, ip, ip, vp, vp, ip, vp]
    for f in ("ccm_launch_convert_wt", "ccm_launch_gemm_fwd", "ccm_launch_softmax_cols",
              "ccm_launch_gemm_fwd_softmax", "ccm_softmax_cols_nblk", "ccm_fwd_softmax_blocks"):
        getattr(L, f).restype = ip
    return L


def ptr(t):
    return C.c_void_p(t.data_ptr())


@pytest.mark.parametrize("L,N,pf,pd,variant", [(75, 1000, 1, 1, -1), (37, 300, 2, 1, -1),
                                               (130, 700, 2, 2, -1), (75, 1000, 1, 2, 5),
                                               (250, 1300, 1, 1, 4)])
def test_fused_matches_unfused(lib, L, N, pf, pd, variant):
    print(f"fused softmax L={L} N={N} forward passes={pf} Dt passes={pd} variant={variant}", flush=True)
    msa, _ = synthetic_msa(L, N, 51, max(1, L // 10))
    rng = np.random.default_rng(52)
    x = init_bias(msa).astype(np.float64)
    nsp = nsingle_padded(L)
    W = rng.normal(0, 0.05, (NA, L, NA, L))
    W = W + W.transpose(2, 3, 0, 1)
    W[:, np.arange(L), :, np.arange(L)] = 0
    x[nsp:].reshape(NA, L, NAP, L)[:, :, :NA, :] = W
    x = x.astype(np.float32)
    w = rng.uniform(0.2, 1.0, N).astype(np.float32)

    Lp = (L + 63) // 64 * 64
    Np = (N + 127) // 128 * 128
    P = NA * Lp
    ncols = Np
    msap = np.full((Np, Lp), 255, np.uint8)
    msap[:N, :L] = msa
    dev = "cuda"
    d_msap = torch.from_numpy(msap).to(dev)
    d_msapt = torch.from_numpy(np.ascontiguousarray(msap.T)).to(dev)
    d_x = torch.from_numpy(x).to(dev)
    d_w = torch.from_numpy(w).to(dev)
    lda = pf * P
    d_wt = torch.zeros(P * lda, dtype=torch.int16, device=dev)  # convert_wt leaves padding rows alone
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib.ccm_launch_convert_wt(ptr(d_x), L, Lp, ptr(d_wt), lda, pf, s) == 0

    lddt = pd * ncols
    # unfused
    d_zt = torch.empty(P * ncols, dtype=torch.float32, device=dev)
    dt_a = torch.zeros(P * lddt, dtype=torch.int16, device=dev)
    nb_a = lib.ccm_softmax_cols_nblk(N)
    fp_a = torch.zeros(nb_a * L, dtype=torch.float64, device=dev)
    g1_a = torch.zeros(nb_a * 20 * L, dtype=torch.float64, device=dev)
    assert lib.ccm_launch_gemm_fwd(ptr(d_wt), lda, Lp, pf, ptr(d_msap), ncols, ptr(d_zt), ncols,
                                   variant, s) == 0
    assert lib.ccm_launch_softmax_cols(ptr(d_zt), ncols, ptr(d_x), ptr(d_msapt), Np, ptr(d_w), L, Lp, N,
                                       ncols, 0, ptr(dt_a), lddt, ncols, pd, 1, ptr(fp_a), ptr(g1_a),
                                       s) == 0
    # fused
    dt_b = torch.zeros(P * lddt, dtype=torch.int16, device=dev)
    nb_b = lib.ccm_fwd_softmax_blocks(N)
    fp_b = torch.zeros(nb_b * L, dtype=torch.float64, device=dev)
    g1_b = torch.zeros(nb_b * 20 * L, dtype=torch.float64, device=dev)
    assert lib.ccm_launch_gemm_fwd_softmax(ptr(d_wt), lda, Lp, pf, ptr(d_msap), ncols, ptr(d_x),
                                           ptr(d_msapt), Np, ptr(d_w), L, N, ptr(dt_b), lddt, ncols, pd,
                                           1, 0, ptr(fp_b), ptr(g1_b), variant, s) == 0
    torch.cuda.synchronize()

    # Dt rows a*Lp + j of the real columns j < L, all ncols sequences (both passes)
    rows = (np.arange(NA)[:, None] * Lp + np.arange(L)[None, :]).ravel()
    A = dt_a.view(P, lddt)[torch.from_numpy(rows).to(dev)]
    B = dt_b.view(P, lddt)[torch.from_numpy(rows).to(dev)]
    ndiff = int((A != B).sum())
    assert ndiff == 0, ndiff

    fa, fb = float(fp_a.sum()), float(fp_b.sum())
    ga = g1_a.view(nb_a, 20, L).sum(0).cpu().numpy()
    gb = g1_b.view(nb_b, 20, L).sum(0).cpu().numpy()
    # the data term the fused partials leave out
    Cnt = np.zeros((20, L))
    for a in range(20):
        Cnt[a] = ((msa == a) * w[:, None].astype(np.float64)).sum(0)
    gb = gb - Cnt
    e = np.linalg.norm(gb - ga) / np.linalg.norm(ga)
    print(f"  Dt equal, f unfused={fa:.15g} fused={fb:.15g}, node gradient rel diff={e:.2e}", flush=True)
    assert abs(fa - fb) <= 1e-12 * abs(fa)
    assert e <= 1e-6


@pytest.mark.parametrize("L,pf", [(1, 1), (37, 2), (130, 1), (250, 2)])
def test_convert_wt_aligned_and_unaligned_x(lib, L, pf):
    """ccm_launch_convert_wt takes 16-byte loads for a 16-byte aligned x and
    4-byte loads otherwise: both write the same Wt, bit for bit."""
    print(f"convert_wt L={L} passes={pf}", flush=True)
    rng = np.random.default_rng(L + pf)
    nsp = nsingle_padded(L)
    n = nsp + NA * NAP * L * L
    x = rng.normal(0, 0.3, n).astype(np.float32)
    Lp = (L + 63) // 64 * 64
    P = NA * Lp
    lda = pf * P
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    buf = torch.from_numpy(np.concatenate([np.zeros(4, np.float32), x])).cuda()
    xa = buf[4:]      # 16-byte aligned (torch allocations are)
    xu = torch.empty(n + 1, dtype=torch.float32, device="cuda")
    xu[1:] = xa       # x at a 4-byte offset
    assert xa.data_ptr() % 16 == 0 and xu[1:].data_ptr() % 16 == 4
    wa = torch.zeros(P * lda, dtype=torch.int16, device="cuda")
    wu = torch.zeros_like(wa)
    assert lib.ccm_launch_convert_wt(C.c_void_p(xa.data_ptr()), L, Lp, ptr(wa), lda, pf, s) == 0
    assert lib.ccm_launch_convert_wt(C.c_void_p(xu[1:].data_ptr()), L, Lp, ptr(wu), lda, pf, s) == 0
    torch.cuda.synchronize()
    assert torch.equal(wa, wu)
    # and the values: hi part of Wt[a*Lp + j][b*Lp + k] is W(b,k,a,j) rounded to BF16
    W = x[nsp:].reshape(NA, L, NAP, L)[:, :, :NA, :]
    hi = wa.view(P, lda)[:, :P].cpu().numpy().view(np.uint16).reshape(NA, Lp, NA, Lp)
    got = (hi[:, :L, :, :L].astype(np.uint32) << 16).view(np.float32)   # [a, j, b, k]
    ref = W.transpose(2, 3, 0, 1)                                          # [a, j, b, k]
    assert np.all(np.abs(got - ref) <= np.abs(ref) * 2.0 ** -8)
    assert not hi[:, :L, :, L:].any()   # columns k >= L are zero

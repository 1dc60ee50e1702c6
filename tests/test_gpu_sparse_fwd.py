"""The forward contraction on 2:4 structured-sparse MFMA
(ccm_launch_sparse_fwd_softmax, the default BF16 forward) against the dense
fused forward (ccm_launch_gemm_fwd_softmax) and the FP64 oracle.

Both forwards multiply the same BF16 couplings by exact 0/1 one-hot values
and accumulate in FP32; they differ only in the order of the FP32 sums, so
f agrees to 1e-6 relative and the gradients to 2e-3 relative L2 between the
two (the BF16 rounding of the probability operand can flip where the
logits differ in the last bits), and each meets the BF16 tolerances against
the FP64 oracle (f 1e-5, gradient 1e-3 relative L2).  Shapes cover L not a
multiple of 64 (padding columns), N not a multiple of 128 or 256 (the
256-sequence tiles' padding rows), a single column, one sequence, and
both forward forms (one pass; hi + lo halves).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as ora  # noqa: E402
from ccmpred_amd import pll as P  # noqa: E402
from ccmpred_amd.msa import synthetic_msa  # noqa: E402
from tests.pll_bruteforce import nsingle_padded, NA, NAP  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def start_x(msa, seed, scale=0.01):
    L = msa.shape[1]
    x = ora.init_bias(msa).astype(np.float64)
    rng = np.random.default_rng(seed)
    W = rng.normal(0, scale, (NA, L, NA, L))
    W = W + W.transpose(2, 3, 0, 1)
    W[:, np.arange(L), :, np.arange(L)] = 0
    x[nsingle_padded(L):].reshape(NA, L, NAP, L)[:, :, :NA, :] += W
    return x.astype(np.float32)


def rel_l2(a, b):
    return float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("prec", ["bf16", "bf16f"])
@pytest.mark.parametrize("L,N,seed", [(1, 1, 21), (7, 300, 22), (64, 256, 23), (70, 1000, 24),
                                      (129, 777, 25), (250, 5000, 2)])
def test_sparse_forward_matches_dense_and_fp64(prec, L, N, seed):
    print(f"sparse forward {prec} L={L} N={N}", flush=True)
    if L < 2:
        rng = np.random.default_rng(seed)
        msa = rng.integers(0, 21, (N, L), dtype=np.uint8)
    else:
        msa, _ = synthetic_msa(L, N, seed, max(1, L // 12))
    x = start_x(msa, seed + 7, 0.05 if L < 100 else 0.01)
    xd = torch.from_numpy(x).cuda()
    out = {}
    for sparse in (True, False):
        with P.PLL(msa, precision=prec, sparse=sparse) as p:
            f, g = p.evaluate(xd)
            out[sparse] = (f, g.cpu().numpy(), p.weights.copy(), p.lambda_single, p.lambda_pair)
    fs, gs, w, l1, l2 = out[True]
    fd, gd = out[False][0], out[False][1]
    f64, g64 = ora.evaluate_f64_par(msa, w.astype(np.float64), x.astype(np.float64), l1, l2)
    ef_sd = abs(fs - fd) / abs(fd)
    eg_sd = rel_l2(gs, gd.astype(np.float64))
    ef, eg = abs(fs - f64) / abs(f64), rel_l2(gs, g64)
    efd, egd = abs(fd - f64) / abs(f64), rel_l2(gd, g64)
    print(f"  sparse vs dense: f {ef_sd:.2e} g {eg_sd:.2e} | vs FP64: sparse f {ef:.2e} g {eg:.2e}, "
          f"dense f {efd:.2e} g {egd:.2e}", flush=True)
    assert np.all(np.isfinite(gs))
    assert ef_sd <= 1e-6 and eg_sd <= 2e-3
    assert ef <= 1e-5 and eg <= 1e-3


@pytest.mark.parametrize("chunk", [512, 1024])
def test_sparse_forward_chunked(chunk):
    """Sequence chunks start the sparse forward's positions at a multiple of
    512 sequences: the chunked evaluation equals the unchunked one up to the
    order of the gradient's sums over chunks."""
    msa, _ = synthetic_msa(90, 2300, 31, 7)
    x = start_x(msa, 38)
    xd = torch.from_numpy(x).cuda()
    with P.PLL(msa, precision="bf16") as whole:
        f0, g0 = whole.evaluate(xd)
        g0 = g0.cpu().numpy().astype(np.float64)
    with P.PLL(msa, precision="bf16", chunk_rows=chunk) as part:
        f1, g1 = part.evaluate(xd)
        g1 = g1.cpu().numpy()
    print(f"chunk {chunk}: f {abs(f1 - f0) / abs(f0):.2e} g {rel_l2(g1, g0):.2e}", flush=True)
    assert abs(f1 - f0) / abs(f0) <= 1e-9 and rel_l2(g1, g0) <= 1e-6


def bf16_bits(v):
    """FP32 -> BF16 bits, round to nearest even (finite values)"""
    u = np.ascontiguousarray(v, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def ws_expected(x, L, Lp, passes, lds, fill):
    """The sparse forward's coupling operand from x in numpy: row c = a Lp + j
    (x-layout column C = a L + j), group G = 7 k + t of column k holds the
    slots (0, W(3t, k), W(3t+1, k), W(3t+2, k)); with two passes each
    instruction's 32 slots (8 groups) are followed by their lo halves."""
    nsp = nsingle_padded(L)
    x2 = x[nsp:].reshape(NA * L, NAP * L)
    ws = np.full((NA * Lp, lds), fill, np.uint16)
    G = np.arange(7 * Lp)
    k, t = G // 7, G % 7
    for a in range(NA):
        for j in range(L):
            C = a * L + j
            vals = np.zeros((7 * Lp, 4), np.float32)
            for pp in range(1, 4):
                ok = k < L
                vals[ok, pp] = x2[(3 * t[ok] + pp - 1) * L + k[ok], C]
            hi = bf16_bits(vals)
            if passes == 1:
                ws[a * Lp + j, :28 * Lp] = hi.reshape(-1)
            else:
                lo = bf16_bits(vals - (hi.astype(np.uint32) << 16).view(np.float32))
                inst = np.concatenate([hi.reshape(-1, 32), lo.reshape(-1, 32)], axis=1)
                ws[a * Lp + j, :56 * Lp] = inst.reshape(-1)
    return ws


@pytest.mark.parametrize("passes", [1, 2])
@pytest.mark.parametrize("L,offset", [(1, 0), (5, 0), (37, 1), (64, 0), (130, 0), (250, 0), (250, 1)])
def test_convert_ws_layout(passes, L, offset):
    """ccm_launch_convert_ws against the layout built in numpy, bit for bit:
    L not a multiple of 64 and of the tile widths, x not 16-byte aligned
    (offset 1 float: the scalar loads), rows of padding columns j >= L and
    slots past the operand left alone."""
    import ctypes as C
    from ccmpred_amd import _lib
    lib = _lib.lib()
    vp = C.c_void_p
    lib.ccm_launch_convert_ws.argtypes = [vp, C.c_int, C.c_int, vp, C.c_long, C.c_int, vp]
    lib.ccm_launch_convert_ws.restype = C.c_int
    print(f"convert_ws passes={passes} L={L} offset={offset}", flush=True)
    Lp = (L + 63) // 64 * 64
    nvp = nsingle_padded(L) + L * L * NA * NAP
    rng = np.random.default_rng(L * 10 + passes + offset)
    x = rng.normal(0, 1, nvp).astype(np.float32) * np.float32(3.0) ** rng.integers(-8, 8, nvp).astype(np.float32)
    lds = 28 * Lp * passes + 8
    fill = 0xABCD
    xd = torch.zeros(nvp + offset, dtype=torch.float32, device="cuda")
    xd[offset:] = torch.from_numpy(x).cuda()
    wsd = torch.full((NA * Lp * lds,), fill, dtype=torch.int32, device="cuda").to(torch.int16)
    st = vp(torch.cuda.current_stream().cuda_stream)
    assert lib.ccm_launch_convert_ws(xd[offset:].data_ptr(), L, Lp, wsd.data_ptr(), lds, passes, st) == 0
    torch.cuda.synchronize()
    got = wsd.cpu().numpy().view(np.uint16).reshape(NA * Lp, lds)
    want = ws_expected(x, L, Lp, passes, lds, fill)
    bad = np.argwhere(got != want)
    assert bad.size == 0, (bad[:5], got[tuple(bad[0])], want[tuple(bad[0])])

"""Differentiation form: the HIP adjoint behind a torch.autograd.Function.

Checks at L=120, N=2000 (x: column log-frequencies + N(0, 0.01^2)
symmetric couplings): the backward gradient against FP64 tied central
differences (h = 1e-5) on 500 random coordinates, and against the oracle's
FP64 and FP32 evaluate_cpu gradients.  Then recomputation of the
probabilities in the backward pass against the stored-probabilities variant,
with the device memory each needs.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as ora  # noqa: E402
from ccmpred_amd import pll as P  # noqa: E402
from ccmpred_amd.autograd import pll_objective  # noqa: E402
from ccmpred_amd.msa import synthetic_msa  # noqa: E402
from tests.pll_bruteforce import nsingle_padded, NA, NAP  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def make_x(msa, seed):
    L = msa.shape[1]
    x = ora.init_bias(msa).astype(np.float64)
    nsp = nsingle_padded(L)
    rng = np.random.default_rng(seed)
    W = rng.normal(0, 0.01, (NA, L, NA, L))
    W = W + W.transpose(2, 3, 0, 1)
    W[:, np.arange(L), :, np.arange(L)] = 0
    x[nsp:].reshape(NA, L, NAP, L)[:, :, :NA, :] += W
    return x.astype(np.float32)


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-4), ("bf16x2", 1e-4), ("bf16", 1e-3)])
def test_autograd_against_fd_and_reference(prec, tol):
    L, N = 120, 2000
    say(f"autograd {prec} L={L} N={N}")
    msa, _ = synthetic_msa(L, N, 5, 24)
    x = make_x(msa, 55)
    with P.PLL(msa, precision=prec) as p:
        w64 = p.weights.astype(np.float64)
        xd = torch.from_numpy(x).cuda().requires_grad_(True)
        f = pll_objective(xd, p, recompute=True)
        f.backward()
        g = xd.grad.cpu().numpy().astype(np.float64)
        rng = np.random.default_rng(7)
        ia, ib = ora.random_tied_coordinates(L, 500, rng)
        fd = ora.tied_fd_f64(msa, w64, x.astype(np.float64), p.lambda_single, p.lambda_pair, ia, ib,
                             h=1e-5)
        f64, g64 = ora.evaluate_f64(msa, w64, x.astype(np.float64), p.lambda_single, p.lambda_pair)
        f32, g32 = ora.evaluate_f32(msa, p.weights, x, p.lambda_single, p.lambda_pair)
    scale = np.sqrt(np.mean(fd ** 2))
    err_fd = np.linalg.norm(g[ia] - fd) / np.linalg.norm(fd)
    err_64 = np.linalg.norm(g - g64) / np.linalg.norm(g64)
    err_32 = np.linalg.norm(g - g32) / np.linalg.norm(g32)
    say(f"  f={float(f.detach()):.9g} f64={f64:.9g}  |g-fd|/|fd| over the probes={err_fd:.2e}  "
        f"rel_l2 vs fp64={err_64:.2e} vs fp32 evaluate_cpu={err_32:.2e}")
    say(f"  max|g-fd|/rms(fd)={np.max(np.abs(g[ia] - fd)) / scale:.2e}")
    # the oracle gradient itself matches the FD probes
    assert np.max(np.abs(g64[ia] - fd)) / scale < 1e-6
    assert abs(float(f.detach()) - f64) / abs(f64) < 1e-5
    assert err_64 <= tol and err_32 <= tol
    assert err_fd <= 2 * tol


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_recompute_matches_stored(prec):
    L, N = 120, 5000
    say(f"recompute vs stored {prec} L={L} N={N}")
    msa, _ = synthetic_msa(L, N, 5, 24)
    x = torch.from_numpy(make_x(msa, 56)).cuda()
    # chunks of 2048 sequences (a multiple of the 512-sequence BF16 softmax block)
    with P.PLL(msa, precision=prec) as stored, P.PLL(msa, precision=prec, chunk_rows=2048) as chunked:
        xs = x.clone().requires_grad_(True)
        fs = pll_objective(xs, stored, recompute=False)
        fs.backward()
        xr = x.clone().requires_grad_(True)
        fr = pll_objective(xr, chunked, recompute=True)
        fr.backward()
        fe, ge = stored.evaluate(x)
        say(f"  f stored={float(fs):.10g} recompute={float(fr):.10g} evaluate={fe:.10g} "
            f"bytes stored={stored.device_bytes} chunked={chunked.device_bytes}")
        assert abs(float(fs) - fe) <= 1e-9 * abs(fe)
        assert abs(float(fr) - fe) <= 1e-7 * abs(fe)
        assert torch.equal(xs.grad, ge)
        rel = ((xr.grad - ge).norm() / ge.norm()).item()
        say(f"  grad rel diff chunked vs stored = {rel:.2e}")
        assert rel < 1e-6
        assert chunked.device_bytes < stored.device_bytes


def test_recompute_large_n_memory():
    """L=500, N=50000: chunked recomputation against the unchunked evaluation;
    reports the device memory of both variants."""
    L, N = 500, 50000
    say(f"large-N recompute L={L} N={N}")
    msa, _ = synthetic_msa(L, N, 6, 60)
    x = torch.from_numpy(make_x(msa, 57)).cuda()
    torch.cuda.synchronize()
    with P.PLL(msa, precision="bf16", chunk_rows=8192) as chunked:
        xr = x.clone().requires_grad_(True)
        fr = pll_objective(xr, chunked, recompute=True)
        fr.backward()
        bytes_chunked = chunked.device_bytes
        g_chunk = xr.grad.clone()
    with P.PLL(msa, precision="bf16") as stored:
        xs = x.clone().requires_grad_(True)
        fs = pll_objective(xs, stored, recompute=False)
        fs.backward()
        bytes_stored = stored.device_bytes
    rel = ((g_chunk - xs.grad).norm() / xs.grad.norm()).item()
    say(f"  f chunked={float(fr):.10g} stored={float(fs):.10g} grad rel diff={rel:.2e} "
        f"device bytes chunked={bytes_chunked / 2**30:.2f} GiB stored={bytes_stored / 2**30:.2f} GiB")
    assert abs(float(fr) - float(fs)) <= 1e-6 * abs(float(fs))
    assert rel < 1e-5
    assert bytes_chunked < bytes_stored

"""The single-alignment CG with its line search on the device
(CCM_LINE_SEARCH_DEVICE: the batched form's graph-captured state machine
driving the caller's evaluate_hip instance and x) against the host loop of
ccm_conjugrad_gpu_ex on the same instance: x, f, the objective history and
the counters bit for bit, for both forward forms, a chunk-free and a
chunked instance (which the device path does not take: the host loop runs),
and an FP32 instance (host loop)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from ccmpred_amd import pll as P  # noqa: E402
from ccmpred_amd.msa import synthetic_msa  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.mark.parametrize("prec,sparse,chunk", [("bf16", True, 0), ("bf16f", True, 0), ("bf16", False, 0),
                                              ("bf16", True, 512), ("fp32", True, 0)])
@pytest.mark.parametrize("L,N,seed", [(75, 1000, 1), (130, 700, 41)])
def test_device_line_search_equals_host(prec, sparse, chunk, L, N, seed):
    msa, _ = synthetic_msa(L, N, seed, max(1, L // 12))
    print(f"device CG {prec} sparse={sparse} chunk={chunk} L={L} N={N}", flush=True)
    out = {}
    with P.PLL(msa, precision=prec, sparse=sparse, chunk_rows=chunk) as p:
        for ls in ("device", "host"):
            xd = torch.from_numpy(P.init_bias(msa)).cuda()
            r = p.conjugrad(xd, max_iterations=30, epsilon=0.0, line_search=ls)
            out[ls] = (r, xd.cpu().numpy())
    (rd, xd_), (rh, xh) = out["device"], out["host"]
    print(f"  device ret={rd['ret']} it={rd['iterations']} ev={rd['evaluations']} f={rd['fx']!r} | "
          f"host ret={rh['ret']} it={rh['iterations']} ev={rh['evaluations']} f={rh['fx']!r}", flush=True)
    assert rh["iterations"] == 30
    assert (rd["ret"], rd["iterations"], rd["evaluations"]) == (rh["ret"], rh["iterations"], rh["evaluations"])
    assert rd["fx"] == rh["fx"]
    assert np.array_equal(rd["fhist"], rh["fhist"])
    assert np.array_equal(xd_, xh)


@pytest.mark.parametrize("prec", ["bf16", "bf16f"])
@pytest.mark.parametrize("L,N,seed", [(250, 600, 7), (418, 300, 8)])
def test_device_line_search_banded_gradient(prec, L, N, seed):
    """At L = 250 and 418 the gradient contraction runs as the banded launch
    (256-row, then 192-row tiles): the device CG's batch of one launches it
    from its descriptor, the host loop through evaluate_hip; the two runs
    agree bit for bit."""
    msa, _ = synthetic_msa(L, N, seed, L // 12)
    print(f"device CG, banded gradient: {prec} L={L} N={N}", flush=True)
    out = {}
    with P.PLL(msa, precision=prec) as p:
        for ls in ("device", "host"):
            xd = torch.from_numpy(P.init_bias(msa)).cuda()
            r = p.conjugrad(xd, max_iterations=20, epsilon=0.0, line_search=ls)
            out[ls] = (r, xd.cpu().numpy())
    (rd, xd_), (rh, xh) = out["device"], out["host"]
    print(f"  device ret={rd['ret']} it={rd['iterations']} ev={rd['evaluations']} f={rd['fx']!r} | "
          f"host ret={rh['ret']} it={rh['iterations']} ev={rh['evaluations']} f={rh['fx']!r}", flush=True)
    assert rh["iterations"] == 20
    assert (rd["ret"], rd["iterations"], rd["evaluations"]) == (rh["ret"], rh["iterations"], rh["evaluations"])
    assert rd["fx"] == rh["fx"]
    assert np.array_equal(rd["fhist"], rh["fhist"])
    assert np.array_equal(xd_, xh)


def test_device_line_search_unaligned_x_takes_host_loop():
    """x not 16-byte aligned: the host loop runs and gives the aligned copy's
    results (the work vectors are zeroed and synchronised on every path)."""
    msa, _ = synthetic_msa(60, 400, 42, 5)
    with P.PLL(msa, precision="bf16") as p:
        x0 = P.init_bias(msa)
        buf = torch.zeros(x0.size + 1, dtype=torch.float32, device="cuda")
        buf[1:] = torch.from_numpy(x0).cuda()
        xo = buf[1:]
        assert xo.data_ptr() % 16 != 0
        ro = p.conjugrad(xo, max_iterations=20, epsilon=0.0)
        xa = torch.from_numpy(x0.copy()).cuda()
        ra = p.conjugrad(xa, max_iterations=20, epsilon=0.0, line_search="host")
        print(f"unaligned f={ro['fx']!r} aligned host f={ra['fx']!r}", flush=True)
        assert ro["iterations"] == ra["iterations"] == 20
        assert ro["fx"] == ra["fx"]
        assert np.array_equal(xo.cpu().numpy(), xa.cpu().numpy())

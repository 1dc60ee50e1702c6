"""ccm_launch_contact_map_noalloc (the allocation-free contact map) against
ccm_launch_contact_map with a work buffer, through the C ABI: bit for bit,
on two calls in a row with the second on a smaller alignment (the fixed
scratch keeps the first call's column means; every element the second call
reads it writes first), with and without APC and rescaling; and the size
limit, which is rejected on the host before anything runs."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from ccmpred_amd import _lib  # noqa: E402
from ccmpred_amd import pll as P  # noqa: E402

pytestmark = pytest.mark.gpu

NA, NAP = 21, 32


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def lib():
    return _lib.lib()


def make_x(L, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros(P.nvar_padded(L), np.float32)
    x[:20 * L] = rng.normal(0, 1, 20 * L)
    W = rng.normal(0, 0.1, (NA, L, NA, L)).astype(np.float32)
    W = W + W.transpose(2, 3, 0, 1)
    W[:, np.arange(L), :, np.arange(L)] = 0
    for k in range(0, L - 3, 5):
        W[:, k, :, k + 3] += 0.5
        W[:, k + 3, :, k] += 0.5
    x[P.nsingle_padded(L):].reshape(NA, L, NAP, L)[:, :, :NA, :] = W
    return x


def run_both(L, x, apc, norm):
    lb = lib()
    xd = torch.from_numpy(x).cuda()
    stream = torch.cuda.current_stream().cuda_stream
    out_na = torch.full((L, L), float("nan"), dtype=torch.float32, device="cuda")
    out_ref = torch.full((L, L), float("nan"), dtype=torch.float32, device="cuda")
    work = torch.empty(max(1, lb.ccm_contact_map_work_len(L)), dtype=torch.float32, device="cuda")
    print(f"noalloc contact map L={L} apc={apc} normalize={norm}", flush=True)
    rc = lb.ccm_launch_contact_map_noalloc(xd.data_ptr(), L, apc, norm, out_na.data_ptr(), stream)
    assert rc == 0
    rc = lb.ccm_launch_contact_map(xd.data_ptr(), L, apc, norm, out_ref.data_ptr(), work.data_ptr(),
                                   stream)
    assert rc == 0
    torch.cuda.synchronize()
    return out_na.cpu().numpy(), out_ref.cpu().numpy()


@pytest.mark.parametrize("apc,norm", [(1, 0), (1, 1), (0, 1), (0, 0)])
def test_noalloc_equals_work_buffer_form(apc, norm):
    for L, seed in ((300, 1), (61, 2)):  # second call smaller than the first
        x = make_x(L, seed)
        got, ref = run_both(L, x, apc, norm)
        assert got.tobytes() == ref.tobytes()
        assert got.tobytes() == P.contact_map(x, L, apc=bool(apc), normalize=bool(norm)).tobytes()


def test_noalloc_size_limit():
    lb = lib()
    mx = lb.ccm_contact_map_noalloc_max_l()
    assert mx == 8192
    # rejected before any device work: the pointers are never dereferenced
    dummy = torch.zeros(16, dtype=torch.float32, device="cuda")
    rc = lb.ccm_launch_contact_map_noalloc(dummy.data_ptr(), mx + 1, 1, 0, dummy.data_ptr(),
                                           torch.cuda.current_stream().cuda_stream)
    assert rc == -5
    assert lb.ccm_launch_contact_map_noalloc(None, 10, 1, 0, dummy.data_ptr(), None) == -2

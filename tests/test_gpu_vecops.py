"""The conjugate-gradient vector kernels through the C ABI.

ccm_launch_dots: up to three FP64-accumulated dot products <a, b_i> in one
pass, against numpy FP64 on the same FP32 inputs (rtol 1e-10: the products
are exact in FP64, only the summation order differs).  Sizes cover the tail
after the 16-byte loads (n not a multiple of 4), n below one block, and
vectors offset by one float (not 16-byte aligned: the scalar loop), and the
CG's own call with b1 == a.  ccm_launch_axpy / direction / negate against
numpy FP32 within one rounding (a fused multiply-add may round once).
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from ccmpred_amd import _lib  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.fixture(scope="module")
def L():
    lib = _lib.lib()
    vp = C.c_void_p
    lib.ccm_launch_dots.argtypes = [vp, vp, vp, vp, C.c_long, C.c_int, vp, vp, vp]
    lib.ccm_launch_dots.restype = C.c_int
    lib.ccm_dots_partial_len.restype = C.c_int
    lib.ccm_launch_axpy.argtypes = [vp, vp, C.c_float, C.c_long, vp]
    lib.ccm_launch_direction.argtypes = [vp, vp, C.c_float, C.c_long, vp]
    lib.ccm_launch_negate.argtypes = [vp, vp, C.c_long, vp]
    for f in (lib.ccm_launch_axpy, lib.ccm_launch_direction, lib.ccm_launch_negate):
        f.restype = C.c_int
    return lib


def stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("n", [1, 3, 255, 257, 4099, (1 << 20) + 5])
@pytest.mark.parametrize("offset", [0, 1])
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_dots(L, n, offset, nd):
    rng = np.random.default_rng(n * 7 + offset * 3 + nd)
    host = rng.normal(0, 1, (4, n + offset)).astype(np.float32)
    dev = torch.from_numpy(host).cuda()
    ptr = [dev[i, offset:].data_ptr() for i in range(4)]
    part = torch.empty(L.ccm_dots_partial_len(), dtype=torch.float64, device="cuda")
    out = torch.zeros(3, dtype=torch.float64, device="cuda")
    for same in ((False, True) if nd >= 2 else (False,)):
        b1 = ptr[0] if same else ptr[2]
        rc = L.ccm_launch_dots(ptr[0], ptr[1], b1, ptr[3], n, nd, part.data_ptr(), out.data_ptr(), stream())
        assert rc == 0
        torch.cuda.synchronize()
        a = host[0, offset:].astype(np.float64)
        ref = [a @ host[1, offset:].astype(np.float64),
               a @ (a if same else host[2, offset:].astype(np.float64)),
               a @ host[3, offset:].astype(np.float64)][:nd]
        got = out.cpu().numpy()[:nd]
        assert np.allclose(got, ref, rtol=1e-10, atol=1e-12 * n), (n, offset, nd, same, got, ref)


@pytest.mark.parametrize("n", [7, 4099, (1 << 18) + 3])
def test_elementwise_steps(L, n):
    rng = np.random.default_rng(n)
    x, s, g = (rng.normal(0, 1, n).astype(np.float32) for _ in range(3))
    alpha, beta = np.float32(0.37), np.float32(-1.25)
    xd, sd, gd = (torch.from_numpy(v.copy()).cuda() for v in (x, s, g))
    assert L.ccm_launch_axpy(xd.data_ptr(), sd.data_ptr(), float(alpha), n, stream()) == 0
    assert L.ccm_launch_direction(sd.data_ptr(), gd.data_ptr(), float(beta), n, stream()) == 0
    nd_ = torch.empty_like(gd)
    assert L.ccm_launch_negate(nd_.data_ptr(), gd.data_ptr(), n, stream()) == 0
    torch.cuda.synchronize()
    ulp = lambda v: np.spacing(np.abs(v).astype(np.float32))  # noqa: E731
    xr = (x.astype(np.float64) + np.float64(alpha) * s).astype(np.float32)
    sr = (-g.astype(np.float64) + np.float64(beta) * s).astype(np.float32)
    assert np.all(np.abs(xd.cpu().numpy() - xr) <= ulp(xr))
    assert np.all(np.abs(sd.cpu().numpy() - sr) <= ulp(sr))
    assert np.array_equal(nd_.cpu().numpy(), -g)

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


def _padded_vectors(ncol, seed):
    """Three vectors in the padded layout of ncol columns, zero in the
    a >= 21 padding of the pair block (and the singles padding random)."""
    from ccmpred_amd.pll import nsingle_padded, nvar_padded
    n, nsp = nvar_padded(ncol), nsingle_padded(ncol)
    rng = np.random.default_rng(seed)
    vs = rng.normal(0, 1, (3, n)).astype(np.float32)
    for v in vs:
        v[nsp:].reshape(21, ncol, 32, ncol)[:, :, 21:, :] = 0
    return n, vs


@pytest.mark.parametrize("ncol", [1, 5, 37, 64, 90])
def test_padded_launchers_equal_unpadded(L, ncol):
    """The padded-layout launchers skip the a >= 21 padding: on vectors that
    are zero there, the dots equal ccm_launch_dots bit for bit (same block
    partition, only zero products left out) and the vector steps equal the
    unpadded ones on every entry they visit and leave the padding alone."""
    vp = C.c_void_p
    L.ccm_launch_dots_padded.argtypes = [vp, vp, vp, vp, C.c_long, C.c_int, C.c_int, vp, vp, vp]
    L.ccm_launch_axpy_padded.argtypes = [vp, vp, C.c_float, C.c_int, C.c_long, vp]
    L.ccm_launch_direction_padded.argtypes = [vp, vp, C.c_float, C.c_int, C.c_long, vp]
    L.ccm_launch_negate_padded.argtypes = [vp, vp, C.c_int, C.c_long, vp]
    L.ccm_launch_pad_scan_x.argtypes = [vp, C.c_int, C.c_long, vp, vp]
    for f in (L.ccm_launch_dots_padded, L.ccm_launch_axpy_padded, L.ccm_launch_direction_padded,
              L.ccm_launch_negate_padded, L.ccm_launch_pad_scan_x):
        f.restype = C.c_int
    n, (x, s, g) = _padded_vectors(ncol, ncol + 3)
    print(f"padded launchers ncol={ncol} n={n}", flush=True)
    part = torch.empty(L.ccm_dots_partial_len(), dtype=torch.float64, device="cuda")
    o1 = torch.zeros(3, dtype=torch.float64, device="cuda")
    o2 = torch.zeros(3, dtype=torch.float64, device="cuda")
    xd, sd, gd = (torch.from_numpy(v.copy()).cuda() for v in (x, s, g))
    assert L.ccm_launch_dots(gd.data_ptr(), sd.data_ptr(), gd.data_ptr(), xd.data_ptr(), n, 3,
                             part.data_ptr(), o1.data_ptr(), stream()) == 0
    assert L.ccm_launch_dots_padded(gd.data_ptr(), sd.data_ptr(), gd.data_ptr(), xd.data_ptr(), n, 3, ncol,
                                    part.data_ptr(), o2.data_ptr(), stream()) == 0
    torch.cuda.synchronize()
    assert torch.equal(o1, o2), (o1, o2)

    # the padding of the outputs keeps a marker value: the padded steps must not write it
    mark = np.float32(7.5)
    pad = np.zeros(n, bool)
    from ccmpred_amd.pll import nsingle_padded
    pad[nsingle_padded(ncol):].reshape(21, ncol, 32, ncol)[:, :, 21:, :] = True
    # the float4 holding the last pair entries of a row may also hold padding,
    # which the padded steps visit (and compute as the unpadded ones do)
    visited = ~pad
    n4 = n // 4
    vis4 = visited.reshape(n4, 4).any(1)
    visited = np.repeat(vis4, 4)
    alpha, beta = np.float32(0.37), np.float32(-1.25)
    ref = {}
    for padded in (False, True):
        y, sv = x.copy(), s.copy()
        y[~visited], sv[~visited] = mark, mark
        yd, sdd = torch.from_numpy(y).cuda(), torch.from_numpy(sv).cuda()
        nd_ = torch.full((n,), float(mark), dtype=torch.float32, device="cuda")
        if padded:
            assert L.ccm_launch_axpy_padded(yd.data_ptr(), sdd.data_ptr(), float(alpha), ncol, n, stream()) == 0
            assert L.ccm_launch_direction_padded(sdd.data_ptr(), gd.data_ptr(), float(beta), ncol, n,
                                                 stream()) == 0
            assert L.ccm_launch_negate_padded(nd_.data_ptr(), gd.data_ptr(), ncol, n, stream()) == 0
        else:
            assert L.ccm_launch_axpy(yd.data_ptr(), sdd.data_ptr(), float(alpha), n, stream()) == 0
            assert L.ccm_launch_direction(sdd.data_ptr(), gd.data_ptr(), float(beta), n, stream()) == 0
            assert L.ccm_launch_negate(nd_.data_ptr(), gd.data_ptr(), n, stream()) == 0
        torch.cuda.synchronize()
        ref[padded] = [t.cpu().numpy() for t in (yd, sdd, nd_)]
    for name, a, b in zip(("axpy", "direction", "negate"), ref[False], ref[True]):
        bad = np.flatnonzero(a[visited] != b[visited])
        print(f"  {name}: {bad.size} visited entries differ; first {[(int(k), float(a[visited][k]), float(b[visited][k])) for k in bad[:4]]}",
              flush=True)
        assert bad.size == 0, name
        assert np.all(b[~visited] == mark), name

    # the scan finds a nonzero value in a skipped float4 only
    flag = torch.zeros(1, dtype=torch.int32, device="cuda")
    assert L.ccm_launch_pad_scan_x(xd.data_ptr(), ncol, n, flag.data_ptr(), stream()) == 0
    torch.cuda.synchronize()
    assert int(flag.item()) == 0
    if (~visited).any():
        k = int(np.flatnonzero(~visited)[-1])
        x2 = x.copy()
        x2[k] = 1e-30
        xd2 = torch.from_numpy(x2).cuda()
        assert L.ccm_launch_pad_scan_x(xd2.data_ptr(), ncol, n, flag.data_ptr(), stream()) == 0
        torch.cuda.synchronize()
        assert int(flag.item()) == 1

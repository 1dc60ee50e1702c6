"""The batched CG vector step, ccm_launch_cg_vec_grouped, through the C ABI.

One grouped launch moves several alignments at once, each as its control
block says:

    x += c_pre s;  s = per vec_mode;  x += c_trial s

with vec_mode KEEP (s unchanged), NEGATE (s = -g, gprev = g), DIRECTION
(s = beta s - g, gprev = g) and RESTART (s = -gprev), each multiply-add
rounded once (FP32 fused multiply-add), and no step where vec_on is 0.
Expected values come from numpy: the product of two FP32 values is exact
in FP64, so round-to-FP32 of the FP64 sum is the fused result unless the
FP64 sum itself was rounded onto an FP32 rounding midpoint; those elements
are recomputed in exact rational arithmetic.

Both code paths of the kernel are driven: the padded layout of the
parameter vector (x_pad_nz = 0, 16-byte aligned vectors: only singles and
pair entries a < 21 are visited and the a >= 21 padding keeps its bits)
and the full loop (x_pad_nz = 1, or vectors not 16-byte aligned: every
element moves).  Block counts per alignment range from one block to more
blocks than the vector has float4s.
"""
import ctypes as C
from fractions import Fraction

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from ccmpred_amd import _lib  # noqa: E402
from tests.pll_bruteforce import nsingle_padded, NA, NAP  # noqa: E402

pytestmark = pytest.mark.gpu

KEEP, NEGATE, DIRECTION, RESTART = 0, 1, 2, 3

vp = C.c_void_p


class CgCtl(C.Structure):
    """ccm_cg_ctl (include/ccmpred_amd_batch.h)"""
    _fields_ = [("vec_on", C.c_int32), ("eval_on", C.c_int32), ("vec_mode", C.c_int32), ("done", C.c_int32),
                ("c_pre", C.c_double), ("c_trial", C.c_double), ("beta", C.c_double)] + \
               [(n, C.c_double) for n in ("f", "f0", "alpha", "moved", "lo", "hi", "dg0", "gg")] + \
               [(n, C.c_int32) for n in ("phase", "it", "ls", "steepest", "status", "nevals", "x_pad_nz",
                                         "pad1")] + \
               [("fx", C.c_double), ("dots", C.c_double * 3)]


class BProb(C.Structure):
    """ccm_bprob (include/ccmpred_amd_batch.h)"""
    _fields_ = [(n, C.c_int32) for n in ("L", "Lp", "P", "N", "ncols", "passes_fwd", "passes", "n_fpart",
                                         "n_g1blk", "n_regpair", "n_regsingle", "n_dots")] + \
               [(n, C.c_int64) for n in ("nsp", "nvp", "ldwt", "lddt", "ldmt")] + \
               [("lam1", C.c_float), ("lam2", C.c_float)] + \
               [(n, vp) for n in ("msap", "msapt", "w", "wt", "dt", "zt", "gt", "hsym", "fpart", "g1part",
                                  "regpair", "regsingle", "dpart", "x", "g", "gprev", "s", "ctl", "log")] + \
               [("log_cap", C.c_int32), ("sparse", C.c_int32), ("sidx", vp), ("ldsidx", C.c_int64)]


class Group(C.Structure):
    """ccm_group (include/ccmpred_amd_batch.h)"""
    _fields_ = [("offs", vp), ("pidx", vp), ("np", C.c_int32), ("total", C.c_uint32), ("qmap", vp)]


def test_struct_sizes():
    # the C layouts: 12 int32, 5 int64, 2 float, 19 pointers, 2 int32, pointer, int64
    assert C.sizeof(BProb) == 48 + 40 + 8 + 19 * 8 + 8 + 8 + 8
    assert C.sizeof(CgCtl) == 16 + 3 * 8 + 8 * 8 + 8 * 4 + 8 + 24
    assert C.sizeof(Group) == 32


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.fixture(scope="module")
def L():
    lib = _lib.lib()
    lib.ccm_launch_cg_vec_grouped.argtypes = [vp, Group, vp]
    lib.ccm_launch_cg_vec_grouped.restype = C.c_int
    lib.ccm_cg_vec_blocks.argtypes = [C.c_long]
    lib.ccm_cg_vec_blocks.restype = C.c_int
    return lib


def fma32(a, b, c):
    """round-to-FP32 of a*b + c, computed exactly (a, b, c FP32 arrays)"""
    a, b, c = (np.atleast_1d(np.asarray(v, np.float32)) for v in np.broadcast_arrays(a, b, c))
    s = np.atleast_1d(a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64))
    r = s.astype(np.float32)
    # FP64 values on an FP32 rounding midpoint: low 29 mantissa bits 1000...0
    bits = s.view(np.uint64) & np.uint64((1 << 29) - 1)
    for i in np.flatnonzero((bits == np.uint64(1 << 28)).ravel()):
        idx = np.unravel_index(i, s.shape)
        exact = Fraction(float(a[idx])) * Fraction(float(b[idx])) + Fraction(float(c[idx]))
        lo = np.float32(float(exact))
        # the FP32 neighbours around the exact value, nearest (ties to even) wins
        cands = [np.nextafter(lo, np.float32(-np.inf)), lo, np.nextafter(lo, np.float32(np.inf))]
        dist = [abs(Fraction(float(v)) - exact) for v in cands]
        best = min(dist)
        win = [v for v, d in zip(cands, dist) if d == best]
        r[idx] = win[0] if len(win) == 1 else [v for v in win if (v.view(np.uint32) & 1) == 0][0]
    return r


def step_ref(x, s, g, gp, mode, cpre, ctr, beta):
    """the vector step in numpy: one rounding per multiply-add"""
    x, s, gp = x.copy(), s.copy(), gp.copy()
    if cpre != 0:
        x = fma32(np.float32(cpre), s, x)
    if mode == NEGATE:
        s = -g
    elif mode == DIRECTION:
        s = fma32(np.float32(beta), s, -g)
    elif mode == RESTART:
        s = -gp
    if mode in (NEGATE, DIRECTION):
        gp = g.copy()
    if ctr != 0:
        x = fma32(np.float32(ctr), s, x)
    return x, s, gp


def active_mask(Lc):
    """elements the padded layout visits: singles, then per pair row the
    float4s that hold entries a < 21"""
    nsp = nsingle_padded(Lc)
    row = NAP * Lc
    act = 4 * ((NA * Lc + 3) // 4)
    m = np.zeros(nsp + NA * Lc * row, bool)
    m[:nsp] = True
    m[nsp:].reshape(NA * Lc, row)[:, :act] = True
    return m


# (L, blocks): blocks None = the library's own count for the vector
ALIGNS = [(5, 1), (37, None), (64, 3), (23, 4096)]


def run_case(L, mode_of, cpre_of, ctr_of, pad_nz, offset, vec_on=lambda i: 1, seed=0):
    """one grouped launch over ALIGNS; returns (before, after, expected) per alignment"""
    rng = np.random.default_rng(seed)
    dev, host, probs, ctls, offs = [], [], [], [], [0]
    for i, (Lc, nb) in enumerate(ALIGNS):
        nsp = nsingle_padded(Lc)
        nvp = nsp + Lc * Lc * NA * NAP
        v = rng.normal(0, 1, (4, nvp + offset)).astype(np.float32)
        d = torch.from_numpy(v).cuda()
        dev.append(d)
        host.append(v[:, offset:].copy())
        ptr = [d[k, offset:].data_ptr() for k in range(4)]  # x, s, g, gprev
        c = CgCtl(vec_on=vec_on(i), vec_mode=mode_of(i), c_pre=cpre_of(i), c_trial=ctr_of(i),
                  beta=0.75 - 0.2 * i, x_pad_nz=pad_nz)
        ctls.append(c)
        b = BProb(L=Lc, nsp=nsp, nvp=nvp, x=ptr[0], s=ptr[1], g=ptr[2], gprev=ptr[3])
        probs.append(b)
        offs.append(offs[-1] + (nb if nb is not None else L.ccm_cg_vec_blocks(nvp)))
    ctl_dev = torch.from_numpy(np.frombuffer(bytes((CgCtl * len(ctls))(*ctls)), np.uint8).copy()).cuda()
    for i, b in enumerate(probs):
        b.ctl = ctl_dev.data_ptr() + i * C.sizeof(CgCtl)
    prob_dev = torch.from_numpy(np.frombuffer(bytes((BProb * len(probs))(*probs)), np.uint8).copy()).cuda()
    offs_dev = torch.tensor(offs, dtype=torch.int32).cuda()
    pidx_dev = torch.arange(len(probs), dtype=torch.int32).cuda()
    grp = Group(offs=offs_dev.data_ptr(), pidx=pidx_dev.data_ptr(), np=len(probs), total=offs[-1], qmap=None)
    torch.cuda.synchronize()
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.ccm_launch_cg_vec_grouped(prob_dev.data_ptr(), grp, st) == 0
    torch.cuda.synchronize()
    out = []
    for i, (Lc, _) in enumerate(ALIGNS):
        after = dev[i].cpu().numpy()[:, offset:]
        x, s, g, gp = host[i]
        c = ctls[i]
        if c.vec_on:
            ex, es, egp = step_ref(x, s, g, gp, c.vec_mode, np.float32(c.c_pre), np.float32(c.c_trial),
                                   np.float32(c.beta))
        else:
            ex, es, egp = x, s, gp
        out.append((host[i], after, (ex, es, g, egp)))
    return out


def check(res, padded):
    for i, (before, after, expect) in enumerate(res):
        Lc = ALIGNS[i][0]
        m = active_mask(Lc) if padded else np.ones(before.shape[1], bool)
        for k, name in enumerate(("x", "s", "g", "gprev")):
            want = np.where(m, expect[k], before[k])
            bad = np.flatnonzero(after[k].view(np.uint32) != want.view(np.uint32))
            assert bad.size == 0, (i, Lc, name, bad[:5], after[k][bad[:5]], want[bad[:5]])


MODES = [KEEP, NEGATE, DIRECTION, RESTART]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("path", ["padded", "full", "unaligned"])
def test_vec_step_modes(L, mode, path):
    pad_nz = 1 if path == "full" else 0
    offset = 1 if path == "unaligned" else 0
    print(f"cg_vec_grouped mode={mode} path={path}", flush=True)
    # per alignment: both multiply-adds, only the trial step, only the
    # pre-step, and neither (a pure direction update)
    cpres = [0.125, 0.0, -0.3, 0.0]
    ctrs = [0.5, 1.25, 0.0, 0.0]
    res = run_case(L, lambda i: mode, lambda i: cpres[i], lambda i: ctrs[i], pad_nz, offset,
                   seed=10 * mode + len(path))
    check(res, padded=(path == "padded"))


def test_vec_step_mixed_modes_and_off(L):
    """different modes side by side in one launch; vec_on = 0 leaves one alignment alone"""
    print("cg_vec_grouped mixed modes", flush=True)
    res = run_case(L, lambda i: MODES[i], lambda i: 0.25, lambda i: -0.5, 0, 0,
                   vec_on=lambda i: 0 if i == 2 else 1, seed=99)
    check(res, padded=True)

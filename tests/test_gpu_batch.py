"""Batched form against the single-alignment path.

The grouped kernels run the device code of the single-alignment launchers
(the sparse forward and, with sparse=False, the dense one)
and the device-side CG control makes the host loop's decisions in the same
FP64 arithmetic, so every alignment of a batch must reproduce the
single-alignment BF16 results bit for bit: objective and gradient of one
evaluation, and x, f and the counters after a CG run (also after the
autotuner has changed the tile variants, whose K order is the same, and
with values in the padding of x, which the batched vector step and dot
products skip).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from ccmpred_amd import pll as P  # noqa: E402
from ccmpred_amd.batch import Batch  # noqa: E402
from ccmpred_amd.msa import synthetic_msa  # noqa: E402
from tests.pll_bruteforce import nsingle_padded, NA, NAP  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def start_x(msa, seed):
    """init_bias + symmetric N(0, 0.01^2) couplings, zero self-edges and padding"""
    L = msa.shape[1]
    x = P.init_bias(msa).astype(np.float64)
    nsp = nsingle_padded(L)
    rng = np.random.default_rng(seed)
    W = rng.normal(0, 0.01, (NA, L, NA, L))
    W = W + W.transpose(2, 3, 0, 1)
    W[:, np.arange(L), :, np.arange(L)] = 0
    x[nsp:].reshape(NA, L, NAP, L)[:, :, :NA, :] += W
    return x.astype(np.float32)


# (L, N, seed): uneven shapes, including L not a multiple of 64 and N not a
# multiple of 128
SHAPES = [(37, 300, 11), (64, 1000, 12), (90, 2500, 13), (50, 129, 14), (128, 700, 15)]


def make_set(shapes):
    msas = [synthetic_msa(L, N, seed, max(1, L // 12))[0] for L, N, seed in shapes]
    xs = [start_x(m, s + 100) for m, (_, _, s) in zip(msas, shapes)]
    return msas, xs


@pytest.mark.parametrize("prec", ["bf16", "bf16x2"])
@pytest.mark.parametrize("sparse", [True, False])
def test_batch_evaluate_bitwise(prec, sparse):
    msas, xs = make_set(SHAPES)
    say(f"batch evaluate {prec} sparse={sparse}: shapes {[(m.shape[1], m.shape[0]) for m in msas]}")
    with Batch(msas, precision=prec, sparse=sparse) as b:
        for i, x in enumerate(xs):
            b.set_x(i, x)
        f = b.evaluate()
        gb = [b.get_g(i) for i in range(b.n)]
        # other tile assignments: register-staged and direct-to-LDS tiles
        nv = b.variants()
        b.set_variants(np.full(b.n, 2), np.full(b.n, 0))
        f2 = b.evaluate()
        gb2 = [b.get_g(i) for i in range(b.n)]
        b.set_variants(np.full(b.n, 5), np.full(b.n, 4))
        f3 = b.evaluate()
        gb3 = [b.get_g(i) for i in range(b.n)]
        b.set_variants(*nv)
        for i, (m, x) in enumerate(zip(msas, xs)):
            with P.PLL(m, precision=prec, sparse=sparse) as p:
                assert np.array_equal(p.weights, b.weights[i])
                fs, gs = p.evaluate(torch.from_numpy(x).cuda())
                gs = gs.cpu().numpy()
            say(f"  [{i}] L={m.shape[1]} N={m.shape[0]} f batch={f[i]:.12g} single={fs:.12g} "
                f"max|dg|={np.max(np.abs(gb[i] - gs)):.3g}")
            assert f[i] == fs and f2[i] == fs and f3[i] == fs
            assert np.array_equal(gb[i], gs)
            assert np.array_equal(gb2[i], gs)
            assert np.array_equal(gb3[i], gs)


@pytest.mark.parametrize("sparse", [True, False])
def test_batch_solve_matches_single(sparse):
    shapes = SHAPES[:4]
    msas, xs = make_set(shapes)
    # values in the a >= 21 padding of one start point: the batched vector
    # step skips the padding, the single-alignment one adds a zero step there
    L0 = msas[0].shape[1]
    pad = xs[0][nsingle_padded(L0):].reshape(NA, L0, NAP, L0)
    pad[:, :, NA:, :] = np.random.default_rng(7).normal(0, 1, (NA, L0, NAP - NA, L0)).astype(np.float32)
    its = 12
    say(f"batch CG {its} iterations sparse={sparse}: shapes {[(m.shape[1], m.shape[0]) for m in msas]}")
    with Batch(msas, precision="bf16", sparse=sparse) as b:
        for i, x in enumerate(xs):
            b.set_x(i, x)
        rounds = b.solve(max_iterations=its, epsilon=0.0, rounds_per_graph=3)
        res = [b.result(i) for i in range(b.n)]
        xb = [b.get_x(i) for i in range(b.n)]
        # contact maps formed on the device equal the host maps of the read-back x
        for i in range(b.n):
            L = msas[i].shape[1]
            for apc, norm in ((True, False), (False, False), (True, True)):
                assert np.array_equal(b.contact_map(i, apc=apc, normalize=norm),
                                      P.contact_map(xb[i], L, apc=apc, normalize=norm)), (i, apc, norm)
    say(f"  rounds={rounds}")
    for i, (m, x) in enumerate(zip(msas, xs)):
        with P.PLL(m, precision="bf16", sparse=sparse) as p:
            xd = torch.from_numpy(x).cuda()
            r = p.conjugrad(xd, max_iterations=its, epsilon=0.0, line_search="host")
            xs_ = xd.cpu().numpy()
        rb = res[i]
        say(f"  [{i}] single ret={r['ret']} it={r['iterations']} ev={r['evaluations']} f={r['fx']:.10g} | "
            f"batch ret={rb['ret']} it={rb['iterations']} ev={rb['evaluations']} f={rb['fx']:.10g} "
            f"max|dx|={np.max(np.abs(xb[i] - xs_)):.3g}")
        assert (rb["ret"], rb["iterations"], rb["evaluations"]) == (r["ret"], r["iterations"], r["evaluations"])
        assert rb["fx"] == r["fx"]
        assert np.array_equal(rb["log"][:, 0], r["fhist"])
        assert np.array_equal(xb[i], xs_)


def test_batch_autotune_then_solve():
    shapes = [(60, 800, 21), (130, 3000, 22), (75, 200, 23), (200, 1500, 24), (45, 5000, 25),
              (150, 400, 26)]
    msas, xs = make_set(shapes)
    its = 6
    say(f"batch autotune + CG {its}: shapes {[(m.shape[1], m.shape[0]) for m in msas]}")
    with Batch(msas, precision="bf16") as b:
        before = b.variants()
        ms = b.autotune(reps=3)
        after = b.variants()
        say(f"  variants before fwd={before[0].tolist()} grad={before[1].tolist()}")
        say(f"  variants after  fwd={after[0].tolist()} grad={after[1].tolist()}  tuned ms={ms}")
        for i, x in enumerate(xs):
            b.set_x(i, x)
        b.solve(max_iterations=its, epsilon=0.0, rounds_per_graph=4)
        res = [b.result(i) for i in range(b.n)]
        xb = [b.get_x(i) for i in range(b.n)]
    for i, (m, x) in enumerate(zip(msas, xs)):
        with P.PLL(m, precision="bf16") as p:
            xd = torch.from_numpy(x).cuda()
            r = p.conjugrad(xd, max_iterations=its, epsilon=0.0, line_search="host")
            xs_ = xd.cpu().numpy()
        say(f"  [{i}] it={res[i]['iterations']}/{r['iterations']} f={res[i]['fx']:.10g}/{r['fx']:.10g}")
        assert res[i]["fx"] == r["fx"] and res[i]["evaluations"] == r["evaluations"]
        assert np.array_equal(xb[i], xs_)

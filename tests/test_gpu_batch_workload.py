"""The batched workload itself: 64 synthetic MSAs, L uniform in [50, 600],
N log-uniform in [200, 20000] (ccmpred_amd.msa.batched_workload), solved
together for 50 CG iterations (epsilon 0) from init_bias in the single-pass
bf16 and the two-pass bf16f form, with autotuned tiles and graph-captured
rounds.

Checked:
  * every alignment runs all 50 iterations;
  * every alignment's x, f and counters equal the single-alignment device
    CG's (ccm_conjugrad_gpu_ex over evaluate_hip, line search on the host)
    bit for bit;
  * f and g of one grouped evaluation against the FP64 oracle
    (evaluate_cpu.c:16-166 restated) for the three smallest alignments: at
    init_bias + symmetric N(0, 0.01^2) couplings before the solve (f 1e-5,
    g 1e-3 relative L2), and f at the solved x (1e-5 bf16f, 1e-4 bf16; near
    the optimum g itself is small, so its relative error there measures
    nothing);
  * the contact map of alignment 40 (L=418, N=348) against the FP64-oracle
    CG map (tests/golden/cg_maps/batch64_40.npy) under the criteria of
    tests/cg_fixtures.py.
"""
import time

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as ora  # noqa: E402
from ccmpred_amd import pll as P  # noqa: E402
from ccmpred_amd.batch import Batch  # noqa: E402
from ccmpred_amd.msa import batched_workload  # noqa: E402
from tests import cg_fixtures as F  # noqa: E402

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
    rng = np.random.default_rng(seed)
    W = rng.normal(0, 0.01, (21, L, 21, L))
    W = W + W.transpose(2, 3, 0, 1)
    W[:, np.arange(L), :, np.arange(L)] = 0
    x[P.nsingle_padded(L):].reshape(21, L, 32, L)[:, :, :21, :] += W
    return x.astype(np.float32)


def fp64(msa, x):
    L = msa.shape[1]
    w = ora.calculate_weights(msa, 0.8).astype(np.float64)
    return ora.evaluate_f64_par(msa, w, x.astype(np.float64), 0.01, float(np.float32(0.2 * (L - 1))))


@pytest.mark.parametrize("prec", ["bf16", "bf16f"])
def test_batched_workload_64(prec):
    msas = batched_workload(64)
    say(f"batched workload: {len(msas)} MSAs, {prec}")
    work = [m.shape[0] * m.shape[1] ** 2 for m in msas]
    small = [int(i) for i in np.argsort(work)[:3]]
    with Batch(msas, precision=prec) as b:
        t0 = time.perf_counter()
        b.autotune(reps=3)
        # one grouped evaluation at a non-optimal x, three alignments against FP64
        for i, m in enumerate(msas):
            b.set_x(i, start_x(m, 100 + i) if i in small else P.init_bias(m))
        fb = b.evaluate()
        for i in small:
            f64, g64 = fp64(msas[i], b.get_x(i))
            ef = abs(fb[i] - f64) / abs(f64)
            eg = np.linalg.norm(b.get_g(i).astype(np.float64) - g64) / np.linalg.norm(g64)
            say(f"  start x, alignment {i} (L={msas[i].shape[1]}, N={msas[i].shape[0]}): "
                f"f rel err {ef:.2e}, g rel L2 {eg:.2e}")
            assert ef <= 1e-5 and eg <= 1e-3
        for i, m in enumerate(msas):
            b.set_x(i, P.init_bias(m))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rounds = b.solve(max_iterations=50, epsilon=0.0, rounds_per_graph=8)
        t2 = time.perf_counter()
        res = [b.result(i) for i in range(b.n)]
        say(f"  autotune {t1 - t0:.1f} s, solve {t2 - t1:.2f} s ({rounds} rounds)")
        its = [r["iterations"] for r in res]
        st = sorted({r["ret"] for r in res})
        say(f"  iterations min {min(its)} max {max(its)}; status {st}")
        assert all(i == 50 for i in its), [(k, r["ret"], r["iterations"]) for k, r in enumerate(res)
                                           if r["iterations"] != 50]

        # alignment 40 against its FP64-oracle CG map
        meta, m64 = F.load("batch64_40")
        assert F.msa_sha256(msas[40]) == meta["msa_sha256"]
        c = F.compare(b.contact_map(40), m64)
        say(f"  alignment 40: f={res[40]['fx']:.10g} (FP64 CG {meta['f']:.10g}) "
            f"spearman={c['rho']:.6f} top-L/5 identical={c['identical']} swapped={c['swapped']}")
        assert c["rho"] >= 0.999
        assert c["identical"] or c["swaps_are_near_ties"]

        # the objective at the solved x against FP64: the single-pass form
        # rounds the larger solved couplings to BF16 in its logits (measured
        # 2.9e-5 on alignment 33), so it has the 1e-4 of the tiny-shape tests
        fb = b.evaluate()
        for i in small:
            f64, _ = fp64(msas[i], b.get_x(i))
            ef = abs(fb[i] - f64) / abs(f64)
            say(f"  solved x, alignment {i}: f rel err {ef:.2e}")
            assert ef <= (1e-4 if prec == "bf16" else 1e-5)

        # the same solves one alignment at a time: bit for bit
        same = 0
        for i, m in enumerate(msas):
            with P.PLL(m, precision=prec) as p:
                xd = torch.from_numpy(P.init_bias(m)).cuda()
                r = p.conjugrad(xd, max_iterations=50, epsilon=0.0, line_search="host")
                xs = xd.cpu().numpy()
            ok = (r["fx"] == res[i]["fx"] and r["iterations"] == res[i]["iterations"]
                  and r["evaluations"] == res[i]["evaluations"] and r["ret"] == res[i]["ret"]
                  and np.array_equal(xs, b.get_x(i)))
            if not ok:
                say(f"  alignment {i}: batched f={res[i]['fx']!r} it={res[i]['iterations']} "
                    f"evals={res[i]['evaluations']} | single f={r['fx']!r} it={r['iterations']} "
                    f"evals={r['evaluations']}")
            same += int(ok)
        say(f"  batched == single-alignment CG bit for bit: {same}/{len(msas)}")
        assert same == len(msas)

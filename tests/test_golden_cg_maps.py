"""The FP64-oracle CG contact maps under tests/golden/cg_maps/ (written by
tests/golden/make_cg_maps.py) against the generators and the oracle here:
each case's alignment is still the one the map was made from (checksum),
the recorded top-L/5 pairs and cut-off follow from the stored map, every run
took all 50 iterations, the maps recover the planted contacts, and the
parallel FP64 driver that made them (ora_cg_pll_f64par) takes the same CG
path as the serial FP64 one on a small alignment."""
import numpy as np
import pytest

from oracle import oracle as ora
from ccmpred_amd.msa import batched_workload, read_psicov, synthetic_msa
from tests import cg_fixtures as F

SYNTH = {"L75_N1000": (75, 1000, 1, 15), "L250_N5000": (250, 5000, 2, 60),
         "L600_N20000": (600, 20000, 4, 60)}


def cases():
    import json
    with open(F.JSON) as fh:
        return sorted(json.load(fh))


@pytest.mark.parametrize("name", cases())
def test_fixture_consistent(name):
    meta, m = F.load(name)
    L = meta["L"]
    assert m.shape == (L, L) and m.dtype == np.float32
    assert np.all(np.diag(m) == 0) and np.allclose(m, m.T, rtol=1e-5, atol=1e-6)
    if name in SYNTH:
        msa, pairs = synthetic_msa(*SYNTH[name])
        # at L/5 <= planted pairs every top-L/5 pair is a planted contact
        planted = set(map(tuple, pairs.tolist()))
        top = set(F.top_pairs(m, L // 5))
        assert len(top & planted) == min(L // 5, len(planted))
    elif name == "1atzA":
        msa = read_psicov(F.GOLDEN + "/1atzA.aln")
    else:
        msa = batched_workload(64)[int(name.rsplit("_", 1)[1])]
    assert F.msa_sha256(msa) == meta["msa_sha256"]
    assert msa.shape == (meta["N"], L)
    assert meta["iterations"] == 50 and meta["ret"] == 2  # max iterations reached
    assert [tuple(p) for p in meta["top_l5"]] == F.top_pairs(m, L // 5)
    assert [tuple(p) for p in meta["top_l5_sep7"]] == F.top_pairs(m, L // 5, 7)
    assert np.isclose(meta["cutoff_score"], F.compare(m, m)["cut"], rtol=0, atol=0)
    assert all(a >= b for a, b in zip(meta["fhist"], meta["fhist"][1:]))


def test_parallel_fp64_driver_same_path():
    msa, _ = synthetic_msa(20, 150, 9, 3)
    w = ora.calculate_weights(msa, 0.8)
    x0 = ora.init_bias(msa)
    a = ora.cg_pll_f32(msa, w, x0, 0.01, np.float32(0.2 * 19), max_iterations=20, evaluator="f64")
    b = ora.cg_pll_f32(msa, w, x0, 0.01, np.float32(0.2 * 19), max_iterations=20, evaluator="f64par")
    assert a["iterations"] == b["iterations"] == 20
    assert a["evaluations"] == b["evaluations"]
    np.testing.assert_allclose(a["fhist"], b["fhist"], rtol=1e-10)
    np.testing.assert_allclose(a["x"], b["x"], rtol=1e-4, atol=1e-6)

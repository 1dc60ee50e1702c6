"""The C-ABI library loads and exports every symbol the public headers
declare (no device calls: runs without a GPU)."""
import os
import re

import numpy as np

from ccmpred_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^(?!typedef)[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", text,
                             flags=re.M):
            names.add(m.group(1))
    # function-pointer typedefs are declarations of types, not symbols
    return {n for n in names if not n.endswith("_t")}


def test_all_declared_symbols_exported():
    lib = _lib.lib()
    declared = _declared()
    assert len(declared) > 40
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    assert declared <= set(_lib.EXPORTED)


def test_host_helpers_match_layout():
    from ccmpred_amd import pll
    for L in (1, 8, 75, 250):
        ns = 20 * L
        assert pll.nsingle_padded(L) == ns + 32 - ns % 32
        assert pll.nvar_padded(L) == pll.nsingle_padded(L) + 21 * 32 * L * L


def test_host_contact_map_matches_oracle():
    from oracle import oracle as ora
    from ccmpred_amd import pll
    rng = np.random.default_rng(0)
    L = 9
    x = rng.normal(0, 0.3, pll.nvar_padded(L)).astype(np.float32)
    for apc in (False, True):
        got = pll.contact_map(x, L, apc=apc)
        ref = ora.sum_submatrices(x, L)
        if apc:
            ref = ora.apc(ref)
        assert got.tobytes() == ref.tobytes()
    got = pll.contact_map(x, L, apc=True, normalize=True)
    ref = ora.normalize(ora.apc(ora.sum_submatrices(x, L)))
    assert got.tobytes() == ref.tobytes()


def test_host_init_bias_matches_oracle():
    from oracle import oracle as ora
    from ccmpred_amd import pll
    from ccmpred_amd.msa import synthetic_msa
    msa, _ = synthetic_msa(30, 100, seed=8, n_pairs=3)
    assert pll.init_bias(msa).tobytes() == ora.init_bias(msa).tobytes()


def test_psicov_reader(tmp_path):
    from ccmpred_amd.msa import read_psicov, encode
    p = tmp_path / "a.aln"
    p.write_text("ARND-\nXBZ\nacdef  \n")
    m = read_psicov(p)
    assert m.shape == (3, 5)
    np.testing.assert_array_equal(m, encode(["ARND-", "XBZ--", "ACDEF"]))

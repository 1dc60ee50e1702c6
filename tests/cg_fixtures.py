"""FP64-oracle CG contact maps (tests/golden/cg_maps/, written by
tests/golden/make_cg_maps.py) and the criteria a device run is held to.

A device CG run of 50 iterations from init_bias (epsilon 0) is compared with
the FP64-evaluated oracle CG of the same algorithm on the same alignment:
  * APC map Spearman correlation over i < j >= 0.999;
  * identical top-L/5 pairs at separation j - i >= 6 (ties by (i, j)), and
    at j - i >= 7 (the separation of the reference's top_couplings.py), where
    the only swaps allowed are between pairs whose FP64 scores lie within
    1 % of the FP64 cut-off score (the L/5-th), i.e. pairs the FP64 map
    itself does not separate by more than 1 %.
"""
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
JSON = os.path.join(GOLDEN, "cg_maps.json")
TIE_REL = 0.01


def msa_sha256(msa):
    return hashlib.sha256(np.ascontiguousarray(msa, dtype=np.uint8).tobytes()).hexdigest()


def has(name):
    """whether the fixture of case name has been written"""
    if not os.path.exists(JSON):
        return False
    with open(JSON) as fh:
        return name in json.load(fh) and os.path.exists(os.path.join(GOLDEN, "cg_maps", f"{name}.npy"))


def load(name):
    with open(JSON) as fh:
        meta = json.load(fh)[name]
    return meta, np.load(os.path.join(GOLDEN, "cg_maps", f"{name}.npy"))


def top_pairs(mat, num, min_sep=6):
    L = mat.shape[0]
    ii, jj = np.triu_indices(L, k=min_sep)
    vals = mat[ii, jj]
    order = np.lexsort((jj, ii, -vals))[:num]
    return [(int(ii[o]), int(jj[o])) for o in order]


def spearman(a, b):
    iu = np.triu_indices(a.shape[0], k=1)
    ra = np.argsort(np.argsort(a[iu], kind="stable"), kind="stable").astype(np.float64)
    rb = np.argsort(np.argsort(b[iu], kind="stable"), kind="stable").astype(np.float64)
    return float(np.corrcoef(ra, rb)[0, 1])


def compare(m_dev, m64, min_sep=6):
    """Spearman, top-L/5 sets at separation j - i >= min_sep and the FP64
    near-tie test of the swaps."""
    L = m64.shape[0]
    k = L // 5
    t_dev, t64 = set(top_pairs(m_dev, k, min_sep)), set(top_pairs(m64, k, min_sep))
    iu = np.triu_indices(L, k=min_sep)
    cut = float(np.sort(m64[iu])[::-1][k - 1])
    swapped = sorted(t_dev ^ t64)
    rel = [abs(float(m64[p]) - cut) / abs(cut) for p in swapped]
    near_band = int(np.sum(np.abs(m64[iu] - cut) <= TIE_REL * abs(cut)))
    return dict(rho=spearman(m_dev, m64), k=k, identical=not swapped, swapped=swapped,
                swap_rel=rel, swaps_are_near_ties=all(r <= TIE_REL for r in rel),
                fp64_pairs_in_band=near_band, cut=cut)

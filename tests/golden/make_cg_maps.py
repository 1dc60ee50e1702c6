#!/usr/bin/env python
"""Writes the FP64-oracle conjugate-gradient contact maps under
tests/golden/cg_maps/ (run in the build container; takes minutes, hours
with --large).

For each case the oracle's host CG (oracle/ccm_oracle.c:ora_cg, the
algorithm of the device CG) runs from init_bias (ccmpred.c:90-131) for 50
iterations with epsilon 0 and the reference's other CG parameters
(ccmpred.c:526-534), every objective and gradient evaluated in FP64
(ora_evaluate_f64_par, a restatement of evaluate_cpu.c:16-166), with the
oracle's reweighting (reweighting.c:9-58).  The APC contact map
(util.c:16-108) of the final x is stored as <case>.npy (float32 L x L),
and cg_maps.json records the run: status, iterations, evaluations, the
objective history, the top-L/5 pairs (j - i >= 6, ties by (i, j)) and the
score margin around the top-L/5 cut-off.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)

from ccmpred_amd.msa import batched_workload, read_psicov, synthetic_msa  # noqa: E402
from oracle import oracle as ora  # noqa: E402

OUT = os.path.join(HERE, "cg_maps")
JSON = os.path.join(HERE, "cg_maps.json")


def lambda_pair_cli(L):
    """The command-line driver's float product 0.2f * (float)(L - 1)."""
    return float(np.float32(np.float32(0.2) * np.float32(L - 1)))


def cases(large):
    out = {
        "L75_N1000": (lambda: synthetic_msa(75, 1000, 1, 15)[0], None),
        "L250_N5000": (lambda: synthetic_msa(250, 5000, 2, 60)[0], None),
        "1atzA": (lambda: read_psicov(os.path.join(HERE, "1atzA.aln")), lambda_pair_cli),
        "batch64_40": (lambda: batched_workload(64)[40], None),
    }
    if large:
        out = {"L600_N20000": (lambda: synthetic_msa(600, 20000, 4, 60)[0], None)}
    return out


def msa_sha256(msa):
    return hashlib.sha256(np.ascontiguousarray(msa, dtype=np.uint8).tobytes()).hexdigest()


def top_pairs(mat, num, min_sep=6):
    L = mat.shape[0]
    ii, jj = np.triu_indices(L, k=min_sep)
    vals = mat[ii, jj]
    order = np.lexsort((jj, ii, -vals))
    return [(int(ii[o]), int(jj[o])) for o in order[:num]], vals[order]


def run(name, make, lam2_fn):
    msa = make()
    N, L = msa.shape
    w = ora.calculate_weights(msa, 0.8)
    lam1 = 0.01
    lam2 = lam2_fn(L) if lam2_fn else float(np.float32(0.2 * (L - 1)))
    x0 = ora.init_bias(msa)
    t0 = time.time()
    r = ora.cg_pll_f32(msa, w, x0, lam1, lam2, max_iterations=50, epsilon=0.0, evaluator="f64par")
    dt = time.time() - t0
    m = ora.apc(ora.sum_submatrices(r["x"], L))
    np.save(os.path.join(OUT, f"{name}.npy"), m.astype(np.float32))
    k = L // 5
    top, sorted_vals = top_pairs(m, k)
    top7, _ = top_pairs(m, k, 7)
    rec = dict(L=int(L), N=int(N), msa_sha256=msa_sha256(msa), lambda_single=lam1, lambda_pair=lam2, ret=int(r["ret"]),
               iterations=int(r["iterations"]), evaluations=int(r["evaluations"]),
               f=float(r["fx"]), fhist=[float(v) for v in r["fhist"]], top_l5=top, top_l5_sep7=top7,
               cutoff_score=float(sorted_vals[k - 1]), next_score=float(sorted_vals[k]),
               seconds=round(dt, 1))
    print(f"{name}: ret={rec['ret']} it={rec['iterations']} evals={rec['evaluations']} "
          f"f={rec['f']:.10g} cut={rec['cutoff_score']:.6g} next={rec['next_score']:.6g} {dt:.0f}s",
          flush=True)
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--large", action="store_true", help="the L=600, N=20000 case only")
    ap.add_argument("--only", default=None)
    ap.add_argument("--annotate", action="store_true",
                    help="only add the alignments' checksums to existing entries")
    args = ap.parse_args()
    os.makedirs(OUT, exist_ok=True)
    meta = json.load(open(JSON)) if os.path.exists(JSON) else {}
    if args.annotate:
        for name, (make, _) in {**cases(False), **cases(True)}.items():
            if name in meta:
                meta[name]["msa_sha256"] = msa_sha256(make())
        with open(JSON, "w") as fh:
            json.dump(meta, fh, indent=1, sort_keys=True)
        return
    for name, (make, lam2_fn) in cases(args.large).items():
        if args.only and name != args.only:
            continue
        meta[name] = run(name, make, lam2_fn)
        with open(JSON, "w") as fh:
            json.dump(meta, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()

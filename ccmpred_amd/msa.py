"""MSA encoding, PSICOV reading and the seeded synthetic MSA generator.

Alphabet and encoding follow the reference's ``sequence.c:10-27``: residues
are indexed in the order ``ARNDCQEGHILKMFPSTWYV-``; the gap and every
non-standard letter (B, J, O, U, X, Z), as well as any non-letter, map to 20.
"""
from __future__ import annotations

import numpy as np

N_ALPHA = 21
GAP = 20
ALPHABET = "ARNDCQEGHILKMFPSTWYV-"

# sequence.c:10-12, letter A..Z -> index
_AMINO_INDICES = np.array(
    [0, 20, 4, 3, 6, 13, 7, 8, 9, 20, 11, 10, 12, 2, 20, 14, 5, 1, 15, 16, 20, 19, 17, 20, 18, 20],
    dtype=np.uint8,
)


def _build_byte_table() -> np.ndarray:
    table = np.full(256, GAP, dtype=np.uint8)
    for off in range(26):
        table[ord("A") + off] = _AMINO_INDICES[off]
        table[ord("a") + off] = _AMINO_INDICES[off]
    return table


BYTE_TO_INDEX = _build_byte_table()


def encode(seqs) -> np.ndarray:
    """Encode equal-length strings/bytes into an N x L uint8 index matrix."""
    rows = [np.frombuffer(s.encode() if isinstance(s, str) else bytes(s), dtype=np.uint8) for s in seqs]
    if not rows:
        return np.zeros((0, 0), dtype=np.uint8)
    ncol = max(len(r) for r in rows)
    out = np.full((len(rows), ncol), GAP, dtype=np.uint8)
    for i, r in enumerate(rows):
        out[i, : len(r)] = BYTE_TO_INDEX[r]
    return out


def decode(msa: np.ndarray) -> list[str]:
    """Inverse of :func:`encode` for indices 0..20."""
    table = np.frombuffer(ALPHABET.encode(), dtype=np.uint8)
    return [table[row].tobytes().decode() for row in np.asarray(msa, dtype=np.uint8)]


def read_psicov(path) -> np.ndarray:
    """Read a PSICOV alignment (one sequence per line).

    Trailing whitespace is stripped and short lines are gap-filled; a FASTA
    or a3m header (``>``) is rejected as in ``sequence.c:48-80``.
    """
    with open(path, "rb") as fh:
        lines = [ln.rstrip() for ln in fh.read().split(b"\n")]
    while lines and lines[-1] == b"":
        lines.pop()
    for ln in lines:
        if b">" in ln:
            raise ValueError("alignment looks like a3m/FASTA; PSICOV format is required")
    return encode(lines)


def planted_pair_count(ncol: int) -> int:
    """Planted contacts for a generated MSA of ``ncol`` columns (15 at L=75, 60 at L>=250)."""
    if ncol <= 75:
        return max(1, ncol // 5)
    return min(60, max(15, ncol // 5))


def synthetic_msa(ncol: int, nrow: int, seed: int, n_pairs: int | None = None,
                  coupling_prob: float = 0.7, min_sep: int = 6):
    """Seeded synthetic MSA with planted contacts.

    Per column j a residue profile p_j ~ Dirichlet(0.5 * 1_20) and a gap rate
    g_j ~ U[0, 0.1] are drawn; x_nj is a gap with probability g_j, else a draw
    from p_j.  ``n_pairs`` column pairs (i, j) with j - i >= ``min_sep`` and no
    shared column are planted: with probability ``coupling_prob`` and x_ni not
    a gap, x_nj is replaced by sigma(x_ni) for a fixed random permutation sigma
    of the 20 residues.  Uses numpy's PCG64 generator.

    Returns (msa uint8 [nrow, ncol], pairs int array [n_pairs, 2]).
    """
    if n_pairs is None:
        n_pairs = planted_pair_count(ncol)
    rng = np.random.Generator(np.random.PCG64(seed))
    profiles = rng.dirichlet(np.full(20, 0.5), size=ncol)
    gap_rate = rng.uniform(0.0, 0.1, size=ncol)
    cum = np.cumsum(profiles, axis=1)
    cum[:, -1] = 1.0
    u = rng.random((nrow, ncol))
    res = np.empty((nrow, ncol), dtype=np.uint8)
    for j in range(ncol):
        res[:, j] = np.minimum(np.searchsorted(cum[j], u[:, j], side="right"), 19)
    gaps = rng.random((nrow, ncol)) < gap_rate[None, :]
    msa = np.where(gaps, np.uint8(GAP), res).astype(np.uint8)

    pairs = []
    used = np.zeros(ncol, dtype=bool)
    tries = 0
    while len(pairs) < n_pairs and tries < 100000:
        tries += 1
        i, j = sorted(rng.integers(0, ncol, size=2).tolist())
        if j - i < min_sep or used[i] or used[j]:
            continue
        used[i] = used[j] = True
        pairs.append((i, j))
    pairs = np.array(sorted(pairs), dtype=np.int64).reshape(-1, 2)

    sigma = rng.permutation(20).astype(np.uint8)
    for i, j in pairs:
        hit = (rng.random(nrow) < coupling_prob) & (msa[:, i] != GAP)
        msa[hit, j] = sigma[msa[hit, i]]
    return msa, pairs


# Workloads of the PLL path: (ncol, nrow, seed, planted pairs)
WORKLOADS = {
    "L75_N1000": (75, 1000, 1, 15),
    "L250_N5000": (250, 5000, 2, 60),
    "L250_N5000_bf16": (250, 5000, 3, 60),
    "L600_N20000": (600, 20000, 4, 60),
    "L120_N2000": (120, 2000, 5, 24),
    "L500_N50000": (500, 50000, 6, 60),
}


def workload_msa(name: str):
    ncol, nrow, seed, npairs = WORKLOADS[name]
    return synthetic_msa(ncol, nrow, seed, npairs)


def batched_workload(nmsa: int = 64, seed: int = 999):
    """The batched workload: ``nmsa`` synthetic MSAs with L uniform in
    [50, 600] and N log-uniform in [200, 20000]; MSA i is generated with seed
    1000 + i and L // 4 planted pairs."""
    rng = np.random.Generator(np.random.PCG64(seed))
    Ls = rng.integers(50, 601, nmsa)
    Ns = np.round(np.exp(rng.uniform(np.log(200), np.log(20000), nmsa))).astype(int)
    return [synthetic_msa(int(L), int(N), 1000 + i, max(1, int(L) // 4))[0]
            for i, (L, N) in enumerate(zip(Ls, Ns))]

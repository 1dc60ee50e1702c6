"""Brute-force FP64 pseudo-likelihood, written from the model definition.

Used only to pin the oracle on tiny alignments.  It works on dense one-hot
tensors with an explicit, max-shifted softmax, so it shares no code or loop
structure with the oracle it checks.
"""
import numpy as np

NA, NAP = 21, 32


def nsingle_padded(L):
    ns = 20 * L
    return ns + NAP - ns % NAP


def unpack(x, L):
    nsp = nsingle_padded(L)
    V = x[: 20 * L].reshape(20, L)                   # V[a, j]
    X2 = x[nsp:].reshape(NA, L, NAP, L)              # [b, k, a(padded), j]
    return V, X2


def lse(msa, x):
    """log sum_a exp(logit(n, j, a)) per sequence and column, [n, j]."""
    msa = np.asarray(msa)
    N, L = msa.shape
    V, X2 = unpack(np.asarray(x, dtype=np.float64), L)
    onehot = np.eye(NA)[msa]
    logits = np.einsum("nkb,bkaj->nja", onehot, X2[:, :, :NA, :])
    logits[:, :, :20] += V.T[None, :, :]
    m = logits.max(axis=2, keepdims=True)
    return (np.log(np.exp(logits - m).sum(axis=2, keepdims=True)) + m)[:, :, 0]


def pll(msa, w, x, lam1, lam2):
    """Returns (f, tied gradient in the padded layout)."""
    msa = np.asarray(msa)
    N, L = msa.shape
    x = np.asarray(x, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    V, X2 = unpack(x, L)
    W = X2[:, :, :NA, :]                              # [b, k, a, j]
    onehot = np.eye(NA)[msa]                          # [n, k, b]
    logits = np.einsum("nkb,bkaj->nja", onehot, W)   # [n, j, a]
    logits[:, :, :20] += V.T[None, :, :]
    m = logits.max(axis=2, keepdims=True)
    e = np.exp(logits - m)
    Z = e.sum(axis=2, keepdims=True)
    P = e / Z
    lse = (np.log(Z) + m)[:, :, 0]
    obs = np.take_along_axis(logits, msa[:, :, None], axis=2)[:, :, 0]
    nll = (w[:, None] * (lse - obs)).sum()
    nsp = nsingle_padded(L)
    reg = lam1 * np.sum(x[: 20 * L] ** 2) + 0.5 * lam2 * np.sum(x[nsp:] ** 2)
    f = nll + reg

    D = w[:, None, None] * (P - onehot)               # dF/dlogits [n, j, a]
    g = np.zeros_like(x)
    g[: 20 * L] = (D[:, :, :20].sum(axis=0).T + 2 * lam1 * V).ravel()
    raw = np.einsum("nkb,nja->bkaj", onehot, D)      # dF/dW without L2
    tied = raw + raw.transpose(2, 3, 0, 1)
    for k in range(L):
        tied[:, k, :, k] = 0.0
    G2 = np.zeros((NA, L, NAP, L))
    G2[:, :, :NA, :] = tied
    g[nsp:] = G2.ravel() + 2 * lam2 * x[nsp:]
    return f, g


def random_symmetric_x(L, rng, scale=0.1, single_scale=0.5):
    """Random x with symmetric couplings, zero self-edge blocks and zero padding."""
    nsp = nsingle_padded(L)
    x = np.zeros(nsp + NA * L * NAP * L)
    x[: 20 * L] = rng.normal(0, single_scale, 20 * L)
    W = rng.normal(0, scale, (NA, L, NA, L))
    W = W + W.transpose(2, 3, 0, 1)
    for k in range(L):
        W[:, k, :, k] = 0.0
    X2 = np.zeros((NA, L, NAP, L))
    X2[:, :, :NA, :] = W
    x[nsp:] = X2.ravel()
    return x

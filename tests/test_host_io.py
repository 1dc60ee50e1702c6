"""Raw coupling output of the library (host code, no GPU): the text format
of io.c:31-103 and the MessagePack "ccm-1" format of io.c:105-196.

* Text: write_raw then read_raw restores a symmetric x with zero self-edge
  blocks and zero padding exactly, except the singles, which io.c:43 prints
  with "%g" (six significant digits): those come back as float("%g" % v).
* MessagePack: the bytes equal what the msgpack packer produces for the
  same document built from x with the reference's indexing (x_single[i*20+a]
  = V(i,a) = x[a*L+i]; pair "i/j" x[a*21+b] = W(b,j,a,i)), with float32
  values; and the decoded document, written back into a zero buffer the way
  read_raw fills it (both mirrored entries), equals the text round trip.
"""
import ctypes as C

import numpy as np
import pytest

from ccmpred_amd import _lib, pll

msgpack = pytest.importorskip("msgpack")

NA, NAP = 21, 32

_libc = C.CDLL(None)
_libc.fopen.restype = C.c_void_p
_libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
_libc.fclose.argtypes = [C.c_void_p]


def _lib_io():
    L = _lib.lib()
    L.ccm_write_raw.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    L.ccm_write_raw.restype = None
    L.ccm_write_raw_msgpack.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    L.ccm_write_raw_msgpack.restype = C.c_int
    return L


def symmetric_x(L, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros(pll.nvar_padded(L), np.float32)
    ns = 20 * L
    x[:ns] = rng.normal(0, 1.3, ns)
    W = rng.normal(0, 0.2, (NA, L, NA, L))
    W = W + W.transpose(2, 3, 0, 1)
    W[:, np.arange(L), :, np.arange(L)] = 0
    x[pll.nsingle_padded(L):].reshape(NA, L, NAP, L)[:, :, :NA, :] = W
    return x


def write_with(fn, path, x, L):
    f = _libc.fopen(str(path).encode(), b"w")
    assert f
    rc = fn(f, x.ctypes.data, L)
    _libc.fclose(f)
    return rc


@pytest.mark.parametrize("L", [1, 2, 7, 40])
def test_raw_text_round_trip(tmp_path, L):
    lib = _lib_io()
    x = symmetric_x(L, L)
    p = tmp_path / "x.raw"
    write_with(lib.ccm_write_raw, p, x, L)
    lines = p.read_text().splitlines()
    assert len(lines) == L + 22 * L * (L - 1) // 2
    y = np.zeros_like(x)
    assert lib.ccm_read_raw(str(p).encode(), L, y.ctypes.data) == 0
    ns, nsp = 20 * L, pll.nsingle_padded(L)
    singles = np.array([float("%g" % v) for v in x[:ns]], np.float32)
    assert np.array_equal(y[:ns], singles)
    assert not y[ns:nsp].any()
    assert np.array_equal(y[nsp:], x[nsp:])
    # a file whose pair header is out of order is rejected
    bad = "\n".join(lines[:L] + ["# 0 2" if L > 2 else "# 9 9"] + lines[L + 1:]) + "\n"
    if L > 1:
        q = tmp_path / "bad.raw"
        q.write_text(bad)
        assert lib.ccm_read_raw(str(q).encode(), L, y.ctypes.data) == -1


def expected_doc(x, L):
    V = x[:20 * L].reshape(20, L)                       # V(i,a) = x[a*L+i]
    X2 = x[pll.nsingle_padded(L):].reshape(NA, L, NAP, L)  # W(b,k,a,j) = X2[b,k,a,j]
    pairs = {}
    for i in range(L):
        for j in range(i + 1, L):
            blk = X2[:, j, :NA, i]                      # [b][a] = W(b,j,a,i)
            pairs[f"{i}/{j}"] = {"i": i, "j": j, "x": [float(v) for v in blk.T.ravel()]}
    return {"format": "ccm-1", "ncol": L,
            "x_single": [float(v) for v in V.T.ravel()], "x_pair": pairs}


@pytest.mark.parametrize("L", [1, 3, 20, 130])
def test_msgpack_bytes_and_round_trip(tmp_path, L):
    lib = _lib_io()
    x = symmetric_x(L, 100 + L)
    p = tmp_path / "x.braw"
    assert write_with(lib.ccm_write_raw_msgpack, p, x, L) == 0
    got = p.read_bytes()
    assert got == msgpack.packb(expected_doc(x, L), use_single_float=True)
    # ncol >= 128 and the 8128 pairs of L=130 take the uint8 / map16 forms
    doc = msgpack.unpackb(got, strict_map_key=False)
    assert doc["format"] == "ccm-1" and doc["ncol"] == L
    assert len(doc["x_pair"]) == L * (L - 1) // 2
    # fill a buffer from the document the way read_raw does and compare it
    # with the text round trip (the couplings print exactly in both)
    y = np.zeros_like(x)
    y[:20 * L] = np.asarray(doc["x_single"], np.float32).reshape(L, 20).T.ravel()
    Y2 = y[pll.nsingle_padded(L):].reshape(NA, L, NAP, L)
    for e in doc["x_pair"].values():
        i, j = e["i"], e["j"]
        blk = np.asarray(e["x"], np.float32).reshape(NA, NA)   # [a][b] = W(b,j,a,i)
        Y2[:, j, :NA, i] = blk.T
        Y2[:, i, :NA, j] = blk
    t = tmp_path / "x.raw"
    write_with(lib.ccm_write_raw, t, x, L)
    z = np.zeros_like(x)
    assert lib.ccm_read_raw(str(t).encode(), L, z.ctypes.data) == 0
    nsp = pll.nsingle_padded(L)
    assert np.array_equal(y[nsp:], z[nsp:])
    assert np.array_equal(y[nsp:], x[nsp:])
    assert np.array_equal(y[:20 * L], x[:20 * L])

"""Packet-ordered MODWT: the oracle's restatement of convert_modwt
(src/wavutils.cpp:532-600) and of the packet-ordered transform
(src/transform.cpp:39-55) against the golden vectors written by a build of
the reference's own sources (tests/golden/make_golden.py)."""
from pathlib import Path

import numpy as np
import pytest

import oracle_bind as O

G = np.load(Path(__file__).resolve().parent / "golden" / "reference_vectors.npz")


def cases(prefix):
    return sorted({k.rsplit("_", 1)[0] for k in G.files if k.startswith(prefix)})


@pytest.mark.parametrize("key", cases("convert_"))
def test_convert_matches_reference(key):
    _, n, l = key.split("_")
    n, J = int(n[1:]), int(l[1:])
    v = G[key + "_in"]
    np.testing.assert_array_equal(O.modwt_convert(v, n, J, 1), G[key + "_to2po"])
    np.testing.assert_array_equal(O.modwt_convert(v, n, J, 2), G[key + "_po2to"])
    # a permutation within each level block, and the two directions invert
    for j in range(J):
        blk = slice(2 * n * j, 2 * n * (j + 1))
        np.testing.assert_array_equal(np.sort(G[key + "_to2po"][blk]), np.sort(v[blk]))
    np.testing.assert_array_equal(O.modwt_convert(G[key + "_to2po"], n, J, 2), v)


@pytest.mark.parametrize("key", cases("modwtpo_"))
def test_po_transform_matches_reference(key):
    _, f, n, l = key.split("_")
    f, n, J = int(f[1:]), int(n[1:]), int(l[1:])
    x = G[key + "_x"]
    # the reference's own packet-ordered forward equals the time-ordered
    # transform followed by the conversion, bit for bit
    np.testing.assert_array_equal(O.modwt_po_fwd(x, J, f), G[key + "_fwd"])
    # its inverse averages packet reconstructions in another order: last-bit
    # differences only
    xr = O.modwt_po_inv(G[key + "_fwd"], n, J, f)
    assert O.relerr(xr, G[key + "_inv"]) <= 1e-14
    assert O.relerr(xr, x) <= 1e-13


def test_packet_mapping_closed_form():
    """Level l's packet `shift` holds the time positions reverse_{l+1}(shift)
    + 2^{l+1} k: tag every time position with its index and read it back."""
    n, J = 64, 5
    to = np.concatenate([np.concatenate([np.arange(n) + 1000 * j, -(np.arange(n) + 1000 * j)])
                         for j in range(J)]).astype(np.float64)
    po = O.modwt_convert(to, n, J, 1)
    for l in range(J):
        P, lenos = 2 ** (l + 1), n >> l
        for shift in range(P):
            rev = int(format(shift, f"0{l + 1}b")[::-1], 2)
            pk = po[2 * n * l + shift * lenos: 2 * n * l + (shift + 1) * lenos]
            t = rev + P * np.arange(lenos // 2)
            np.testing.assert_array_equal(pk[0::2], t + 1000 * l)
            np.testing.assert_array_equal(pk[1::2], -(t + 1000 * l))

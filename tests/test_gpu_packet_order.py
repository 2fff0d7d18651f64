"""Packet-ordered MODWT on the device: wh_modwt_convert against the golden
conversions of the reference's convert_modwt and against the oracle (bit
for bit: a permutation), the packet-ordered transforms through the `.C`
symbol RgpuTransform (ttype 2) against the reference's own packet-ordered
Haar / D4 (golden) and the oracle (C6, LA8), and a full-size round trip."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle_bind as O
from wavecuda_amd import _lib as W
from wavecuda_amd import device as D

pytestmark = pytest.mark.gpu

G = np.load(Path(__file__).resolve().parent / "golden" / "reference_vectors.npz")
FLEN = {1: 2, 2: 4, 3: 6, 4: 8}


def keys(prefix):
    return sorted({k.rsplit("_", 1)[0] for k in G.files if k.startswith(prefix)})


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("key", keys("convert_"))
def test_convert_matches_reference(key, dtype):
    _, n, l = key.split("_")
    n, J = int(n[1:]), int(l[1:])
    print(f"convert {key} {dtype}", flush=True)
    v = torch.tensor(G[key + "_in"], dtype=dtype, device="cuda")
    po = D.modwt_convert(v, n, J, "Haar", to_po=True).cpu().numpy()
    to = D.modwt_convert(v, n, J, "Haar", to_po=False).cpu().numpy()
    np.testing.assert_array_equal(po, G[key + "_to2po"].astype(po.dtype))
    np.testing.assert_array_equal(to, G[key + "_po2to"].astype(to.dtype))


@pytest.mark.parametrize("n,J", [(32, 5), (2048, 11), (1 << 13, 9), (1 << 16, 16)])
def test_convert_matches_oracle(n, J):
    """Tiles smaller than the level, and levels with more packets than a
    tile has positions (2^{l+1} > 2048)."""
    print(f"convert n={n} J={J}", flush=True)
    rng = np.random.default_rng(n + J)
    v = rng.standard_normal(2 * n * J)
    t = torch.tensor(v, device="cuda")
    po = D.modwt_convert(t, n, J, "Haar", True)
    np.testing.assert_array_equal(po.cpu().numpy(), O.modwt_convert(v, n, J, 1))
    back = D.modwt_convert(po, n, J, "Haar", False)
    np.testing.assert_array_equal(back.cpu().numpy(), v)


def rgpu(x, xmod, n, sense, J, ttype, f):
    ip = lambda v: C.byref(C.c_int(int(v)))
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    W.lib().RgpuTransform(dp(x), dp(xmod), ip(n), ip(sense), ip(J), ip(ttype), ip(f),
                          ip(FLEN[f]))


@pytest.mark.parametrize("key", keys("modwtpo_"))
def test_po_transform_matches_reference(key):
    _, f, n, l = key.split("_")
    f, n, J = int(f[1:]), int(n[1:]), int(l[1:])
    x = G[key + "_x"].copy()
    xm = np.zeros(2 * n * J)
    rgpu(x, xm, n, 1, J, 2, f)
    assert O.relerr(xm, G[key + "_fwd"]) <= 1e-12
    xr = np.zeros(n)
    rgpu(xr, G[key + "_fwd"].copy(), n, 0, J, 2, f)
    assert O.relerr(xr, G[key + "_inv"]) <= 1e-12


@pytest.mark.parametrize("f", [3, 4])
def test_po_transform_c6_la8(f):
    n, J = 4096, 6
    x = np.random.default_rng(f).standard_normal(n)
    xm = np.zeros(2 * n * J)
    rgpu(x.copy(), xm, n, 1, J, 2, f)
    assert O.relerr(xm, O.modwt_po_fwd(x, J, f)) <= 1e-12
    xr = np.zeros(n)
    rgpu(xr, xm.copy(), n, 0, J, 2, f)
    assert O.relerr(xr, x) <= 1e-12
    t = torch.tensor(x, device="cuda")
    np.testing.assert_array_equal(D.modwt_inv_po(D.modwt_fwd_po(t, J, f), n, J, f).cpu().numpy(),
                                  D.modwt_inv(D.modwt_fwd(t, J, f), n, J, f).cpu().numpy())


def test_full_size_round_trip_and_mapping():
    """n = 2^24, J = 10: TO -> PO -> TO is the identity, and on a sample of
    packet-order entries the value (a time-order index tag) is the one the
    bit-reversal mapping names."""
    n, J = 1 << 24, 10
    print("round trip n=2^24 J=10", flush=True)
    tag = torch.arange(2 * n * J, dtype=torch.float64, device="cuda")
    po = D.modwt_convert(tag, n, J, "LA8", True)
    back = D.modwt_convert(po, n, J, "LA8", False)
    assert torch.equal(back, tag)
    g = torch.Generator(device="cuda").manual_seed(0)
    p = torch.randint(0, 2 * n * J, (1 << 20,), device="cuda", generator=g)
    l = p // (2 * n)
    q = p - l * 2 * n
    lenos = n >> l
    shift = q // lenos
    within = q - shift * lenos
    k, sd = within // 2, within % 2
    rev = torch.zeros_like(shift)  # the l + 1 low bits of shift, reversed
    for b in range(J):
        bit = (shift >> b) & 1
        rev = rev + torch.where(l >= b, bit << (l - b).clamp(min=0), torch.zeros_like(bit))
    t = rev + (2 << l) * k
    expect = l * 2 * n + sd * n + t
    assert torch.equal(po[p], expect.to(torch.float64))


def test_convert_argument_checks():
    v = torch.zeros(2 * 1024 * 4, dtype=torch.float64, device="cuda")
    rc = W.lib().wh_modwt_convert(v.data_ptr(), v.data_ptr(), 1024, 4, 1, 1, W.WH_F64, None)
    assert rc == W.WH_ERR_ARG
    w = torch.zeros_like(v)
    assert W.lib().wh_modwt_convert(v.data_ptr(), w.data_ptr(), 1024, 4, 1, 2, W.WH_F64,
                                    None) == W.WH_ERR_ARG
    assert W.lib().wh_modwt_convert(v.data_ptr() + 8, w.data_ptr(), 1024, 4, 1, 1, W.WH_F64,
                                    None) == W.WH_ERR_ARG

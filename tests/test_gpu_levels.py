"""Coefficients by level on the device: wh_coeff_labels against the
oracle's restatement of WSTtoDT's labels (bit for bit, up to n = 2^22),
wh_coeff_level's gathers against indexing with getCoeffLevel's positions,
its argument checks, and the R-surface mirrors label_detail_scaling and
WSTtoDT over a GPUTransform result."""
import numpy as np
import pytest
import torch

import oracle_bind as O
from wavecuda_amd import _lib as W
from wavecuda_amd import device as D
from wavecuda_amd import rsurface as R

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,J", [(8, 1), (8, 3), (1024, 1), (1024, 5), (1024, 9),
                                 (1 << 22, 12), (1 << 22, 21)])
def test_dwt_labels_match_oracle(n, J):
    print(f"dwt labels n={n} J={J}", flush=True)
    lev, tr, kind = D.coeff_labels(n, J, 0, "D4" if n > 8 else "Haar")
    ol, ot, ok = O.coeff_labels(n, J, 0)
    np.testing.assert_array_equal(lev.cpu().numpy(), ol)
    np.testing.assert_array_equal(tr.cpu().numpy(), ot)
    np.testing.assert_array_equal(kind.cpu().numpy(), ok)


@pytest.mark.parametrize("n,J", [(16, 1), (1024, 5), (1 << 16, 8), (1 << 20, 10)])
def test_modwt_labels_match_oracle(n, J):
    print(f"modwt labels n={n} J={J}", flush=True)
    lev, tr, kind = D.coeff_labels(n, J, 1, "LA8" if n >= 1024 else "Haar")
    ol, ot, ok = O.coeff_labels(n, J, 1)
    np.testing.assert_array_equal(lev.cpu().numpy(), ol)
    np.testing.assert_array_equal(tr.cpu().numpy(), ot)
    np.testing.assert_array_equal(kind.cpu().numpy(), ok)


@pytest.mark.parametrize("case", ["unaligned", "no_translate", "n2"])
def test_modwt_labels_scalar_and_partial(case):
    """The label columns through the C ABI with the element-wise kernel
    (columns one element off their 16-byte alignment, or n = 2) and with a
    column left out (NULL)."""
    n, J = (2, 1) if case == "n2" else (1 << 12, 6)
    m = 2 * n * J
    print(f"modwt labels {case} n={n} J={J}", flush=True)
    off = 1 if case == "unaligned" else 0
    lev = torch.full((m + off,), -7, dtype=torch.int32, device="cuda")
    tr = torch.full((m + off,), -7.0, dtype=torch.float64, device="cuda")
    kind = torch.full((m + off,), 9, dtype=torch.uint8, device="cuda")
    trp = None if case == "no_translate" else tr.data_ptr() + 8 * off
    W.check(W.lib().wh_coeff_labels(n, J, W.MODWT, W.HAAR, lev.data_ptr() + 4 * off, trp,
                                    kind.data_ptr() + off, None), "wh_coeff_labels")
    torch.cuda.synchronize()
    ol, ot, ok = O.coeff_labels(n, J, 1)
    np.testing.assert_array_equal(lev[off:].cpu().numpy(), ol)
    np.testing.assert_array_equal(kind[off:].cpu().numpy(), ok)
    if trp is None:
        assert (tr == -7.0).all()
    else:
        np.testing.assert_array_equal(tr[off:].cpu().numpy(), ot)
    if off:
        assert lev[0].item() == -7 and tr[0].item() == -7.0 and kind[0].item() == 9


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("n,filt,J", [(1024, "D4", 8), (1 << 16, "LA8", 13), (64, "Haar", 6)])
def test_dwt_level_gather(dtype, n, filt, J):
    print(f"dwt gather n={n} {filt} J={J} {dtype}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(n)
    x = torch.randn(n, dtype=dtype, device="cuda", generator=g)
    y = D.dwt_fwd(x, J, filt)
    yh = y.cpu().numpy()
    for l in range(1, J + 1):
        got = D.coeff_level(y, n, J, 0, filt, l, "d").cpu().numpy()
        pos = O.coeff_positions(n, J, 0, l, "d")
        np.testing.assert_array_equal(got, yh[pos - 1])
    got = D.coeff_level(y, n, J, 0, filt, J, "s").cpu().numpy()
    np.testing.assert_array_equal(got, yh[O.coeff_positions(n, J, 0, J, "s") - 1])


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_modwt_level_gather(dtype):
    n, J, filt = 1 << 14, 8, "C6"
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.randn(n, dtype=dtype, device="cuda", generator=g)
    w = D.modwt_fwd(x, J, filt)
    wh = w.cpu().numpy()
    for l in range(1, J + 1):
        np.testing.assert_array_equal(D.coeff_level(w, n, J, 1, filt, l, "s").cpu().numpy(),
                                      wh[2 * n * (l - 1): 2 * n * (l - 1) + n])
        np.testing.assert_array_equal(D.coeff_level(w, n, J, 1, filt, l, "d").cpu().numpy(),
                                      wh[2 * n * (l - 1) + n: 2 * n * l])


def test_level_argument_checks():
    y = torch.zeros(1024, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        D.coeff_level(y, 1024, 5, 0, "D4", 0)
    with pytest.raises(ValueError):
        D.coeff_level(y, 1024, 5, 0, "D4", 6)
    with pytest.raises(Exception, match="nlevels only"):
        D.coeff_level(y, 1024, 5, 0, "D4", 2, "s")


@pytest.mark.parametrize("ttype,filt,nl", [("DWT", "D4", 0), ("MODWT", "LA8", 4),
                                           ("DWT", "Haar", 1)])
def test_wsttodt_and_labels(ttype, filt, nl):
    n = 256
    rng = np.random.default_rng(7)
    x = rng.standard_normal(n)
    w = R.GPUTransform(x, "FWD", nl, ttype, filt)
    J = w.nlevels
    tt = 0 if ttype == "DWT" else 1
    ol, ot, ok = O.coeff_labels(n, J, tt)
    lab = R.label_detail_scaling(w)
    np.testing.assert_array_equal(lab, np.where(ok == 1, "s", "d"))
    df = R.WSTtoDT(w)
    assert list(df.columns) == ["W", "CoeffType", "Level", "Translate"]
    assert len(df) == ol.size
    np.testing.assert_array_equal(df["W"].to_numpy(), w.w[: ol.size])
    np.testing.assert_array_equal(df["Level"].to_numpy(), ol)
    np.testing.assert_array_equal(df["Translate"].to_numpy(), ot)
    np.testing.assert_array_equal(df["CoeffType"].to_numpy(), np.where(ok == 1, "s", "d"))
    dd = R.WSTtoDT(w, scaling=False)
    assert (dd["CoeffType"] == "d").all() and len(dd) == int((ok == 0).sum())
    pl = R.WSTtoDT(w, forPlotting=True)
    assert list(pl.columns) == ["W", "CoeffType", "Level", "Translate", "minmax"]
    pads = pl[pl["minmax"] > 0]
    assert len(pads) == 2 * J and (pads["Translate"] == 0).all()
    for l in range(1, J + 1):
        m = np.abs(w.w[: ol.size][(ol == l) & (ok == 0)]).max()
        got = pads[pads["Level"] == l].sort_values("minmax")["W"].to_numpy()
        np.testing.assert_array_equal(got, [-m, m])

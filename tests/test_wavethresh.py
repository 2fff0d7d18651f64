"""WSTtoWavethresh (R/rfunctions.R:864-893): the wd() arguments and the
putD / putC calls it makes, in order.

CPU: the layout table (wh_wavethresh_layout needs no device) against the R
code's own arithmetic: wavethresh's level count for n values is log2 n; the
loop l = 1..nlevels puts the level-l details at wavethresh level
log2(n) - l; putC then uses the loop variable's last value.  GPU: the packed
vectors equal the level gathers (getCoeffLevel's positions) of the oracle's
transform.  The wavethresh package itself (the wd object it builds from
these) is third-party R code absent here: its side is parity unpinned.
"""
import numpy as np
import pytest

import oracle_bind as O
from wavecuda_amd import rsurface as R

FILTERS = ["Haar", "D4", "C6", "LA8"]
FCODE = {"Haar": 1, "D4": 2, "C6": 3, "LA8": 4}


@pytest.mark.parametrize("ttype", ["DWT", "MODWT"])
@pytest.mark.parametrize("filt", FILTERS)
@pytest.mark.parametrize("n,nl", [(1024, 0), (1024, 5), (256, 1), (1 << 16, 8)])
def test_layout_follows_the_R_loop(ttype, filt, n, nl):
    J = O.levels(n, nl, FCODE[filt])
    lay = R.wavethresh_layout(n, J, ttype, filt)
    lg = int(np.log2(n))
    want = []
    off = 0
    for l in range(1, J + 1):
        c = (n >> l) if ttype == "DWT" else n
        want.append((lg - l, "d", off, c))
        off += c
    c = (n >> J) if ttype == "DWT" else n
    want.append((lg - J, "s", off, c))  # putC at nlevels - l with l = nlevels
    assert lay["calls"] == want
    # filter.number / family switches (R/rfunctions.R:866-867)
    assert lay["filter.number"] == FCODE[filt]
    assert lay["family"] == ("DaubExPhase" if filt in ("Haar", "D4") else "DaubLeAsymm")
    assert lay["type"] == ("wavelet" if ttype == "DWT" else "station")
    assert lay["warnings"] == (1 if filt == "C6" else 0) | (2 if ttype == "MODWT" else 0)
    total = sum(c[3] for c in lay["calls"])
    assert total == (n if ttype == "DWT" else n * (J + 1))


def test_layout_rejects_bad_shapes():
    with pytest.raises(ValueError):
        R.wavethresh_layout(1000, 0, "DWT", "Haar")
    with pytest.raises(ValueError):
        R.wavethresh_layout(256, 7, "DWT", "LA8")


@pytest.mark.gpu
@pytest.mark.parametrize("ttype", ["DWT", "MODWT"])
@pytest.mark.parametrize("filt", FILTERS)
def test_pack_equals_level_gathers(ttype, filt):
    import warnings

    from wavecuda_amd.signals import noisy_doppler

    n = 4096
    nl = 0 if ttype == "DWT" else 6
    x, _ = noisy_doppler(n, 4.0, seed=FCODE[filt] + 40)
    XW = R.GPUTransform(x, "FWD", nl, ttype, filt)
    J = XW.nlevels
    with warnings.catch_warnings(record=True) as rec:
        warnings.simplefilter("always")
        res = R.WSTtoWavethresh(XW)
    nwarn = (filt == "C6") + (ttype == "MODWT")
    assert len(rec) == nwarn
    ref = O.dwt(x, J, FCODE[filt]) if ttype == "DWT" else O.modwt_fwd(x, J, FCODE[filt])
    puts = res["puts"]
    assert [p["fn"] for p in puts] == ["putD"] * J + ["putC"]
    lg = int(np.log2(n))
    for l in range(1, J + 1):
        p = puts[l - 1]
        assert p["level"] == lg - l
        pos = O.coeff_positions(n, J, 0 if ttype == "DWT" else 1, l, "d")
        if ttype == "DWT":
            want = ref[pos - 1]
        else:
            want = ref[2 * n * (l - 1) + n: 2 * n * l]
        # the pack is a gather of the GPU transform: equal to it bit for
        # bit, and within 1e-12 of the oracle's coefficients
        got_gpu = (XW.w[pos - 1] if ttype == "DWT" else XW.w[2 * n * (l - 1) + n: 2 * n * l])
        np.testing.assert_array_equal(p["v"], got_gpu)
        assert O.relerr(p["v"], want) <= 1e-12
    pc = puts[-1]
    assert pc["level"] == lg - J
    want = ref[:: 1 << J] if ttype == "DWT" else ref[2 * n * (J - 1): 2 * n * (J - 1) + n]
    assert O.relerr(pc["v"], want) <= 1e-12

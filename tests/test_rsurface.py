"""The R surface (GPUTransform, GPUReconstruct, GPUSmooth, GPUThreshold)
restated over the `.C` symbols: argument checks (CPU, no device needed),
the returned WST object field by field, and results against the CPU oracle
(the reference's CPUTransform / CPUSmooth path) and its golden vectors.

The reference's own usage example (README.md:18-30: D4 DWT of a doppler
signal of n = 1024 with noise) is the first GPU case.
"""
import math

import numpy as np
import pytest

import oracle_bind as O
from wavecuda_amd import rsurface as R
from wavecuda_amd.signals import noisy_doppler

FILTERS = ["Haar", "D4", "C6", "LA8"]
FCODE = {"Haar": 1, "D4": 2, "C6": 3, "LA8": 4}


# ---- checks: no device needed ---------------------------------------------

def test_trans_checks_match_reference_rules():
    x = np.zeros(1024)
    a = R.check_trans_inputs(x, "FWD", 0, "DWT", "D4")
    assert a["nlevels"] == 9 and a["len"] == 1024 and a["ttype"] == 0 and a["filt"] == 2
    assert a["filtlen"] == 4 and a["sense"] == 1 and a["xmod"].size == 0
    m = R.check_trans_inputs(x, "FWD", 4, "MODWT", "LA8")
    assert m["xmod"].size == 2 * 1024 * 10  # 2 n log2 n, as R allocates
    b = R.check_trans_inputs(np.zeros(2 * 1024 * 4), "BWD", 4, "MODWT", "Haar")
    assert b["len"] == 1024 and b["x"].size == 1024
    with pytest.raises(ValueError, match="Unexpected transform type"):
        R.check_trans_inputs(x, "FWD", 0, "PWT", "D4")
    with pytest.raises(ValueError, match="Unexpected filter"):
        R.check_trans_inputs(x, "FWD", 0, "DWT", "D8")
    with pytest.raises(ValueError, match="Unexpected direction"):
        R.check_trans_inputs(x, "UP", 0, "DWT", "D4")
    with pytest.raises(ValueError, match="power of 2"):
        R.check_trans_inputs(np.zeros(1000), "FWD", 0, "DWT", "D4")
    with pytest.raises(ValueError, match="between 1 and 9"):
        R.check_trans_inputs(x, "FWD", 10, "DWT", "D4")
    with pytest.raises(ValueError, match="infinite length"):
        R.check_trans_inputs(np.zeros(2048), "BWD", 0, "MODWT", "Haar")
    with pytest.raises(ValueError, match="really a MODWT vector"):
        R.check_trans_inputs(np.zeros(3000), "BWD", 2, "MODWT", "Haar")


def test_smooth_and_thresh_checks():
    x = np.zeros(256)
    a = R.check_smooth_inputs(x, 5, "MODWT", "LA8", "cv", None, "hard", 2, 5, 0.01)
    assert (a["min_level"], a["max_level"], a["hardness"], a["threshtype"]) == (1, 4, 0, 2)
    with pytest.raises(ValueError, match="manual, univ or cv"):
        R.check_smooth_inputs(x, 5, "MODWT", "LA8", "sure", None, "hard", 1, 5, 0.01)
    with pytest.raises(ValueError, match="greater than 0"):
        R.check_smooth_inputs(x, 5, "MODWT", "LA8", "manual", -1.0, "hard", 1, 5, 0.01)
    with pytest.raises(ValueError, match="Tolerance"):
        R.check_smooth_inputs(x, 5, "MODWT", "LA8", "cv", None, "hard", 1, 5, 0.0)
    with pytest.raises(ValueError, match="hard or soft"):
        R.check_smooth_inputs(x, 5, "MODWT", "LA8", "univ", None, "firm", 1, 5, 0.01)
    with pytest.raises(ValueError, match="range 1...nlevels"):
        R.check_smooth_inputs(x, 5, "MODWT", "LA8", "univ", None, "soft", 0, 5, 0.01)
    # nlevels = 0 makes any max.level out of range in the reference's check
    with pytest.raises(ValueError, match="range 1...nlevels"):
        R.check_smooth_inputs(x, 0, "MODWT", "LA8", "univ", None, "soft", 1, 1, 0.01)
    wst = R.WST(x=x, ttype="DWT", filt="D4", filtlen=4, nlevels=3, len=256, w=np.zeros(256))
    t = R.check_thresh_inputs(wst, "soft", 0.5, 1, 3)
    assert (t["min_level"], t["max_level"], t["thresh"], t["hardness"]) == (0, 2, 0.5, 1)
    with pytest.raises(ValueError, match="scalar"):
        R.check_thresh_inputs(wst, "soft", [0.5, 0.2], 1, 3)
    with pytest.raises(ValueError, match="range 1...nlevels"):
        R.check_thresh_inputs(wst, "soft", 0.5, 1, 4)


# ---- on the GPU -----------------------------------------------------------

@pytest.mark.gpu
def test_readme_example_d4_doppler():
    x, _ = noisy_doppler(1024, 5.0, seed=10)
    w = R.GPUTransform(x, "FWD", 0, "DWT", "D4")
    assert isinstance(w, R.WST)
    assert (w.ttype, w.filt, w.filtlen, w.nlevels, w.len) == ("DWT", "D4", 4, 9, 1024)
    np.testing.assert_array_equal(w.x, x)
    assert O.relerr(w.w, O.dwt(x, 9, 2)) <= 1e-12
    assert O.relerr(R.GPUReconstruct(w), x) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("filt", FILTERS)
@pytest.mark.parametrize("ttype", ["DWT", "MODWT"])
def test_transform_wst_fields_and_values(filt, ttype):
    n = 4096
    x, _ = noisy_doppler(n, 4.0, seed=FCODE[filt])
    nl = 0 if ttype == "DWT" else 6
    w = R.GPUTransform(x, "FWD", nl, ttype, filt)
    J = O.levels(n, nl, FCODE[filt])
    assert (w.ttype, w.filt, w.nlevels, w.len) == (ttype, filt, J, n)
    if ttype == "DWT":
        assert w.w.size == n
        assert O.relerr(w.w, O.dwt(x, J, FCODE[filt])) <= 1e-12
    else:
        assert w.w.size == 2 * n * int(math.log2(n))  # R's allocation, zero-padded
        assert not np.any(w.w[2 * n * J:])
        assert O.relerr(w.w[:2 * n * J], O.modwt_fwd(x, J, FCODE[filt])) <= 1e-12
    assert O.relerr(R.GPUReconstruct(w), x) <= 1e-12
    if ttype == "DWT":
        xb = R.GPUTransform(w.w, "BWD", J, ttype, filt)
        assert O.relerr(xb, x) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("filt", ["Haar", "LA8"])
@pytest.mark.parametrize("thresh_type", ["univ", "cv", "manual"])
@pytest.mark.parametrize("hard_soft", ["soft", "hard"])
def test_smooth_matches_cpu_path(filt, thresh_type, hard_soft):
    n = 1 << 12
    x, _ = noisy_doppler(n, 3.0, seed=21)
    J = 8
    thresh = 0.3 if thresh_type == "manual" else None
    y = R.GPUSmooth(x, J, "MODWT", filt, thresh_type, thresh, hard_soft, 2, 7)
    tt = {"manual": 0, "univ": 1, "cv": 2}[thresh_type]
    ys, _, rc = O.smooth(x, J, 1, FCODE[filt], tt, thresh, 1 if hard_soft == "soft" else 0, 1, 6)
    assert rc == 0
    assert O.relerr(y, ys) <= 1e-12


@pytest.mark.gpu
def test_threshold_wst():
    n = 2048
    x, _ = noisy_doppler(n, 3.0, seed=5)
    w = R.GPUTransform(x, "FWD", 5, "MODWT", "D4")
    wt = R.GPUThreshold(w, 0.2, "soft", 2, 4)
    ref = O.threshold(w.w[:2 * n * 5], n, 5, 1, 2, 0.2, 1, 1, 3)
    np.testing.assert_array_equal(wt.w[:2 * n * 5], ref)
    assert (wt.ttype, wt.nlevels, wt.len) == (w.ttype, w.nlevels, w.len)


@pytest.mark.gpu
@pytest.mark.parametrize("ttype", ["DWT", "MODWT"])
def test_cpu_named_functions(ttype):
    """CPUTransform / CPUReconstruct / CPUThreshold / CPUSmooth over the
    Rcpu* symbols: the same WST fields and values as the oracle (the
    reference's CPU path), and the GPU-named functions' results."""
    n, filt = 2048, "LA8"
    x, _ = noisy_doppler(n, 3.0, seed=6)
    nl = 0 if ttype == "DWT" else 5
    w = R.CPUTransform(x, "FWD", nl, ttype, filt)
    g = R.GPUTransform(x, "FWD", nl, ttype, filt)
    J = O.levels(n, nl, 4)
    assert (w.ttype, w.filt, w.filtlen, w.nlevels, w.len) == (ttype, filt, 8, J, n)
    np.testing.assert_array_equal(w.w, g.w)
    want = O.dwt(x, J, 4) if ttype == "DWT" else O.modwt_fwd(x, J, 4)
    assert O.relerr(w.w[: want.size], want) <= 1e-12
    assert O.relerr(R.CPUReconstruct(w), x) <= 1e-12
    wt = R.CPUThreshold(w, 0.2, "hard", 1, J - 1)
    ref = O.threshold(w.w[: want.size], n, J, 0 if ttype == "DWT" else 1, 4, 0.2, 0, 0, J - 2)
    np.testing.assert_array_equal(wt.w[: want.size], ref)
    ys = R.CPUSmooth(x, J, ttype, filt, "univ", None, "soft", 1, J)
    np.testing.assert_array_equal(ys, R.GPUSmooth(x, J, ttype, filt, "univ", None, "soft", 1, J))
    yo, _, rc = O.smooth(x, J, 0 if ttype == "DWT" else 1, 4, 1, None, 1, 0, J - 1)
    assert rc == 0 and O.relerr(ys, yo) <= 1e-12

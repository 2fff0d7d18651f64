"""Coefficients by level: the oracle's restatement of getCoeffLevel and
WSTtoDT's labels (R/rplotting.R:10-46, R/rfunctions.R:902-1004), pinned
against the DWT layout itself and against closed forms; the R-surface
mirror's positions and argument errors (no device needed)."""
import numpy as np
import pytest

import oracle_bind as O
from wavecuda_amd import rsurface as R


def dwt_cases():
    for n, L in ((16, 2), (64, 4), (256, 8), (1024, 6)):
        f = {2: 1, 4: 2, 6: 3, 8: 4}[L]
        Jmax = O.levels(n, 0, f)
        for J in sorted({1, max(1, Jmax // 2), Jmax}):
            yield n, f, J


@pytest.mark.parametrize("n,f,J", list(dwt_cases()))
def test_positions_hold_their_level(n, f, J):
    """A J-level DWT keeps d_l where an l-level DWT put it: the values at the
    positions labelled (l, "d") equal the l-level transform's, for every l;
    a constant signal puts c 2^{J/2} exactly at the positions labelled "s"
    and zero elsewhere."""
    rng = np.random.default_rng(n + J)
    x = rng.standard_normal(n)
    yJ = O.dwt(x, J, f)
    seen = np.zeros(n, dtype=bool)
    for l in range(1, J + 1):
        pos = O.coeff_positions(n, J, 0, l, "d")
        assert pos is not None and pos.size == n >> l
        yl = O.dwt(x, l, f)
        np.testing.assert_array_equal(yJ[pos - 1], yl[pos - 1])
        # and not the s_l positions, which deeper levels rewrite
        if l < J:
            spos = np.arange(0, n, 2 ** l)
            assert not np.allclose(yJ[spos], yl[spos])
        seen[pos - 1] = True
    spos = O.coeff_positions(n, J, 0, J, "s")
    assert spos.size == n >> J
    seen[spos - 1] = True
    assert seen.all()  # every coefficient is some level's d or the s
    c = 0.75
    yc = O.dwt(np.full(n, c), J, f)
    np.testing.assert_allclose(yc[spos - 1], c * 2 ** (J / 2), rtol=1e-12)
    rest = np.ones(n, dtype=bool)
    rest[spos - 1] = False
    assert np.abs(yc[rest]).max() < 1e-12


@pytest.mark.parametrize("n,f,J", list(dwt_cases()))
def test_dwt_labels_closed_form(n, f, J):
    """The oracle follows WSTtoDT's loops; the labels equal the closed form:
    position i % 2^J == 0 is s_J[i / 2^J], otherwise d_l with l - 1 the
    lowest set bit of i; translate 2^{l-1} (m + 1/2) for index m."""
    level, translate, kind = O.coeff_labels(n, J, 0)
    i = np.arange(n)
    s = (i % (1 << J)) == 0
    low = np.array([(int(v) & -int(v)).bit_length() if v else 0 for v in i])
    l_exp = np.where(s, J, low)
    m_exp = i >> l_exp
    np.testing.assert_array_equal(kind, s.astype(np.uint8))
    np.testing.assert_array_equal(level, l_exp)
    np.testing.assert_array_equal(translate, 2.0 ** (l_exp - 1) * (m_exp + 0.5))
    # label.detail.scaling's own rule, (1:len) %% 2^nlevels == 1
    np.testing.assert_array_equal(kind == 1, (np.arange(1, n + 1) % 2 ** J) == 1)


@pytest.mark.parametrize("n,J", [(16, 1), (64, 3), (256, 6)])
def test_modwt_labels_and_positions(n, J):
    level, translate, kind = O.coeff_labels(n, J, 1)
    assert level.size == 2 * n * J
    blk = np.arange(2 * n * J) // n
    np.testing.assert_array_equal(level, blk // 2 + 1)
    np.testing.assert_array_equal(kind, (blk % 2 == 0).astype(np.uint8))
    np.testing.assert_array_equal(translate, np.tile(np.arange(1, n + 1), 2 * J))
    # getCoeffLevel gives 1..len for a MODWT whatever the level and type
    for l in range(1, J + 1):
        for t in ("d", "s"):
            np.testing.assert_array_equal(O.coeff_positions(n, J, 1, l, t),
                                          np.arange(1, n + 1))


def test_getcoefflevel_mirror_matches_oracle():
    for n, f, J in dwt_cases():
        w = R.WST(x=None, ttype="DWT", filt="D4", filtlen=4, nlevels=J, len=n, w=np.zeros(n))
        for l in range(1, J + 1):
            np.testing.assert_array_equal(R.getCoeffLevel(w, l, "d"),
                                          O.coeff_positions(n, J, 0, l, "d"))
        np.testing.assert_array_equal(R.getCoeffLevel(w, J, "s"),
                                      O.coeff_positions(n, J, 0, J, "s"))
        if J > 1:
            assert O.coeff_positions(n, J, 0, 1, "s") is None
            with pytest.raises(ValueError, match="nlevels only"):
                R.getCoeffLevel(w, 1, "s")
    m = R.WST(x=None, ttype="MODWT", filt="LA8", filtlen=8, nlevels=3, len=64, w=np.zeros(384))
    np.testing.assert_array_equal(R.getCoeffLevel(m, 2, "d"), np.arange(1, 65))
    with pytest.raises(ValueError, match="between 1 and 3"):
        R.getCoeffLevel(m, 0)
    with pytest.raises(ValueError, match="between 1 and 3"):
        R.getCoeffLevel(m, 4)
    with pytest.raises(ValueError, match="between 1 and 3"):
        R.getCoeffLevel(m, 1.5)
    with pytest.raises(ValueError, match="WST type required"):
        R.getCoeffLevel(np.zeros(4), 1)

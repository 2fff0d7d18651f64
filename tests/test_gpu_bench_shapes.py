"""Oracle parity at the shapes bench.py times, under every plan the tuner
can pick for them.

* MODWT J = 10 at n = 2^20 and 2^24, all four filters: the forward against
  the oracle (all 2nJ values), the inverse on random coefficients and on
  per-level thresholded coefficients (perfect reconstruction cannot pin an
  inverse: the MODWT is redundant, so every left inverse reconstructs
  consistent coefficients), FP64 <= 1e-12 and FP32 <= 1e-5 against the FP64
  oracle.  Each candidate plan of the tuner's search space is forced in
  turn (wh_debug_modwt_set_plan) and must give the same result bit for bit;
  the plan the tuner then picks for the headline shape is one of them.
* DWT full depth at n = 2^20 and 2^24, all filters, forward and inverse
  (random and thresholded coefficients).
* The differentiable-smoothing launch: 512 signals of n = 2^14, LA8,
  J = 8, FP32 (and FP64), one many-signal launch; forward, thresholding
  inverse and both adjoints on sampled signals.  The adjoints are checked
  against the oracle's own transforms through identities the mathematics
  fixes: the forward's adjoint is sum_j 2^j R_j(cot S_j, cot D_j) with R_j
  the j-level inverse (all finer details zero), and the inverse's adjoint
  is the forward with D_j scaled by 2^-j, S_J by 2^-J and the other S
  slots zero.

The oracle runs its OpenMP variant here, which tests/test_oracle.py holds
bit-identical to the serial restatement.
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle_bind as O
from wavecuda_amd import _lib as W
from wavecuda_amd import device as D
from wavecuda_amd.signals import doppler

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5
FILTERS = [1, 2, 3, 4]
J10 = 10


def _x(n, seed):
    return np.random.Generator(np.random.PCG64(seed)).standard_normal(n)


def _block_err(got: torch.Tensor, ref: torch.Tensor, n: int) -> float:
    """Max over n-long blocks of max|got - ref| / max|ref| (on the GPU)."""
    g = got.double().view(-1, n)
    r = ref.view(-1, n)
    den = r.abs().amax(dim=1).clamp_min(1e-300)
    return ((g - r).abs().amax(dim=1) / den).max().item()


def _lams(J):
    return np.array([0.35 + 0.07 * l for l in range(J)])


def _thr(J, lams, dev="cuda"):
    return dict(value=0.0, hardness=W.SOFT, minlevel=0, maxlevel=J - 1,
                levels=torch.tensor(lams, dtype=torch.float64, device=dev))


def _check_plans(inverse, n, f, dtype, run, check):
    """Run `run()` under every distinct candidate plan; `check` the first
    result against the oracle and every other bit for bit against it."""
    cands = D.modwt_plan_candidates(inverse, n, J10, f, dtype)
    assert cands, "no candidate plans"
    first = None
    try:
        for plan, knobs in cands.items():
            D.modwt_set_plan(inverse, n, J10, f, dtype, knobs)
            assert D.modwt_plan(inverse, n, J10, f, dtype) == plan
            print(f"n=2^{n.bit_length() - 1} f={f} {dtype} {'inv' if inverse else 'fwd'} "
                  f"knobs={knobs} {plan}", flush=True)
            out = run()
            torch.cuda.synchronize()
            if first is None:
                check(out)
                first = out.clone()
            else:
                assert torch.equal(out, first), f"plan {plan} differs from the first plan"
    finally:
        D.modwt_set_plan(inverse, n, J10, f, dtype, None)
    return len(cands)


@pytest.mark.parametrize("lg", [20, 24])
@pytest.mark.parametrize("filt", FILTERS)
def test_modwt_j10_every_plan(filt, lg):
    n = 1 << lg
    x = _x(n, 100 + lg + filt)
    ref = O.modwt_fwd(x, J10, filt, omp=True)
    ref_d = torch.from_numpy(ref).cuda()
    del ref
    for dt, tol in ((torch.float64, TOL64), (torch.float32, TOL32)):
        xd = torch.tensor(x, dtype=dt, device="cuda")
        out = torch.empty(2 * n * J10, dtype=dt, device="cuda")

        def check_fwd(o, tol=tol):
            assert _block_err(o, ref_d, n) <= tol

        k = _check_plans(False, n, filt, dt, lambda: D.modwt_fwd(xd, J10, filt, out=out),
                         check_fwd)
        print("forward plans checked:", k, flush=True)
        del out, xd
        torch.cuda.empty_cache()

    # inverse: random (inconsistent) coefficients, then thresholded ones
    w64 = _x(2 * n * J10, 7 + filt)
    for dt, nt, tol in ((torch.float64, np.float64, TOL64), (torch.float32, np.float32, TOL32)):
        # the oracle sees the values the kernels see (the rounded inputs)
        w = w64.astype(nt).astype(np.float64)
        lams = _lams(J10).astype(nt).astype(np.float64)
        want = torch.from_numpy(O.modwt_inv(w, n, J10, filt, omp=True)).cuda()
        wt = O.threshold_levels_(w.copy(), n, J10, 1, filt, lams, W.SOFT)
        want_t = torch.from_numpy(O.modwt_inv(wt, n, J10, filt, omp=True)).cuda()
        del wt
        wd = torch.tensor(w, dtype=dt, device="cuda")
        del w
        xr = torch.empty(n, dtype=dt, device="cuda")
        th = _thr(J10, lams)

        def check_inv(o, tol=tol):
            assert _block_err(o, want, n) <= tol

        def check_inv_t(o, tol=tol):
            assert _block_err(o, want_t, n) <= tol

        _check_plans(True, n, filt, dt, lambda: D.modwt_inv(wd, n, J10, filt, out=xr),
                     check_inv)
        _check_plans(True, n, filt, dt,
                     lambda: D.modwt_inv(wd, n, J10, filt, thresh=th, out=xr), check_inv_t)
        del wd, xr
        torch.cuda.empty_cache()


def test_headline_tuned_plan():
    """bench.py tunes the headline shape (LA8, J = 10, n = 2^24, FP64) before
    timing; the plan it then runs is one of the candidates checked above,
    and gives the oracle's forward and inverse."""
    n, f = 1 << 24, 4
    cands = {inv: D.modwt_plan_candidates(inv, n, J10, f) for inv in (False, True)}
    D.modwt_tune(n, J10, f)
    try:
        for inverse in (False, True):
            plan = D.modwt_plan(inverse, n, J10, f)
            print("tuned", "inv" if inverse else "fwd", plan, flush=True)
            assert plan in cands[inverse]
        x = _x(n, 77)
        ref = torch.from_numpy(O.modwt_fwd(x, J10, f, omp=True)).cuda()
        xm = D.modwt_fwd(torch.from_numpy(x).cuda(), J10, f)
        assert _block_err(xm, ref, n) <= TOL64
        del xm
        w = _x(2 * n * J10, 78)
        want = torch.from_numpy(O.modwt_inv(w, n, J10, f, omp=True)).cuda()
        got = D.modwt_inv(torch.from_numpy(w).cuda(), n, J10, f)
        assert _block_err(got, want, n) <= TOL64
    finally:
        for inverse in (False, True):
            D.modwt_set_plan(inverse, n, J10, f, torch.float64, None)


def _dwt_level_err(got, ref, J):
    """Max over DWT levels (and the scaling block) of the norm-wise error."""
    n = ref.numel()
    errs = []
    for j in range(1, J + 1):
        s = 1 << (j - 1)
        a, b = got[s::2 * s].double(), ref[s::2 * s]
        errs.append(((a - b).abs().max() / b.abs().max().clamp_min(1e-300)).item())
    a, b = got[:: 1 << J].double(), ref[:: 1 << J]
    errs.append(((a - b).abs().max() / b.abs().max().clamp_min(1e-300)).item())
    return max(errs)


@pytest.mark.parametrize("lg", [20, 24])
@pytest.mark.parametrize("filt", FILTERS)
def test_dwt_full_depth(filt, lg):
    n = 1 << lg
    J = O.levels(n, 0, filt)
    x = _x(n, 300 + lg + filt)
    ref = torch.from_numpy(O.dwt(x, J, filt, 1, omp=True)).cuda()
    w64 = _x(n, 400 + filt)
    lams64 = np.array([0.2 + 0.05 * l for l in range(J)])
    for dt, nt, tol in ((torch.float64, np.float64, TOL64), (torch.float32, np.float32, TOL32)):
        y = D.dwt_fwd(torch.tensor(x, dtype=dt, device="cuda"), J, filt)
        assert _dwt_level_err(y, ref, J) <= tol
        # the oracle sees the values the kernels see: hard thresholding is
        # discontinuous, so a coefficient within FP32 rounding of its
        # threshold must be decided on the same (rounded) numbers
        w = w64.astype(nt).astype(np.float64)
        lams = lams64.astype(nt).astype(np.float64)
        want = torch.from_numpy(O.dwt(w, J, filt, 0, omp=True)).cuda()
        want_t = torch.from_numpy(
            O.dwt(O.threshold_levels_(w.copy(), n, J, 0, filt, lams, W.HARD), J, filt, 0,
                  omp=True)).cuda()
        th = dict(value=0.0, hardness=W.HARD, minlevel=0, maxlevel=J - 1,
                  levels=torch.tensor(lams, dtype=torch.float64, device="cuda"))
        wd = torch.tensor(w, dtype=dt, device="cuda")
        xr = D.dwt_inv(wd, J, filt)
        e = ((xr.double() - want).abs().max() / want.abs().max()).item()
        assert e <= tol
        xt = D.dwt_inv(wd, J, filt, thresh=th)
        e = ((xt.double() - want_t).abs().max() / want_t.abs().max()).item()
        assert e <= tol


# ------------------------------------------- the differentiable-smoothing launch
B5, N5, J5 = 512, 1 << 14, 8
SAMPLE = [0, 1, 77, 255, 256, 301, 510, 511]


def _config5_signals():
    rng = np.random.Generator(np.random.PCG64(55))
    clean = doppler(N5)
    sd = clean.std(ddof=1)
    snr = rng.uniform(1.0, 10.0, size=B5)
    return clean[None, :] + (sd / snr)[:, None] * rng.standard_normal((B5, N5))


def _fwd_adj_oracle(cot, n, J, f):
    """sum_j 2^j R_j(S_j = cot S_j, D_j = cot D_j, finer D = 0)."""
    x = np.zeros(n)
    for j in range(1, J + 1):
        v = np.zeros(2 * n * j)
        v[2 * n * (j - 1): 2 * n * j] = cot[2 * n * (j - 1): 2 * n * j]
        x += 2.0 ** j * O.modwt_inv(v, n, j, f)
    return x


def _inv_adj_oracle(g, n, J, f):
    """The forward of g with D_j scaled by 2^-j, S_J by 2^-J, other S zero."""
    w = O.modwt_fwd(g, J, f)
    out = np.zeros_like(w)
    for j in range(1, J + 1):
        out[2 * n * (j - 1) + n: 2 * n * j] = w[2 * n * (j - 1) + n: 2 * n * j] / 2.0 ** j
    out[2 * n * (J - 1): 2 * n * (J - 1) + n] = w[2 * n * (J - 1): 2 * n * (J - 1) + n] / 2.0 ** J
    return out


def _row_err(got, want):
    return float(np.max(np.abs(got - want)) / np.max(np.abs(want)))


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_config5_many_signal_launch(dt):
    f = W.LA8
    tol = TOL32 if dt == torch.float32 else TOL64
    esz = W.WH_F32 if dt == torch.float32 else W.WH_F64
    xs = _config5_signals()
    x = torch.tensor(xs, dtype=dt, device="cuda")
    xref = x.double().cpu().numpy()  # the oracle sees the values the GPU sees
    for inverse in (False, True):
        plan = D.modwt_plan(inverse, N5, J5, f, dt, count=B5)
        print("config 5", dt, "inv" if inverse else "fwd", plan, flush=True)
        # every level in one fused launch over all 512 signals
        assert plan.count("{") == 1 and " g=8 " in plan
    w = torch.empty(B5, 2 * N5 * J5, dtype=dt, device="cuda")
    W.check(W.lib().wh_modwt_fwd_many(x.data_ptr(), w.data_ptr(), N5, B5, J5, f, esz,
                                      D._stream()), "fwd_many")
    lams = np.array([0.02 + 0.01 * l for l in range(J5)])
    lam_d = torch.tensor(lams, device="cuda")
    th = W.make_thresh(True, W.SOFT, 0, J5 - 1, 0.0, lam_d.data_ptr())
    y = torch.empty(B5, N5, dtype=dt, device="cuda")
    W.check(W.lib().wh_modwt_inv_many(w.data_ptr(), y.data_ptr(), N5, B5, J5, f, esz,
                                      ctypes.byref(th), D._stream()), "inv_many")
    # adjoints on cotangents drawn at random
    rng = np.random.Generator(np.random.PCG64(9))
    gw = torch.tensor(rng.standard_normal((B5, 2 * N5 * J5)), dtype=dt, device="cuda")
    gx = torch.empty(B5, N5, dtype=dt, device="cuda")
    W.check(W.lib().wh_modwt_fwd_adj(gw.data_ptr(), gx.data_ptr(), N5, B5, J5, f, esz,
                                     D._stream()), "fwd_adj")
    gy = torch.tensor(rng.standard_normal((B5, N5)), dtype=dt, device="cuda")
    gc = torch.empty(B5, 2 * N5 * J5, dtype=dt, device="cuda")
    W.check(W.lib().wh_modwt_inv_adj(gy.data_ptr(), gc.data_ptr(), N5, B5, J5, f, esz,
                                     D._stream()), "inv_adj")
    torch.cuda.synchronize()
    del lam_d
    for b in SAMPLE:
        ref = O.modwt_fwd(xref[b], J5, f)
        assert _row_err(w[b].double().cpu().numpy(), ref) <= tol, b
        c = w[b].double().cpu().numpy()
        want = O.modwt_inv(O.threshold_levels_(c.copy(), N5, J5, 1, f, lams, W.SOFT), N5, J5, f)
        assert _row_err(y[b].double().cpu().numpy(), want) <= tol, b
        cot = gw[b].double().cpu().numpy()
        assert _row_err(gx[b].double().cpu().numpy(), _fwd_adj_oracle(cot, N5, J5, f)) <= tol, b
        g = gy[b].double().cpu().numpy()
        assert _row_err(gc[b].double().cpu().numpy(), _inv_adj_oracle(g, N5, J5, f)) <= tol, b


# ------------------------------------------------------- wrapper checks
def test_wrappers_reject_short_or_mistyped_buffers():
    """The tensor wrappers check sizes and dtypes before handing raw
    pointers to the kernels."""
    n, J = 1024, 4
    x = torch.randn(n, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        D.modwt_fwd(x, J, "LA8", out=torch.empty(2 * n * J - 1, dtype=torch.float64,
                                                 device="cuda"))
    with pytest.raises(ValueError):
        D.modwt_fwd(x, J, "LA8", out=torch.empty(2 * n * J, dtype=torch.float32, device="cuda"))
    w = D.modwt_fwd(x, J, "LA8")
    with pytest.raises(ValueError):
        D.modwt_inv(w, n, J, "LA8", out=torch.empty(n // 2, dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError):
        D.dwt_inv(x, J, "D4", out=torch.empty(n, dtype=torch.float32, device="cuda"))
    with pytest.raises(ValueError):
        D.threshold_(w[: n * J], n, J, 1, "LA8", 0.1, 1, 0, J - 1)
    with pytest.raises(ValueError):
        D.univ_thresh(w[: n], n, J, 1, "LA8")
    with pytest.raises(ValueError):
        D.interp_mse(x, x[: n // 2 - 1].clone(), x[: n // 2].clone())
    with pytest.raises(ValueError):
        D.interp_mse(x, x[: n // 2].float(), x[: n // 2].clone())
    with pytest.raises(ValueError):
        D.modwt_inv(w, n, J, "LA8", thresh=dict(levels=torch.ones(J - 1, device="cuda")))

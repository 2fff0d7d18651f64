"""GPU parity of the batched smoothing pipeline (wh_batch_*).

Every signal of a batch must equal (a) the CPU oracle's universal-threshold
smoothing of that signal (reference src/rfunctions.cu:82-87) within 1e-12
norm-wise in FP64 (thresholds within 1e-12 relative: the GPU transforms
contract products into FMAs, so a MAD can differ in the last bit), and
(b) the single-signal GPU path (wh_smooth, UNIV) -- the same per-element
arithmetic and the same exact order statistics, so bit for bit.  FP32 runs
are held to 1e-5 against the FP64 oracle.  Launch shapes chosen by the
tuner and graph replays must reproduce the eager result bit for bit.
"""
import numpy as np
import pytest
import torch

import oracle_bind as O
from wavecuda_amd import _lib as W
from wavecuda_amd import device as D
from wavecuda_amd.batch import BatchPlan
from wavecuda_amd.signals import batch_workload, noisy_doppler

pytestmark = pytest.mark.gpu


def same_t(a, b):
    return abs(a - b) <= 1e-12 * abs(b)


def mixed_table(kmin=4, kmax=12, count=96, seed=5):
    """Every filter x transform over a range of lengths, levels varied."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lens, filters, ttypes, levels = [], [], [], []
    for i in range(count):
        f = 1 + i % 4
        tt = (i // 4) % 2
        L = W.FILTER_LENGTHS[f]
        k = int(rng.integers(max(kmin, L.bit_length()), kmax + 1))
        n = 1 << k
        jmax = O.levels(n, 0, f)
        if tt == 1:
            jmax = min(jmax, 8)
        J = int(rng.integers(1, jmax + 1)) if i % 3 else (0 if tt == 0 else jmax)
        lens.append(n)
        filters.append(f)
        ttypes.append(tt)
        levels.append(J)
    lens = np.array(lens, np.int32)
    x = np.concatenate([noisy_doppler(int(n), 2.0 + (i % 7), seed=100 + i)[0]
                        for i, n in enumerate(lens)])
    return x, lens, np.array(filters, np.int32), np.array(ttypes, np.int32), \
        np.array(levels, np.int32)


def oracle_smooth(xs, n, f, tt, lv, hard):
    J = O.levels(n, int(lv), int(f))
    y, t, rc = O.smooth(xs, J, int(tt), int(f), W.UNIV, None, hard, 0, J - 1)
    assert rc == 0
    return y, t


@pytest.mark.parametrize("hard", [W.SOFT, W.HARD])
def test_batch_matches_oracle_and_single(hard):
    x, lens, filters, ttypes, levels = mixed_table()
    plan = BatchPlan(lens, filters, ttypes, levels, hardness=hard)
    print("plan", plan.info(), flush=True)
    xg = torch.tensor(x, device="cuda")
    y, t = plan.smooth(xg)
    torch.cuda.synchronize()
    yc = y.cpu().numpy()
    tc = t.cpu().numpy()
    o = plan.offsets
    for i in range(len(lens)):
        n, f, tt, lv = int(lens[i]), int(filters[i]), int(ttypes[i]), int(levels[i])
        xs = x[o[i]:o[i + 1]]
        ys, ts = oracle_smooth(xs, n, f, tt, lv, hard)
        case = f"signal {i}: n={n} filter={f} ttype={tt} levels={lv}"
        assert O.relerr(yc[o[i]:o[i + 1]], ys) <= 1e-12, case
        assert same_t(tc[i], ts), case
        ysg, tsg = D.smooth(xg[o[i]:o[i + 1]].clone(), lv, tt, f, W.UNIV, hardness=hard)
        assert tsg == tc[i], case
        assert torch.equal(ysg, y[o[i]:o[i + 1]]), case


def test_batch_f32_against_f64_oracle():
    x, lens, filters, ttypes, levels = mixed_table(kmin=6, count=48, seed=9)
    plan = BatchPlan(lens, filters, ttypes, levels, dtype=torch.float32)
    xg = torch.tensor(x, dtype=torch.float32, device="cuda")
    y, t = plan.smooth(xg)
    yc = y.double().cpu().numpy()
    o = plan.offsets
    for i in range(len(lens)):
        xs = xg[o[i]:o[i + 1]].double().cpu().numpy()  # the FP32 input, exactly
        ys, ts = oracle_smooth(xs, int(lens[i]), filters[i], ttypes[i], levels[i], W.SOFT)
        assert O.relerr(yc[o[i]:o[i + 1]], ys) <= 1e-5, f"signal {i}"
        assert abs(t[i].item() - ts) <= 1e-5 * ts, f"signal {i}"


def test_batch_tune_and_graph_replays():
    x, lens, filters, ttypes, levels = mixed_table(count=64, seed=11)
    plan = BatchPlan(lens, filters, ttypes, levels)
    xg = torch.tensor(x, device="cuda")
    y0, t0 = plan.smooth(xg)
    y1, t1 = plan.tune(xg)
    info = plan.info()
    print("tuned", info, flush=True)
    assert info["tuned"]
    assert torch.equal(y0, y1) and torch.equal(t0, t1)
    # a second plan of the same composition starts from the tuning table
    assert BatchPlan(lens, filters, ttypes, levels).info()["tuned"]
    yg = torch.zeros_like(xg)
    tg = torch.zeros_like(t0)
    plan.capture(xg, yg, tg)
    for r in range(3):
        yg.zero_()
        tg.zero_()
        plan.replay()
        torch.cuda.synchronize()
        assert torch.equal(yg, y0), f"replay {r}"
        assert torch.equal(tg, t0), f"replay {r}"
        # eager work between replays
        ye, _ = plan.smooth(xg)
        torch.cuda.synchronize()
        assert torch.equal(ye, y0)


def test_batch_edge_sizes():
    # shortest legal signals, one-level DWTs (MAD over x[1..n-1]), deepest
    # MODWT, a single long DWT that runs levels through scratch
    lens = np.array([2, 4, 8, 16, 256, 1 << 15, 64, 2], np.int32)
    filters = np.array([1, 2, 4, 4, 3, 2, 1, 1], np.int32)
    ttypes = np.array([0, 0, 1, 1, 1, 0, 1, 1], np.int32)
    levels = np.array([1, 1, 1, 2, 5, 0, 6, 1], np.int32)
    x = np.concatenate([noisy_doppler(int(n), 4.0, seed=7 + i)[0] for i, n in enumerate(lens)])
    plan = BatchPlan(lens, filters, ttypes, levels)
    y, t = plan.smooth(torch.tensor(x, device="cuda"))
    yc = y.cpu().numpy()
    o = plan.offsets
    for i in range(len(lens)):
        ys, ts = oracle_smooth(x[o[i]:o[i + 1]], int(lens[i]), filters[i], ttypes[i], levels[i],
                               W.SOFT)
        assert O.relerr(yc[o[i]:o[i + 1]], ys) <= 1e-12, f"signal {i}"
        assert same_t(t[i].item(), ts), f"signal {i}"


def test_batch_rejects_bad_tables():
    ok = dict(lens=[256], filters=[4], ttypes=[1], levels=[4])
    with pytest.raises(W.WaveHipError, match="depth above 8"):
        BatchPlan(lens=[1024], filters=[1], ttypes=[1], levels=[9])
    with pytest.raises(W.WaveHipError):
        BatchPlan(**{**ok, "levels": [7]})  # beyond the filter's maximum
    with pytest.raises(W.WaveHipError):
        BatchPlan(**{**ok, "lens": [100]})
    with pytest.raises(W.WaveHipError):
        BatchPlan(**{**ok, "filters": [5]})
    with pytest.raises(W.WaveHipError):
        BatchPlan(**{**ok, "ttypes": [2]})
    with pytest.raises(W.WaveHipError):
        BatchPlan(**{**ok, "lens": [4]})  # shorter than the filter
    plan = BatchPlan(**ok)
    with pytest.raises(W.WaveHipError):
        plan.smooth(torch.zeros(255, dtype=torch.float64, device="cuda"))


def test_batch_workload_sample():
    """The 8192-signal mixed workload: every signal against the single-signal
    GPU path would take minutes of launches, so a seeded sample of signals is
    checked against the oracle and the single-signal path; all signals are
    checked for finiteness and the thresholds for positivity."""
    x, lens, filters, ttypes, levels = batch_workload()
    plan = BatchPlan(lens, filters, ttypes, levels)
    xg = torch.tensor(x, device="cuda")
    y, t = plan.smooth(xg)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(y).all()) and bool((t > 0).all())
    o = plan.offsets
    pick = np.random.Generator(np.random.PCG64(3)).choice(len(lens), 24, replace=False)
    for i in sorted(pick):
        n, f, tt, lv = int(lens[i]), int(filters[i]), int(ttypes[i]), int(levels[i])
        ys, ts = oracle_smooth(x[o[i]:o[i + 1]], n, f, tt, lv, W.SOFT)
        yi = y[o[i]:o[i + 1]].cpu().numpy()
        case = f"signal {i}: n={n} filter={f} ttype={tt} levels={lv}"
        assert O.relerr(yi, ys) <= 1e-12, case
        assert same_t(t[i].item(), ts), case
        ysg, tsg = D.smooth(xg[o[i]:o[i + 1]].clone(), lv, tt, f, W.UNIV)
        assert torch.equal(ysg, y[o[i]:o[i + 1]]) and tsg == t[i].item(), case


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batch_workload_all_signals_equal_single(dtype):
    """Every one of the 8192 signals of the mixed workload equals the
    single-signal GPU smoothing (wh_smooth, UNIV) bit for bit, thresholds
    included."""
    x, lens, filters, ttypes, levels = batch_workload()
    plan = BatchPlan(lens, filters, ttypes, levels, dtype=dtype)
    xg = torch.tensor(x, dtype=dtype, device="cuda")
    y, t = plan.smooth(xg)
    tc = t.cpu().numpy()
    o = plan.offsets
    bad = []
    for i in range(len(lens)):
        yi, ti = D.smooth(xg[o[i]:o[i + 1]], int(levels[i]), int(ttypes[i]), int(filters[i]),
                          W.UNIV)
        if not (torch.equal(yi, y[o[i]:o[i + 1]]) and ti == tc[i]):
            bad.append(i)
            if len(bad) > 5:
                break
    assert not bad, f"signals differing from the single-signal path: {bad}"


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batch_selection_routes(dtype):
    """Detail coefficients that stress the workgroup's exact median / MAD
    (noise; noise with one moderate or one huge spike, which spreads the
    values' exponents; integer-valued signals, whose details tie; a
    constant and a signed-zero signal, whose details are (nearly) all
    equal) at 2^10, 2^14 and 2^16 samples must give thresholds and outputs
    bit for bit equal to the single-signal path, and on the noise-like
    signals the oracle's thresholds."""
    rng = np.random.Generator(np.random.PCG64(77))
    sigs, kinds = [], []
    for n in (1 << 10, 1 << 14, 1 << 16):
        noise = rng.standard_normal(n)
        spike = noise.copy()
        spike[n // 3] = 60.0
        huge = noise.copy()
        huge[n // 5] = 1.0e6
        ties = np.round(rng.standard_normal(n))
        const = np.full(n, 3.0)
        zeros = np.where(rng.random(n) < 0.5, -0.0, 0.0)
        for name, s in (("noise", noise), ("spike", spike), ("huge", huge), ("ties", ties),
                        ("const", const), ("zeros", zeros)):
            sigs.append(s)
            kinds.append((name, n))
    count = len(sigs)
    lens = np.array([k[1] for k in kinds], np.int32)
    filters = np.array([(W.LA8, W.D4, W.HAAR, W.C6)[i % 4] for i in range(count)], np.int32)
    ttypes = np.array([(i // 2) % 2 for i in range(count)], np.int32)
    levels = np.array([0 if t == 0 else 6 for t in ttypes], np.int32)
    x = np.concatenate(sigs)
    plan = BatchPlan(lens, filters, ttypes, levels, dtype=dtype)
    xg = torch.tensor(x, dtype=dtype, device="cuda")
    y, t = plan.smooth(xg)
    tc = t.cpu().numpy()
    o = plan.offsets
    for i in range(count):
        case = f"signal {i}: {kinds[i][0]} n={kinds[i][1]} filter={filters[i]} ttype={ttypes[i]}"
        yi, ti = D.smooth(xg[o[i]:o[i + 1]], int(levels[i]), int(ttypes[i]), int(filters[i]),
                          W.UNIV)
        assert torch.equal(yi, y[o[i]:o[i + 1]]), case
        assert ti == tc[i], case
        if dtype == torch.float64 and kinds[i][0] in ("noise", "spike", "huge"):
            _, ts = oracle_smooth(x[o[i]:o[i + 1]], int(lens[i]), filters[i], ttypes[i],
                                  levels[i], W.SOFT)
            assert same_t(tc[i], ts), case


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batch_every_launch_shape(dtype):
    """Every launch shape the tuner can pick (128, 256 or 512 threads per
    signal; a few workgroups or one per signal) gives the default shape's
    outputs and thresholds bit for bit, and those equal the single-signal
    path's."""
    x, lens, filters, ttypes, levels = mixed_table(count=64, seed=21)
    plan = BatchPlan(lens, filters, ttypes, levels, dtype=dtype)
    xg = torch.tensor(x, dtype=dtype, device="cuda")
    y0, t0 = plan.smooth(xg)
    o = plan.offsets
    for i in range(len(lens)):
        yi, ti = D.smooth(xg[o[i]:o[i + 1]], int(levels[i]), int(ttypes[i]), int(filters[i]),
                          W.UNIV)
        assert torch.equal(yi, y0[o[i]:o[i + 1]]) and ti == t0[i].item(), f"signal {i}"
    for nt in (128, 256, 512):
        for wg in (3, len(lens)):
            plan.set_shape(nt, wg)
            info = plan.info()
            assert info["threads"] == nt and info["workgroups"] == wg
            y, t = plan.smooth(xg)
            case = f"{nt} threads x {wg} workgroups"
            assert torch.equal(y, y0), case
            assert torch.equal(t, t0), case
    with pytest.raises(W.WaveHipError):
        plan.set_shape(384, 3)
    with pytest.raises(W.WaveHipError):
        plan.set_shape(256, len(lens) + 1)

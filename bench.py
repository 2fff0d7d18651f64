"""Benchmark of the wavelet hot path on one MI355X.

Step = one LA8 time-ordered MODWT with J = 10 levels, forward then inverse,
over one n = 2^24 FP64 signal already resident in HBM (white noise).  The
headline value is whole-job throughput in Msamples/s.  Alongside:

  roofline      the dominant kernel (the fused forward MODWT) priced against
                MI355X's 8 TB/s HBM: algorithmic bytes per launch
                (8 (1 + 2J) B per sample: read x once, write S_j and D_j for
                every level) / its average duration, timed with HIP events
                on the stream the kernel runs on;
  batched       8192 mixed signals smoothed by one persistent launch
                (replayed from a HIP graph) against the plain loop of one
                smoothing call per signal, FP64 and FP32, outputs compared
                bit for bit;
  sweep         every filter, DWT full depth and MODWT J = 10, n = 2^20 and
                2^24, forward and inverse (FP64);
  small         D4 DWT forward + inverse at n = 1024 (latency);
  smooth        single-signal universal / CV smoothing at n = 2^18;
  differentiable  per-level thresholds fitted by Adam on 512 x 2^14 FP32
                signals, against CV smoothing;
  host_c_abi    the reference-compatible RgpuTransform on host arrays
                (PCIe copies included), LA8 MODWT J = 10, n = 2^24;
  cpu_baseline  the CPU oracle (the reference's serial CPU path restated)
                timed on this host on a bounded sample of the headline
                workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                       [--no-batch] [--no-extra] [--plain-sample N]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "LA8 DWT/MODWT fwd+inv Msamples/s & %HBM peak; batched vs plain at same tol"


def modwt_bytes_fwd(n, J, esz=8):
    return esz * n * (1 + 2 * J)


def modwt_bytes_inv(n, J, esz=8):
    return esz * n * (J + 2)


def parse_knobs(text: str) -> tuple:
    """'g1max,first_per_cu,deep_per_cu,threads' -> the planner's knob tuple."""
    k = tuple(int(v) for v in text.split(","))
    if len(k) != 4:
        raise SystemExit(f"plan knobs {text!r}: expected four comma-separated integers")
    return k


def plan_knobs(D, inverse: bool, n: int, J: int, filt: int) -> str | None:
    """The knobs (as parse_knobs reads them) of the MODWT plan now in force
    for this shape, found among the tuner's candidates; the plan itself is
    left as it was."""
    current = D.modwt_plan(inverse, n, J, filt)
    cands = D.modwt_plan_candidates(inverse, n, J, filt)  # resets to the default rules
    k = cands.get(current)
    default = D.modwt_plan(inverse, n, J, filt)
    if current != default:
        if k is None:
            raise SystemExit("the MODWT plan in force is not among the tuner's candidates")
        D.modwt_set_plan(inverse, n, J, filt, knobs=k)
    if D.modwt_plan(inverse, n, J, filt) != current:
        raise SystemExit("could not restore the MODWT plan in force")
    return None if k is None else ",".join(str(v) for v in k)


def plan_kernels(plan: str, inverse: bool) -> str:
    """The fused MODWT kernels a plan string launches, in launch order
    (LA8, FP64): modwt_fwd_k<T, filter, R, threads, levels, map, store> /
    modwt_inv_k<T, filter, R, threads, levels, map>."""
    import re

    groups = re.findall(r"\{j0=(\d+) g=(\d+) R=(\d+) NT=(\d+)", plan)
    names = []
    for _, g, r, nt in (reversed(groups) if inverse else groups):
        if inverse:
            names.append(f"modwt_inv_k<double, 4, {r}, {nt}, {g}, {2 if r == '1' else 0}>")
        else:
            names.append(f"modwt_fwd_k<double, 4, {r}, {nt}, {g}, 0, 0>")
    return " then ".join(names)


def traffic_from_profiles():
    """HBM bytes per launch of the forward kernel measured by rocprofv3
    FETCH_SIZE / WRITE_SIZE (scaled by a copy-kernel calibration), if the
    committed summary exists."""
    p = ROOT / "profiles" / "modwt_la8_n2p24_J10_traffic.txt"
    if not p.exists():
        return None
    m = re.search(r"^hbm_bytes_per_launch:\s*([0-9.eE+]+)", p.read_text(), re.M)
    return float(m.group(1)) if m else None


def cpu_baseline(n_sample=1 << 22, J=10, filt=4, omp=False, budget_s=10.0):
    """The oracle's LA8 MODWT fwd+inv on n_sample samples: single-threaded
    (the reference's serial CPU path, restated) or its OpenMP variant on the
    host's threads.  The reference's own CPU build has no LA8 MODWT, so the
    restatement is the baseline for this workload."""
    sys.path.insert(0, str(ROOT / "tests"))
    import numpy as np

    import oracle_bind as O  # test infrastructure: the checker, timed as a baseline

    x = np.random.Generator(np.random.PCG64(1)).standard_normal(n_sample)
    O.modwt_fwd(x[:1024], J if J <= 7 else 7, filt, omp=omp)  # load / warm
    reps, t_total = 0, 0.0
    while t_total < budget_s and reps < (4 if not omp else 16):
        t0 = time.perf_counter()
        xm = O.modwt_fwd(x, J, filt, omp=omp)
        xr = O.modwt_inv(xm, n_sample, J, filt, omp=omp)
        t_total += time.perf_counter() - t0
        reps += 1
        del xm, xr
    rate = reps * n_sample / t_total / 1e6
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)) if omp else 1
    return {
        "value": round(rate, 3),
        "unit": "Msamples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"LA8 MODWT J={J} fwd+inv on n=2^{n_sample.bit_length() - 1} FP64, "
                  f"{reps} reps, {'OpenMP, ' + str(threads) + ' threads' if omp else 'single thread'}"
                  f", oracle/wvoracle.c",
    }


def cpu_config1(budget_s=1.0):
    """The oracle's D4 DWT forward + inverse of the n = 1024 doppler signal
    (9 levels, full depth), single thread, timed in C (one transform is a
    few microseconds)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import numpy as np

    import oracle_bind as O  # test infrastructure: the checker, timed as a baseline
    from wavecuda_amd.signals import noisy_doppler

    x = noisy_doppler(1024, 5.0, seed=10)[0]
    w = np.zeros(1024)
    reps = 1000
    t = O.oracle().wvo_time_dwt_fwd_inv(O._p(x), O._p(w), 1024, 9, 2, reps)
    while t < budget_s:
        reps *= 2
        t = O.oracle().wvo_time_dwt_fwd_inv(O._p(x), O._p(w), 1024, 9, 2, reps)
    return {"value": round(t / reps * 1e6, 3), "unit": "us per fwd+inv", "cores": 1,
            "kind": "port",
            "sample": f"D4 DWT fwd+inv, n=1024, 9 levels, doppler+noise SNR 5, {reps} reps, "
                      "single thread, oracle/wvoracle.c"}


def cpu_config3():
    """The oracle's CPUSmooth on every smoothing case the GPU section runs
    (n = 2^18 doppler + noise, MODWT J = 10, Haar / LA8, universal / CV,
    soft / hard, SNR 5 / 2), single thread, one call each."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_bind as O
    from wavecuda_amd.signals import noisy_doppler

    n = 1 << 18
    rows = []
    for snr in (5.0, 2.0):
        xs, _ = noisy_doppler(n, snr, seed=10)
        for f, fname in ((1, "Haar"), (4, "LA8")):
            for tt, tname in ((1, "univ"), (2, "cv")):
                for hard in (1, 0):
                    t0 = time.perf_counter()
                    _, th, rc = O.smooth(xs, 10, 1, f, tt, None, hard, 0, 9)
                    dt = time.perf_counter() - t0
                    rows.append({"snr": snr, "filter": fname, "thresh": tname,
                                 "hardness": "soft" if hard else "hard",
                                 "ms": round(dt * 1e3, 2), "threshold": th, "rc": rc})
    return {"unit": "ms per smoothing call", "cores": 1, "kind": "port",
            "sample": "all 16 cases of the smoothing section, one call each, single thread, "
                      "oracle/wvoracle.c (the restated CPUSmooth path)",
            "cases": rows}


def cpu_config4(sample=1024):
    """The oracle's universal-threshold smoothing, signal by signal, over the
    first `sample` signals of the 8192-signal mixed batch, single thread."""
    sys.path.insert(0, str(ROOT / "tests"))
    import numpy as np

    import oracle_bind as O
    from wavecuda_amd.signals import batch_workload

    x, lens, filters, ttypes, levels = batch_workload()
    off = np.concatenate([[0], np.cumsum(lens.astype(np.int64))])
    m = min(sample, len(lens))
    t0 = time.perf_counter()
    for i in range(m):
        J = O.levels(int(lens[i]), int(levels[i]), int(filters[i]))
        O.smooth(x[off[i]:off[i + 1]], J, int(ttypes[i]), int(filters[i]), 1, None, 1, 0, J - 1)
    dt = time.perf_counter() - t0
    rate = float(off[m]) / dt / 1e6
    return {"value": round(rate, 3), "unit": "Msamples/s", "cores": 1, "kind": "port",
            "sample": f"first {m} of the 8192 signals ({int(off[m])} samples), universal soft "
                      "smoothing signal by signal, single thread, oracle/wvoracle.c",
            "full_workload_s_est": round(float(off[-1]) / (rate * 1e6), 1)}


def batch_bytes(lens, ttypes, levels, filters, esz):
    """Words a fused smoothing pipeline must move per signal, summed: the
    DWT reads x, writes and re-reads its n coefficients, writes y (4 n);
    the MODWT reads x, writes D_1..D_J and S_J, re-reads them, writes y
    ((2J + 4) n)."""
    from wavecuda_amd import _lib as W

    tot = 0
    for n, tt, lv, f in zip(lens, ttypes, levels, filters):
        n = int(n)
        if tt == 0:
            tot += 4 * n
        else:
            J = W.lib().wh_resolve_levels(n, int(lv), int(f))
            tot += (2 * J + 4) * n
    return tot * esz


def both_transforms(x, lens, filters, ttypes, levels):
    """The batch with every signal twice, once per transform: DWT at full
    depth and MODWT at J = min(8, k - 2)."""
    import numpy as np

    k = np.log2(lens).astype(np.int32)
    off = np.concatenate([[0], np.cumsum(lens.astype(np.int64))])
    xs = np.concatenate([np.concatenate([x[off[i]:off[i + 1]], x[off[i]:off[i + 1]]])
                         for i in range(len(lens))])
    l2 = np.repeat(lens, 2)
    f2 = np.repeat(filters, 2)
    t2 = np.tile(np.array([0, 1], dtype=np.int32), len(lens))
    v2 = np.where(t2 == 1, np.minimum(8, np.repeat(k, 2) - 2), 0).astype(np.int32)
    return xs, l2, f2, t2, v2


def batch_section(steps, warmup, plain_sample, both=False):
    """Batched smoothing of the mixed 8192-signal workload (one persistent
    launch, replayed from a HIP graph) against the plain per-signal GPU
    loop (one universal-threshold smoothing call per signal).  both=True:
    every signal smoothed by both transforms (16384 items)."""
    import numpy as np
    import torch

    from wavecuda_amd import _lib as W
    from wavecuda_amd import device as D
    from wavecuda_amd.batch import BatchPlan
    from wavecuda_amd.signals import batch_workload

    x, lens, filters, ttypes, levels = batch_workload()
    desc = "DWT full depth or MODWT J=min(8,k-2) (one drawn per signal)"
    if both:
        x, lens, filters, ttypes, levels = both_transforms(x, lens, filters, ttypes, levels)
        desc = "both DWT full depth and MODWT J=min(8,k-2) for every signal (16384 items)"
    out = {"workload": "8192 doppler+noise signals, n=2^k k~U{8..16}, filter~U{Haar,D4,C6,LA8}, "
                       f"{desc}, SNR~U[1,10]; forward -> universal soft threshold -> inverse",
           "signals": int(len(lens)), "samples": int(lens.astype(np.int64).sum())}
    for dtype, tag in ((torch.float64, "f64"), (torch.float32, "f32")):
        esz = 8 if dtype == torch.float64 else 4
        plan = BatchPlan(lens, filters, ttypes, levels, dtype=dtype)
        xg = torch.tensor(x, dtype=dtype, device="cuda")
        y, t = plan.tune(xg)
        plan.capture(xg, y, t)
        for _ in range(warmup):
            plan.replay()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(steps)]
        for a, b in ev:
            a.record()
            plan.replay()
            b.record()
        torch.cuda.synchronize()
        ms = min(a.elapsed_time(b) for a, b in ev)
        ms_avg = sum(a.elapsed_time(b) for a, b in ev) / steps
        # plain loop over a leading sample of the signals (in input order),
        # outputs compared bit for bit with the batched ones
        o = plan.offsets
        m = min(plain_sample, len(lens))
        ys = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(m):
            yi, ti = D.smooth(xg[o[i]:o[i + 1]], int(levels[i]), int(ttypes[i]),
                              int(filters[i]), W.UNIV)
            ys.append((yi, ti))
        torch.cuda.synchronize()
        plain_s = time.perf_counter() - t0
        same = all(torch.equal(yi, y[o[i]:o[i + 1]]) and ti == t[i].item()
                   for i, (yi, ti) in enumerate(ys))
        frac_samples = float(o[m]) / float(o[-1])
        plain_full_ms = plain_s * 1e3 / frac_samples if m < len(lens) else plain_s * 1e3
        nbytes = batch_bytes(lens, ttypes, levels, filters, esz)
        info = plan.info()
        out[tag] = {
            "batched_ms": round(ms_avg, 3),
            "batched_ms_min": round(ms, 3),
            "msamples_per_s": round(float(o[-1]) / (ms_avg * 1e-3) / 1e6, 1),
            "hbm_frac": round(nbytes / (ms_avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "pipeline_bytes": nbytes,
            "plain_loop_signals": m,
            "plain_loop_ms": round(plain_s * 1e3, 1),
            "plain_loop_ms_full_est": round(plain_full_ms, 1),
            "speedup_vs_plain": round(plain_full_ms / ms_avg, 1),
            "bitwise_equal_to_plain": bool(same),
            "launch": {k: info[k] for k in ("threads", "workgroups", "lds_bytes", "tuned")},
        }
        del plan, xg, y, t, ys
        torch.cuda.empty_cache()
    return out


def _time_ms(fn, steps, warmup=2):
    """Average milliseconds of fn() over `steps` runs (HIP events on torch's
    current stream, the stream the wrappers launch on)."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def sweep_section(steps):
    """Single-signal transforms, white noise: every filter, DWT at full depth
    and MODWT J = 10, n = 2^20 and 2^24 in FP64 and n = 2^24 in FP32;
    forward and inverse timed separately, priced at their algorithmic
    bytes."""
    import torch

    from wavecuda_amd import _lib as W
    from wavecuda_amd import device as D

    rows = []
    for lg, dt in ((20, torch.float64), (24, torch.float64), (24, torch.float32)):
        n = 1 << lg
        esz = 8 if dt == torch.float64 else 4
        x = torch.randn(n, dtype=dt, device="cuda",
                        generator=torch.Generator(device="cuda").manual_seed(lg))
        for f in (W.HAAR, W.D4, W.C6, W.LA8):
            Jd = D.resolve_levels(n, 0, f)
            y = torch.empty_like(x)
            xr = torch.empty_like(x)
            tf = _time_ms(lambda: D.dwt_fwd(x, Jd, f, out=y), steps)
            ti = _time_ms(lambda: D.dwt_inv(y, Jd, f, out=xr), steps)
            b = 2 * esz * n
            rows.append({"transform": "DWT", "filter": W.FILTER_NAMES[f], "n": n, "levels": Jd,
                         "dtype": "f64" if esz == 8 else "f32",
                         "fwd_ms": round(tf, 4), "inv_ms": round(ti, 4),
                         "fwd_GBs": round(b / tf / 1e6, 1), "inv_GBs": round(b / ti / 1e6, 1),
                         "msamples_per_s": round(n / (tf + ti) / 1e3, 1)})
            del y, xr
            J = 10
            xm = torch.empty(2 * n * J, dtype=dt, device="cuda")
            xr = torch.empty_like(x)
            tf = _time_ms(lambda: D.modwt_fwd(x, J, f, out=xm), steps)
            ti = _time_ms(lambda: D.modwt_inv(xm, n, J, f, out=xr), steps)
            bf, bi = modwt_bytes_fwd(n, J, esz), modwt_bytes_inv(n, J, esz)
            rows.append({"transform": "MODWT", "filter": W.FILTER_NAMES[f], "n": n, "levels": J,
                         "dtype": "f64" if esz == 8 else "f32",
                         "fwd_ms": round(tf, 4), "inv_ms": round(ti, 4),
                         "fwd_GBs": round(bf / tf / 1e6, 1), "inv_GBs": round(bi / ti / 1e6, 1),
                         "msamples_per_s": round(n / (tf + ti) / 1e3, 1)})
            del xm, xr
        del x
        torch.cuda.empty_cache()
    return rows


def small_section(steps):
    """D4 DWT forward + inverse, full depth (9 levels), n = 1024 doppler +
    noise at SNR 5: a launch-latency case, reported in microseconds."""
    import torch

    from wavecuda_amd import device as D
    from wavecuda_amd.signals import noisy_doppler

    x = torch.tensor(noisy_doppler(1024, 5.0, seed=10)[0], device="cuda")
    y = torch.empty_like(x)
    xr = torch.empty_like(x)

    def fi():
        D.dwt_fwd(x, 0, "D4", out=y)
        D.dwt_inv(y, 0, "D4", out=xr)
    ms = _time_ms(fi, steps * 10, warmup=10)
    # the same pair recorded ten times into a HIP graph and replayed: the
    # GPU time without the per-call launch overhead of the host
    reps = 10
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fi()
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(reps):
            fi()
    ms_g = _time_ms(g.replay, max(5, steps), warmup=3) / reps
    ok = bool(torch.allclose(xr, x, rtol=0, atol=1e-12 * x.abs().max().item()))
    return {"workload": "D4 DWT fwd+inv, n=1024, 9 levels, FP64", "us_per_fwd_inv": round(ms * 1e3, 2),
            "us_per_fwd_inv_graph": round(ms_g * 1e3, 2), "graph_reconstruction_ok": ok}


def levels_section(steps):
    """Coefficient layouts at the headline size (LA8, n = 2^24, J = 10, FP64):
    time order <-> packet order (every value read and written once: 16 B per
    value) and WSTtoDT's label columns (13 B written per coefficient)."""
    import torch

    from wavecuda_amd import device as D

    n, J = 1 << 24, 10
    m = 2 * n * J
    g = torch.Generator(device="cuda").manual_seed(1)
    to = torch.randn(m, dtype=torch.float64, device="cuda", generator=g)
    po = torch.empty_like(to)
    back = torch.empty_like(to)
    t_to = _time_ms(lambda: D.modwt_convert(to, n, J, "LA8", True, out=po), steps)
    t_po = _time_ms(lambda: D.modwt_convert(po, n, J, "LA8", False, out=back), steps)
    ok = bool(torch.equal(back, to))
    del to, po, back
    t_lab = _time_ms(lambda: D.coeff_labels(n, J, 1, "LA8"), steps)
    b = 16.0 * m
    return {"workload": "LA8 MODWT J=10 coefficients, n=2^24, FP64 (2nJ = 335.5 M values)",
            "to_po_ms": round(t_to, 4), "to_po_GBs": round(b / (t_to * 1e-3) / 1e9, 1),
            "po_to_ms": round(t_po, 4), "po_to_GBs": round(b / (t_po * 1e-3) / 1e9, 1),
            "hbm_frac": round(b / (max(t_to, t_po) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "round_trip_exact": ok,
            "labels_ms": round(t_lab, 4),
            "labels_GBs": round(13.0 * m / (t_lab * 1e-3) / 1e9, 1)}


def smooth_section():
    """Single-signal smoothing, n = 2^18 doppler + noise, MODWT J = 10, Haar
    and LA8, soft and hard, universal and cross-validated thresholds, SNR 5
    and 2: wall time of one synchronous call and the MSE against the clean
    signal."""
    import torch

    from wavecuda_amd import _lib as W
    from wavecuda_amd import device as D
    from wavecuda_amd.signals import noisy_doppler

    rows = []
    n = 1 << 18
    for snr in (5.0, 2.0):
        xs, clean = noisy_doppler(n, snr, seed=10)
        x = torch.tensor(xs, device="cuda")
        cl = torch.tensor(clean, device="cuda")
        for f in (W.HAAR, W.LA8):
            for tt, tname in ((W.UNIV, "univ"), (W.CV, "cv")):
                for hard in (W.SOFT, W.HARD):
                    D.smooth(x, 10, W.MODWT, f, tt, hardness=hard)  # warm
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    y, t = D.smooth(x, 10, W.MODWT, f, tt, hardness=hard)
                    wall = time.perf_counter() - t0
                    rows.append({"snr": snr, "filter": W.FILTER_NAMES[f], "thresh": tname,
                                 "hardness": "soft" if hard == W.SOFT else "hard",
                                 "ms": round(wall * 1e3, 3), "threshold": t,
                                 "mse_vs_clean": torch.mean((y - cl) ** 2).item()})
    return rows


def host_section():
    """The reference-compatible .C entry point RgpuTransform on host arrays
    (what R's .C call does): LA8 MODWT J = 10 forward then inverse of
    n = 2^24 FP64, PCIe copies included (x in, 2 n J coefficients out, and
    back for the inverse)."""
    import ctypes as C

    import numpy as np

    from wavecuda_amd import _lib as W

    n, J = 1 << 24, 10
    x = np.random.Generator(np.random.PCG64(3)).standard_normal(n)
    xm = np.zeros(2 * n * J)
    xr = np.zeros(n)
    dp = C.POINTER(C.c_double)

    def ip(v):
        return C.byref(C.c_int(v))

    def call(xa, xma, sense):
        W.lib().RgpuTransform(xa.ctypes.data_as(dp), xma.ctypes.data_as(dp), ip(n), ip(sense),
                              ip(J), ip(1), ip(W.LA8), ip(8))
    # one untimed pair first: it grows the library's device buffers and
    # touches the host arrays' pages (a caller reusing its arrays, as here,
    # pays neither again; first-touch of fresh arrays is the caller's cost)
    call(x, xm, 1)
    call(xr, xm, 0)
    fwd, inv = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        call(x, xm, 1)
        t1 = time.perf_counter()
        call(xr, xm, 0)
        t2 = time.perf_counter()
        fwd.append(t1 - t0)
        inv.append(t2 - t1)
    err = float(np.max(np.abs(xr - x)) / np.max(np.abs(x)))
    tf, ti = sorted(fwd)[1], sorted(inv)[1]
    return {"workload": "RgpuTransform (host arrays, as R's .C passes them), LA8 MODWT J=10, "
                        "n=2^24 FP64; median of 3 forward+inverse pairs",
            "fwd_ms": round(tf * 1e3, 1), "inv_ms": round(ti * 1e3, 1),
            "msamples_per_s_fwd_plus_inv": round(n / (tf + ti) / 1e6, 1),
            "reconstruction_relerr": err}


def diff_section(steps=300):
    """Per-level soft thresholds fitted by Adam: 512 doppler + noise signals
    (SNR ~ U[1, 10]) of n = 2^14 in FP32, LA8 MODWT J = 8, loss = mean
    squared error of the smoothed signals against the clean ones, lambda
    initialised to the universal threshold, 300 Adam steps at lr 1e-2.
    Compared with cross-validated smoothing (one CV threshold per signal)."""
    import numpy as np
    import torch

    from wavecuda_amd import _lib as W
    from wavecuda_amd import autograd as A
    from wavecuda_amd import device as D
    from wavecuda_amd.signals import doppler

    B, n, J = 512, 1 << 14, 8
    rng = np.random.Generator(np.random.PCG64(55))
    clean = doppler(n)
    sd = clean.std(ddof=1)
    snr = rng.uniform(1.0, 10.0, size=B)
    xs = clean[None, :] + (sd / snr)[:, None] * rng.standard_normal((B, n))
    x = torch.tensor(xs, dtype=torch.float32, device="cuda")
    cl = torch.tensor(np.broadcast_to(clean, (B, n)).copy(), dtype=torch.float32, device="cuda")
    A.fit_level_thresholds(x, cl, J, W.LA8, steps=3, lr=1e-2)  # untimed warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lam, losses = A.fit_level_thresholds(x, cl, J, W.LA8, steps=steps, lr=1e-2)
    torch.cuda.synchronize()
    fit_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    mse_cv = []
    for b in range(B):
        y, _ = D.smooth(x[b], J, W.MODWT, W.LA8, W.CV)
        mse_cv.append(torch.mean((y - cl[b]) ** 2).item())
    cv_s = time.perf_counter() - t0
    return {"workload": "512 x n=2^14 FP32 doppler+noise (SNR~U[1,10]), LA8 MODWT J=8, "
                        "per-level soft thresholds, 300 Adam steps lr 1e-2",
            "ms_per_adam_step": round(fit_s * 1e3 / steps, 3),
            "mse_universal_init": losses[0], "mse_fitted": losses[-1],
            "mse_cv_gpusmooth": float(np.mean(mse_cv)),
            "cv_smooth_ms_per_signal": round(cv_s * 1e3 / B, 3),
            "lambda": [round(v, 5) for v in lam.cpu().tolist()]}


def copy_calibration(nbytes=1 << 30, reps=4):
    """A plain device-to-device copy of nbytes, for scaling the rocprofv3
    FETCH_SIZE / WRITE_SIZE counters: the copy moves exactly nbytes in and
    nbytes out."""
    import torch

    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    ms = _time_ms(lambda: b.copy_(a), reps, warmup=1)
    print(json.dumps({"copy_bytes": nbytes, "reps": reps, "ms": round(ms, 4),
                      "GBs_read_plus_write": round(2 * nbytes / ms / 1e6, 1)}))


def summary(out):
    """The headline figures again, compact, at the end of the line (a
    reader of the line's tail sees them)."""
    s = {"msamples_per_s": out["value"], "ms_per_step": out["ms_per_step"],
         "step_hbm_frac": out["step_hbm_frac"], "fwd_frac": out["roofline"]["frac"],
         "inv_frac": out["inverse_kernel"]["frac"]}
    for key, tag in (("batched", "batch"), ("batched_both_transforms", "batch_both")):
        b = out.get(key)
        if b:
            for dt in ("f64", "f32"):
                s[f"{tag}_{dt}_ms"] = b[dt]["batched_ms"]
                s[f"{tag}_{dt}_hbm_frac"] = b[dt]["hbm_frac"]
                s[f"{tag}_{dt}_speedup_vs_plain"] = b[dt]["speedup_vs_plain"]
                s[f"{tag}_{dt}_bitwise_equal"] = b[dt]["bitwise_equal_to_plain"]
    if "host_c_abi" in out:
        s["host_c_abi_msamples_per_s"] = out["host_c_abi"]["msamples_per_s_fwd_plus_inv"]
    if "small" in out:
        s["d4_n1024_us"] = out["small"]["us_per_fwd_inv"]
        s["d4_n1024_us_graph"] = out["small"]["us_per_fwd_inv_graph"]
    if "cpu_baseline" in out:
        s["cpu_msamples_per_s"] = out["cpu_baseline"]["value"]
        s["cpu_openmp_msamples_per_s"] = out["cpu_baseline_openmp"]["value"]
        c = out["cpu_baseline_configs"]
        s["cpu_d4_n1024_us"] = c["d4_dwt_n1024"]["value"]
        s["cpu_batch_msamples_per_s"] = c["batch_8192"]["value"]
    return s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--log2n", type=int, default=24)
    ap.add_argument("--levels", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="use the planner's default MODWT plans instead of tuning them first")
    ap.add_argument("--fwd-knobs", default=None,
                    help="set the forward MODWT plan instead of tuning it: "
                         "'g1max,first_per_cu,deep_per_cu,threads' as plan.knobs reports them")
    ap.add_argument("--inv-knobs", default=None,
                    help="set the inverse MODWT plan instead of tuning it (same form)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the sweep, small, smoothing and differentiable sections")
    ap.add_argument("--plain-sample", type=int, default=8192,
                    help="signals of the batch run through the plain per-signal loop")
    ap.add_argument("--cpu-batch-sample", type=int, default=1024,
                    help="signals of the batch smoothed by the CPU oracle for its baseline")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch

    from wavecuda_amd import _lib as W
    from wavecuda_amd import device as D

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    W.check(W.lib().wh_init(0), "wh_init")
    if args.calibrate_copy:
        copy_calibration()
        return
    n = 1 << args.log2n
    J = args.levels
    filt = W.LA8
    g = torch.Generator(device="cuda").manual_seed(1234)
    x = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    xmod = torch.empty(2 * n * J, dtype=torch.float64, device="cuda")
    xr = torch.empty(n, dtype=torch.float64, device="cuda")

    set_plans = args.fwd_knobs is not None or args.inv_knobs is not None
    if set_plans:
        # a plan fixed from the command line (counter runs repeat the plan a
        # timed run chose: the tuner's timings under counter collection differ)
        for inverse, text in ((False, args.fwd_knobs), (True, args.inv_knobs)):
            if text is not None:
                D.modwt_set_plan(inverse, n, J, filt, knobs=parse_knobs(text))
    elif not args.no_tune:
        D.modwt_tune(n, J, filt)  # one-time plan selection, before any timing

    def step():
        D.modwt_fwd(x, J, filt, out=xmod)
        D.modwt_inv(xmod, n, J, filt, out=xr)

    # the reconstruction check runs before the warm-up: the card idles for
    # tens of ms while it runs (first calls of torch kernels), and after such
    # a pause the inverse runs up to 25 % slower for some 20 steps, so the
    # warm-up steps lead straight into the timed ones
    step()
    torch.cuda.synchronize()
    err = ((xr - x).abs().max() / x.abs().max()).item()
    if not err <= 1e-12:
        print(f"reconstruction error {err} exceeds 1e-12", file=sys.stderr)
        sys.exit(1)

    # timed region: K steps, an event at every kernel boundary on the current
    # stream (one event ends step k and starts step k + 1)
    bounds = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    mids = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bounds[0].record()
    for k in range(args.steps):
        D.modwt_fwd(x, J, filt, out=xmod)
        mids[k].record()
        D.modwt_inv(xmod, n, J, filt, out=xr)
        bounds[k + 1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fwd_ms = sum(bounds[k].elapsed_time(mids[k]) for k in range(args.steps)) / args.steps
    inv_ms = sum(mids[k].elapsed_time(bounds[k + 1]) for k in range(args.steps)) / args.steps
    ms_per_step = wall * 1e3 / args.steps

    # the same step with a 1 GiB write between steps (outside the timed
    # events), so the signal is read from HBM rather than from the 256 MB
    # last-level cache it can stay in when the same step repeats
    junk = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
    cold = []
    for _ in range(5):
        junk.fill_(1.0)
        c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        c0.record()
        D.modwt_fwd(x, J, filt, out=xmod)
        D.modwt_inv(xmod, n, J, filt, out=xr)
        c1.record()
        torch.cuda.synchronize()
        cold.append(c0.elapsed_time(c1))
    del junk
    cold_ms = sorted(cold)[len(cold) // 2]

    fwd_plan, inv_plan = D.modwt_plan(False, n, J, filt), D.modwt_plan(True, n, J, filt)
    fwd_knobs, inv_knobs = plan_knobs(D, False, n, J, filt), plan_knobs(D, True, n, J, filt)
    value = n / (ms_per_step * 1e-3) / 1e6
    bf, bi = modwt_bytes_fwd(n, J), modwt_bytes_inv(n, J)
    fwd_gbs = bf / (fwd_ms * 1e-3) / 1e9
    inv_gbs = bi / (inv_ms * 1e-3) / 1e9
    step_gbs = (bf + bi) / (ms_per_step * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"LA8 MODWT J={J} forward+inverse, one signal of n=2^{args.log2n}, "
                        "FP64 white noise resident in HBM",
            "transform": "MODWT (time ordered)",
            "filter": "LA8",
            "n": n,
            "levels": J,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": plan_kernels(fwd_plan, False),
            "achieved": round(fwd_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(fwd_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles(),
            "algorithmic_bytes_per_launch": bf,
            "avg_launch_ms": round(fwd_ms, 4),
        },
        "inverse_kernel": {
            "kernel": plan_kernels(inv_plan, True),
            "achieved_GBs": round(inv_gbs, 1),
            "frac": round(inv_gbs / HBM_PEAK_GBS, 4),
            "algorithmic_bytes_per_launch": bi,
            "avg_launch_ms": round(inv_ms, 4),
        },
        "step_hbm_frac": round(step_gbs / HBM_PEAK_GBS, 4),
        "cold_ms_per_step": round(cold_ms, 4),
        "cold_msamples_per_s": round(n / (cold_ms * 1e-3) / 1e6, 2),
        "plan": {"fwd": fwd_plan, "inv": inv_plan,
                 "tuned": not (args.no_tune or set_plans), "set": set_plans,
                 "knobs": {"fwd": fwd_knobs, "inv": inv_knobs}},
        "reconstruction_relerr": err,
    }
    if not args.no_batch:
        out["batched"] = batch_section(max(3, args.steps // 5), 2, args.plain_sample)
        out["batched_both_transforms"] = batch_section(max(3, args.steps // 5), 2,
                                                       min(args.plain_sample, 2048), both=True)
    if not args.no_extra:
        out["sweep"] = sweep_section(max(5, args.steps // 5))
        out["small"] = small_section(max(5, args.steps // 5))
        out["layouts"] = levels_section(max(5, args.steps // 10))
        out["smooth"] = smooth_section()
        out["differentiable"] = diff_section()
        out["host_c_abi"] = host_section()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(J=J)
        out["cpu_baseline_openmp"] = cpu_baseline(J=J, omp=True, budget_s=5.0)
        out["cpu_baseline_configs"] = {
            "d4_dwt_n1024": cpu_config1(),
            "smoothing_n2p18": cpu_config3(),
            "batch_8192": cpu_config4(args.cpu_batch_sample),
        }
    out["summary"] = summary(out)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

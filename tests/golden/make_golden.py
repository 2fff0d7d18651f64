"""Write tests/golden/reference_vectors.npz from a build of the reference's
own CPU sources (oracle/_ref/libwvref.so, made by `make -C oracle ref` in a
container that holds the reference tree).

Every entry is data: seeded inputs and the reference's outputs for them,
through the same entry points the reference's R glue calls
(src/rfunctions.cu:3-100).  Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parents[1]))

import oracle_bind as O  # noqa: E402
from wavecuda_amd.signals import noisy_doppler  # noqa: E402

FLEN = {1: 2, 2: 4, 3: 6, 4: 8}
P = C.POINTER(C.c_double)


def p(a):
    return a.ctypes.data_as(P)


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libwvref.so is missing: run `make -C oracle ref` first")
    R = O.ref()
    out = {}
    rng = np.random.Generator(np.random.PCG64(2024))

    def put(key, arr):
        out[key] = np.asarray(arr)

    # ---- DWT forward / inverse, all filters (src/transform.cpp:14-38)
    for f in (1, 2, 3, 4):
        for n in (8, 16, 32, 64, 256, 1024):
            for nl in (0, 1, 2):
                J = R.ref_check_len_levels(n, nl, FLEN[f])
                if J == 0 or n < FLEN[f]:
                    continue
                x = rng.standard_normal(n)
                y = x.copy()
                assert R.ref_transform(p(y), None, n, 1, nl, 0, f, FLEN[f]) == 0
                z = y.copy()
                assert R.ref_transform(p(z), None, n, 0, nl, 0, f, FLEN[f]) == 0
                k = f"dwt_f{f}_n{n}_l{nl}"
                put(k + "_x", x)
                put(k + "_fwd", y)
                put(k + "_inv", z)

    # ---- the D4 n=1024 doppler case, full depth
    xd, _ = noisy_doppler(1024, 5.0, seed=10)
    y = xd.copy()
    R.ref_transform(p(y), None, 1024, 1, 0, 0, 2, 4)
    z = y.copy()
    R.ref_transform(p(z), None, 1024, 0, 0, 0, 2, 4)
    put("doppler1024_d4_x", xd)
    put("doppler1024_d4_fwd", y)
    put("doppler1024_d4_inv", z)

    # ---- time-ordered MODWT, the filters the reference has (Haar, D4)
    for f in (1, 2):
        for n in (8, 16, 64, 256, 1024):
            for nl in (1, 2, 0):
                J = R.ref_check_len_levels(n, nl, FLEN[f])
                if J == 0 or (n == 1024 and nl == 0):
                    continue
                x = rng.standard_normal(n)
                xm = np.zeros(2 * n * J)
                assert R.ref_transform(p(x.copy()), p(xm), n, 1, J, 1, f, FLEN[f]) == 0
                xr = np.zeros(n)
                xm2 = xm.copy()
                assert R.ref_transform(p(xr), p(xm2), n, 0, J, 1, f, FLEN[f]) == 0
                k = f"modwt_f{f}_n{n}_l{nl}"
                put(k + "_x", x)
                put(k + "_fwd", xm)
                put(k + "_inv", xr)
                put(k + "_invxmod", xm2)

    # ---- thresholding (RcpuThreshold), values on and off the threshold
    for ttype in (0, 1):
        for f in ((1, 2, 3, 4) if ttype == 0 else (1, 2)):
            n = 128
            J = R.ref_check_len_levels(n, 0, FLEN[f])
            x = rng.standard_normal(n)
            if ttype == 0:
                c = x.copy()
                R.ref_transform(p(c), None, n, 1, 0, 0, f, FLEN[f])
            else:
                c = np.zeros(2 * n * J)
                R.ref_transform(p(x.copy()), p(c), n, 1, J, 1, f, FLEN[f])
            t = float(np.quantile(np.abs(c), 0.5))
            c[5] = t          # |c| == t exactly
            c[7] = -t
            for hard in (0, 1):
                for (mn, mx) in ((0, J - 1), (1, min(2, J - 1)), (J - 1, J - 1)):
                    cc = c.copy()
                    if ttype == 0:
                        R.ref_threshold(p(cc), None, n, J, 0, f, FLEN[f], t, hard, mn, mx)
                    else:
                        R.ref_threshold(p(np.zeros(n)), p(cc), n, J, 1, f, FLEN[f], t, hard, mn, mx)
                    k = f"thr_t{ttype}_f{f}_h{hard}_m{mn}_{mx}"
                    put(f"thr_t{ttype}_f{f}_in", c)
                    put(f"thr_t{ttype}_f{f}_t", np.array([t]))
                    put(k + "_out", cc)

    # ---- median / MAD (src/utils.cpp:166-203), odd / even / ties
    for i, n in enumerate((1, 2, 3, 10, 11, 1000, 1001)):
        x = np.round(rng.standard_normal(n) * 4) / 4  # ties
        put(f"median_{i}_x", x)
        put(f"median_{i}_med", np.array([R.ref_median(p(x.copy()), n)]))
        put(f"median_{i}_mad", np.array([R.ref_mad(p(x.copy()), n)]))

    # ---- universal threshold (src/thresh.cpp:170-201)
    for ttype in (0, 1):
        for f in ((2, 4) if ttype == 0 else (1, 2)):
            n = 256
            J = R.ref_check_len_levels(n, 0, FLEN[f])
            x, _ = noisy_doppler(n, 3.0, seed=5)
            if ttype == 0:
                c = x.copy()
                R.ref_transform(p(c), None, n, 1, 0, 0, f, FLEN[f])
                xm = np.zeros(1)
            else:
                xm = np.zeros(2 * n * J)
                R.ref_transform(p(x.copy()), p(xm), n, 1, J, 1, f, FLEN[f])
                c = np.zeros(n)
            for (mn, mx) in ((0, 0), (0, J - 1), (1, 3), (2, J - 1)):
                u = R.ref_univ_thresh(p(c), p(xm), n, J, ttype, f, FLEN[f], mn, mx)
                put(f"univ_t{ttype}_f{f}_coef", c if ttype == 0 else xm)
                put(f"univ_t{ttype}_f{f}_m{mn}_{mx}_u", np.array([u]))

    # ---- interpolation error (src/thresh.cpp:344-384)
    for n in (4, 8, 1024):
        xn = rng.standard_normal(n)
        ye = rng.standard_normal(n // 2)
        yo = rng.standard_normal(n // 2)
        put(f"imse_n{n}_xn", xn)
        put(f"imse_n{n}_ye", ye)
        put(f"imse_n{n}_yo", yo)
        put(f"imse_n{n}_m", np.array([R.ref_interp_mse(p(xn), p(ye), p(yo), n)]))

    # ---- smoothing (RcpuSmooth): manual / universal / CV
    cases = []
    for f in (1, 2, 3, 4):
        cases.append((0, f, 1, 0, 1, 0, 6, 1024, 5.0))  # DWT univ soft
        cases.append((0, f, 0, 0, 0, 1, 5, 1024, 5.0))  # DWT manual hard
    for f in (1, 2):
        for n, snr in ((1024, 5.0), (4096, 2.0)):
            J = 8 if f == 1 else 7
            cases.append((1, f, 2, J, 1, 0, J - 1, n, snr))  # MODWT CV soft
            cases.append((1, f, 2, J, 0, 1, J - 1, n, snr))  # MODWT CV hard
            cases.append((1, f, 1, J, 1, 0, J - 1, n, snr))  # MODWT univ soft
        cases.append((1, f, 0, 5, 1, 0, 4, 1024, 5.0))       # MODWT manual
    for n, snr in ((1024, 5.0), (4096, 2.0)):
        cases.append((0, 1, 2, 0, 1, 0, int(np.log2(n)) - 1, n, snr))  # DWT Haar CV full depth
        cases.append((0, 1, 2, 0, 0, 2, int(np.log2(n)) - 1, n, snr))
    for i, (tt, f, th, nl, hard, mn, mx, n, snr) in enumerate(cases):
        x, clean = noisy_doppler(n, snr, seed=100 + i)
        y = x.copy()
        t = np.array([0.25 if th == 0 else 0.0])
        rc = R.ref_smooth(p(y), n, nl, tt, f, FLEN[f], th, p(t), hard, mn, mx, 0.01)
        assert rc == 0
        k = f"smooth_{i}"
        # the input is noisy_doppler(n, snr, seed=100 + i) (wavecuda_amd.signals)
        put(k + "_args", np.array([tt, f, th, nl, hard, mn, mx, n], dtype=np.int64))
        put(k + "_snr", np.array([snr]))
        put(k + "_y", y)
        put(k + "_t", t)

    # ---- packet-ordered MODWT (src/transform.cpp:39-55: Haar, D4) and the
    # PO <-> TO conversion (src/wavutils.cpp:532-600), from a stream of
    # their own so the vectors above do not change
    rpo = np.random.Generator(np.random.PCG64(4242))
    for f in (1, 2):
        for n in (16, 64, 256, 1024):
            for nl in (1, 2, 0):
                J = R.ref_check_len_levels(n, nl, FLEN[f])
                if J == 0 or (n == 1024 and nl == 0):
                    continue
                x = rpo.standard_normal(n)
                xm = np.zeros(2 * n * J)
                assert R.ref_transform(p(x.copy()), p(xm), n, 1, J, 2, f, FLEN[f]) == 0
                xr = np.zeros(n)
                assert R.ref_transform(p(xr), p(xm.copy()), n, 0, J, 2, f, FLEN[f]) == 0
                k = f"modwtpo_f{f}_n{n}_l{J}"
                put(k + "_x", x)
                put(k + "_fwd", xm)
                put(k + "_inv", xr)
    for n, J in ((8, 3), (64, 6), (256, 4), (1024, 10), (4096, 5)):
        v = rpo.standard_normal(2 * n * J)
        po = np.zeros_like(v)
        assert R.ref_convert_modwt(p(v), p(po), n, J, 1) == 0
        to = np.zeros_like(v)
        assert R.ref_convert_modwt(p(v), p(to), n, J, 2) == 0
        k = f"convert_n{n}_l{J}"
        put(k + "_in", v)
        put(k + "_to2po", po)
        put(k + "_po2to", to)

    path = HERE / "reference_vectors.npz"
    np.savez_compressed(path, **out)
    print(f"wrote {len(out)} arrays to {path} ({path.stat().st_size} bytes)")


if __name__ == "__main__":
    main()

"""Tail extrapolation of the IFA port transient by the matrix-pencil method.

The learned early stop (earlystop.py) predicts the port tail from a 1500-step prefix.
This module measures what a physics-based predictor reaches instead, and at which
prefix length: after the source pulse the port voltage and current ring down as a sum
of damped exponentials sharing their poles, which the matrix-pencil method (Y. Hua and
T. K. Sarkar, "Matrix pencil method for estimating parameters of exponentially
damped/undamped sinusoids in noise", IEEE Trans. ASSP 38(5), 1990) fits to a window of
samples.  The fitted sum continues the traces to the full length, and S11 follows from
the port DFTs as the reference forms them (js_fdtd/src/dft.js:17-38,
js_fdtd/src/sparameters.js:65-101).

    python -m fdtdhip.extrapolate --first 4000 --count 64 --out study.json

runs each variant for the full length in FP64 (the reference result) and in FP32 (the
early-stopped run reads its first N steps) on the GPU, then reports, for several N, the
worst S11 error over 1000 frequencies of the matrix-pencil continuation and of plain
truncation, and the largest speed-up an early stop at N could give.
"""

from __future__ import annotations

import argparse
import json
import time

import numpy as np

STEPS = 7000
FREQS = {"maxFreq": 6e9, "step": 6e6}     # the early stop's frequency grid


def matrix_pencil(ys, tol=1e-6, L=None):
    """Poles z_k shared by the sequences ys (each of M samples): the nonzero eigenvalues
    of the pencil (Y2, Y1) of their stacked Hankel matrices, reduced to the K singular
    vectors whose singular values exceed tol times the largest."""
    M = len(ys[0])
    L = M // 2 if L is None else L
    H = np.vstack([np.lib.stride_tricks.sliding_window_view(np.asarray(y, dtype=float), L + 1) for y in ys])
    _, s, Vh = np.linalg.svd(H, full_matrices=False)
    K = max(1, int(np.sum(s > tol * s[0])))
    Vk = Vh[:K].conj().T
    return np.linalg.eigvals(np.linalg.pinv(Vk[:-1]) @ Vk[1:])


def fit_amplitudes(y, z):
    """Least-squares amplitudes a with y[n] = sum_k a_k z_k^n."""
    A = np.asarray(z)[None, :] ** np.arange(len(y))[:, None]
    a, *_ = np.linalg.lstsq(A, np.asarray(y, dtype=complex), rcond=None)
    return a


def extend(v, i, start, n, steps=STEPS, dec=8, tol=1e-6):
    """The traces v, i known for steps [0, n), continued to `steps` by damped exponentials
    fitted (shared poles, every dec-th sample) to the window [start, n).  Growing poles are
    moved onto the unit circle: the continued run may not blow up."""
    idx = np.arange(start, n, dec)
    yv, yi = np.asarray(v, float)[idx], np.asarray(i, float)[idx]
    sv = max(np.abs(yv).max(), 1e-300)
    si = max(np.abs(yi).max(), 1e-300)
    z = matrix_pencil([yv / sv, yi / si], tol)
    z = np.where(np.abs(z) > 1.0, z / np.abs(z), z)
    av, ai = fit_amplitudes(yv, z), fit_amplitudes(yi, z)
    zt = np.exp(np.outer((np.arange(n, steps) - start) / dec, np.log(z.astype(complex))))
    return (np.concatenate([np.asarray(v, float)[:n], (zt @ av).real]),
            np.concatenate([np.asarray(i, float)[:n], (zt @ ai).real]))


def port_dft(v, i, dt, freqs):
    """V(f), I(f) of whole traces: sample n at (n + 1) dt, the current half a step earlier."""
    w = 2 * np.pi * np.asarray(freqs)
    t = (np.arange(len(v)) + 1) * dt
    return dt * (np.exp(-1j * np.outer(w, t)) @ v), dt * (np.exp(-1j * np.outer(w, t - dt / 2)) @ i)


def s11_db(V, I, z=50.0):
    return 20 * np.log10(np.abs((0.5 * (V - z * I)) / (0.5 * (V + z * I))))


def study(first, count, prefixes=(1500, 2000, 2500, 3000, 3500, 4000), seed=4, batch=32,
          workdir="/tmp/es-extrapolate", start=1100, dec=8, tol=1e-6):
    from . import earlystop as es
    from .sweep import LazyBuilds, Sweep, frequency_grid
    paths, _ = es.variant_scenarios(first, count, seed, workdir)
    runs, secs = {}, {}
    for prec in ("fp64", "fp32"):
        builds = LazyBuilds(paths, f"{workdir}/{prec}", fp32=prec == "fp32", jobs=16, ahead=2 * batch)
        res, s, _ = Sweep(prec, "matlab", keep_traces=True, farfield=False, s11=False, batch=batch,
                          progress_every=16).run(builds)
        builds.close()
        runs[prec], secs[prec] = res, s
        print(f"[extrapolate] {prec}: {count} variants in {s:.1f} s", flush=True)
    freqs = frequency_grid(FREQS)
    w = 2 * np.pi * freqs
    mp = np.zeros((len(prefixes), count))
    tr = np.zeros((len(prefixes), count))
    f32 = np.zeros(count)
    fit_s = 0.0
    for q, (r64, r32) in enumerate(zip(runs["fp64"], runs["fp32"])):
        dt = r64["dt"]
        # port_dft's sums with the phase matrix shared by every trace of the variant.
        E = np.exp(-1j * np.outer(w, (np.arange(STEPS) + 1) * dt))
        half = np.exp(1j * w * dt / 2)

        def s11(v, i):
            return s11_db(dt * (E @ v), dt * half * (E @ i))

        ref = s11(r64["v"], r64["i"])
        f32[q] = np.abs(s11(r32["v"], r32["i"]) - ref).max()
        for p, n in enumerate(prefixes):
            t0 = time.perf_counter()
            ve, ie = extend(r32["v"], r32["i"], start, n, dec=dec, tol=tol)
            fit_s += time.perf_counter() - t0
            mp[p, q] = np.abs(s11(ve, ie) - ref).max()
            z = np.zeros(STEPS - n)
            tr[p, q] = np.abs(s11(np.concatenate([r32["v"][:n], z]), np.concatenate([r32["i"][:n], z])) - ref).max()
        if (q + 1) % 8 == 0:
            print(f"[extrapolate] {q + 1}/{count} variants scored", flush=True)

    def stats(e):
        return {"median_db": round(float(np.median(e)), 3), "p90_db": round(float(np.quantile(e, 0.9)), 3),
                "within_0p1db": round(float(np.mean(e < 0.1)), 3)}

    return {"variants": count, "first": first, "window_start": start, "decimation": dec, "tol": tol,
            "fp64_s_per_variant": round(secs["fp64"] / count, 4), "fp32_s_per_variant": round(secs["fp32"] / count, 4),
            "fp32_full_max_err_db": float(f32.max()),
            "fit_host_s_per_variant": round(fit_s / (count * len(prefixes)), 4),
            "prefixes": [{"n": n, "speedup_bound": round(secs["fp64"] / (secs["fp32"] * n / STEPS), 2),
                          "matrix_pencil": stats(mp[p]), "truncated": stats(tr[p])} for p, n in enumerate(prefixes)]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--first", type=int, default=4000)
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    text = json.dumps(study(a.first, a.count, seed=a.seed))
    if a.out:
        with open(a.out, "w") as f:
            f.write(text)
    print(text, flush=True)


if __name__ == "__main__":
    main()

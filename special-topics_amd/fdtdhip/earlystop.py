"""Learned early stop of the IFA port transient.

A run is stopped after N = 1500 of its 7000 steps; a small network predicts the
remaining port-voltage and port-current tail from the first N steps, and S11 is taken
from the measured prefix plus the predicted tail.  The network also estimates its own
S11 error; when that estimate exceeds a threshold calibrated on a validation split (so
that, with fallbacks counted as exact, 95% of validation variants end within 0.1 dB of
the full run) the full-length FP64 run is used instead.

Representation (all per channel, V and I, each scaled by its prefix max |x|):
  input  x: the 1500-step prefix mean-pooled by 4  -> 375 samples
  target  : the tail, steps 1500..6995, mean-pooled by 8 -> 687 samples,
            plus log10 of the residual-energy ratio.
Pooled samples sit at window centres; for signals band-limited to 6 GHz the pooled
tail reproduces the full-rate DFT sum to about 0.02 dB (midpoint rule, phase step
2*pi*6 GHz*8*dt = 0.24 rad).

    python -m fdtdhip.earlystop gen-train --first 0 --count 2000 --out train0.npz
    python -m fdtdhip.earlystop gen-heldout --first 4000 --count 500 --out heldout.npz
    python -m fdtdhip.earlystop evaluate --weights W.npz --heldout heldout.npz
    python -m fdtdhip.earlystop tts-batched --weights W.npz --first 4600 --count 64

Data generation and the batched time-to-solution step `--batch` variants together in
one batched launch per phase (fdtdhip.sim.HipMulti).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from .sweep import LazyBuilds, Sweep, build_all, frequency_grid

N_PREFIX = 1500
POOL_IN = 4
POOL_OUT = 8
STEPS = 7000
TAIL = (STEPS - N_PREFIX) // POOL_OUT * POOL_OUT     # 5496
N_IN = N_PREFIX // POOL_IN                           # 375
N_OUT = TAIL // POOL_OUT                             # 687


def pool(x: np.ndarray, k: int) -> np.ndarray:
    n = x.shape[-1] // k * k
    return x[..., :n].reshape(*x.shape[:-1], -1, k).mean(axis=-1)


def features(v: np.ndarray, i: np.ndarray):
    """Network input [2, N_IN] and the channel scales from the first N_PREFIX steps."""
    scale = np.array([np.max(np.abs(v[:N_PREFIX])), np.max(np.abs(i[:N_PREFIX]))])
    scale = np.where(scale > 0, scale, 1.0)
    x = np.stack([pool(v[:N_PREFIX] / scale[0], POOL_IN), pool(i[:N_PREFIX] / scale[1], POOL_IN)])
    return x.astype(np.float32), scale


def targets(v: np.ndarray, i: np.ndarray, scale: np.ndarray):
    """Pooled normalised tail [2, N_OUT] and log10 residual-energy ratio."""
    vn, inn = v / scale[0], i / scale[1]
    tail = np.stack([pool(vn[N_PREFIX:N_PREFIX + TAIL], POOL_OUT), pool(inn[N_PREFIX:N_PREFIX + TAIL], POOL_OUT)])
    e = vn ** 2 + inn ** 2
    ratio = e[N_PREFIX:].sum() / max(e.sum(), 1e-300)
    return tail.astype(np.float32), np.float32(np.log10(max(ratio, 1e-30)))


def tail_dft(tail: np.ndarray, scale: np.ndarray, dt: float, freqs: np.ndarray):
    """DFT contribution of a pooled tail (un-normalised) with the conventions of
    js_fdtd/src/dft.js:17-38: shift 0 for V, -dt/2 for I; returns (V(f), I(f)).  Each
    pooled sample stands for its window's sum at the window centre (midpoint rule)."""
    ph = _tail_phases(dt, freqs)
    return ph[0] @ (tail[0] * scale[0]), ph[1] @ (tail[1] * scale[1])


_PHASES = {}


def _tail_phases(dt: float, freqs: np.ndarray):
    """dt * POOL_OUT * exp(-j w t) at the pooled-window centres, for V and for I
    (cached: it depends on dt and the frequency grid only)."""
    key = (float(dt), freqs.size, float(freqs[0]), float(freqs[-1]))
    if key not in _PHASES:
        centres = N_PREFIX + POOL_OUT * np.arange(N_OUT) + (POOL_OUT - 1) / 2
        w = 2 * np.pi * freqs
        _PHASES[key] = [dt * POOL_OUT * np.exp(-1j * np.outer(w, (centres + 1) * dt + shift))
                        for shift in (0.0, -dt / 2)]
    return _PHASES[key]


def prefix_dft(v: np.ndarray, i: np.ndarray, dt: float, freqs: np.ndarray):
    w = 2 * np.pi * freqs
    n = np.arange(N_PREFIX)
    V = dt * (np.exp(-1j * np.outer(w, (n + 1) * dt)) @ v[:N_PREFIX])
    I = dt * (np.exp(-1j * np.outer(w, (n + 1) * dt - dt / 2)) @ i[:N_PREFIX])
    return V, I


def s11_from(V, I, z=50.0):
    return (0.5 * (V - z * I)) / (0.5 * (V + z * I))


# ------------------------------------------------------------------------ data --

def variant_scenarios(first, count, seed, workdir, steps=STEPS):
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, os.path.join(here, "scenarios"))
    import gen_ifa_variants
    os.makedirs(workdir, exist_ok=True)
    paths, params = [], []
    for idx in range(first, first + count):
        sc = gen_ifa_variants.scenario(seed, idx, steps)
        path = os.path.join(workdir, sc["name"] + (f"-{steps}" if steps != STEPS else "") + ".json")
        with open(path, "w") as f:
            json.dump(sc, f)
        paths.append(path)
        params.append([sc["parameters"][k] for k in ("arm_length", "arm_height", "feed_short", "ground_w", "ground_l")])
    return paths, np.array(params)


def gen_train(first, count, out, seed=4, workdir="/tmp/es-data", batch=64, jobs=16):
    """FP32 full runs (matlab boundary) of variants [first, first+count), `batch` at a
    time in one batched launch: network input and targets, plus what the S11-error
    target needs -- the full-rate prefix traces and the full run's S11."""
    paths, params = variant_scenarios(first, count, seed, workdir)
    builds = LazyBuilds(paths, workdir, fp32=True, jobs=jobs, ahead=2 * batch)
    res, secs, cells = Sweep("fp32", "matlab", keep_traces=True, farfield=False, s11=True,
                             progress_every=1, batch=batch).run(builds)
    builds.close()
    xs, scales, tails, logr, dims, dts = [], [], [], [], [], []
    for r in res:
        x, sc = features(r["v"], r["i"])
        t, lr = targets(r["v"], r["i"], sc)
        xs.append(x); scales.append(sc); tails.append(t); logr.append(lr); dims.append(r["dims"]); dts.append(r["dt"])
    np.savez_compressed(out, index=np.arange(first, first + count), params=params, x=np.array(xs),
                        scale=np.array(scales), tail=np.array(tails), log_ratio=np.array(logr),
                        dims=np.array(dims), dt=np.array(dts),
                        v_prefix=np.array([r["v"][:N_PREFIX] for r in res], dtype=np.float32),
                        i_prefix=np.array([r["i"][:N_PREFIX] for r in res], dtype=np.float32),
                        s11=np.array([r["s11"] for r in res], dtype=np.complex64))
    return {"variants": count, "run_s": round(secs, 1), "mcell_updates_per_s": round(cells / secs / 1e6, 1)}


def gen_heldout(first, count, out, seed=4, workdir="/tmp/es-heldout", batch=64, jobs=16):
    """Per held-out variant: the full-length FP64 S11 (the reference result) and the FP32 prefix
    traces at full rate (what an early-stopped run sees)."""
    paths, params = variant_scenarios(first, count, seed, workdir)
    b64 = LazyBuilds(paths, os.path.join(workdir, "f64"), fp32=False, jobs=jobs, ahead=2 * batch)
    res64, secs64, cells64 = Sweep("fp64", "matlab", keep_traces=False, farfield=False,
                                   progress_every=1, batch=batch).run(b64)
    b64.close()
    b32 = LazyBuilds(paths, os.path.join(workdir, "f32"), fp32=True, jobs=jobs, ahead=2 * batch)
    res32, secs32, _ = Sweep("fp32", "matlab", keep_traces=True, farfield=False, s11=False, steps=N_PREFIX,
                             progress_every=1, batch=batch).run(b32)
    b32.close()
    np.savez_compressed(out, index=np.arange(first, first + count), params=params,
                        s11_fp64=np.array([r["s11"] for r in res64]),
                        v_prefix=np.array([r["v"] for r in res32]), i_prefix=np.array([r["i"] for r in res32]),
                        dims=np.array([r["dims"] for r in res32]), dt=np.array([r["dt"] for r in res32]))
    return {"variants": count, "fp64_run_s": round(secs64, 1),
            "fp64_mcell_updates_per_s": round(cells64 / secs64 / 1e6, 1), "fp32_prefix_run_s": round(secs32, 1)}


# ------------------------------------------------------------------ inference --

class Predictor:
    """The HIP inference kernels (fdtd_hip_es_*) with weights from earlystop_net.train,
    and the fallback threshold calibrated with them."""

    def __init__(self, weights_path):
        import ctypes as C
        from . import native
        d = np.load(weights_path)
        self.w = np.ascontiguousarray(d["weights_bf16"], dtype=np.uint16)
        self.b = np.ascontiguousarray(d["biases"], dtype=np.float32)
        self.threshold = float(json.loads(str(d["meta"]))["calibration"]["threshold"])
        self.C, self.native, self.L = C, native, native.lib()
        self.h = C.c_void_p()
        native.check(self.L.fdtd_hip_es_create(self.w.ctypes.data_as(C.POINTER(C.c_uint16)), self.w.size,
                                               self.b.ctypes.data_as(C.POINTER(C.c_float)), self.b.size,
                                               C.byref(self.h)), "es_create")

    def __call__(self, x: np.ndarray) -> np.ndarray:
        C = self.C
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 2, N_IN)
        y = np.zeros((len(x), 2 * N_OUT + 2), dtype=np.float32)
        self.native.check(self.L.fdtd_hip_es_predict(self.h, len(x), x.ctypes.data_as(C.POINTER(C.c_float)),
                                                     y.ctypes.data_as(C.POINTER(C.c_float))), "es_predict")
        return y

    def close(self):
        if self.h:
            self.L.fdtd_hip_es_destroy(self.h)
            self.h = None


def decode(y: np.ndarray, x: np.ndarray):
    """Network output -> (pooled tail [2, N_OUT] in prefix-scale units, log10 residual
    energy ratio, log10 of the estimated S11 error in dB)."""
    from .earlystop_net import end_amplitude
    a = end_amplitude(x)
    return y[: 2 * N_OUT].reshape(2, N_OUT) * a[:, None], float(y[2 * N_OUT]), float(y[2 * N_OUT + 1])


def evaluate(weights_path, heldout_path):
    """Early-stop S11 of every held-out variant against its full-length FP64 S11, with
    the calibrated fallback, and the accuracy / fallback-rate curve over thresholds."""
    d = np.load(heldout_path)
    dt = d["dt"]
    freqs = frequency_grid({"maxFreq": 6e9, "step": 6e6})
    feats = [features(v, i) for v, i in zip(d["v_prefix"], d["i_prefix"])]
    X = np.stack([f[0] for f in feats])
    net = Predictor(weights_path)
    Y = net(X)
    th = net.threshold
    net.close()
    errs, est, res_err, band_err = [], [], [], []
    for k, (x, scale) in enumerate(feats):
        tail, _, le = decode(Y[k], x)
        Vp, Ip = prefix_dft(d["v_prefix"][k], d["i_prefix"][k], dt[k], freqs)
        Vt, It = tail_dft(tail, scale, dt[k], freqs)
        s_es = s11_from(Vp + Vt, Ip + It)
        db_es, db_64 = 20 * np.log10(np.abs(s_es)), 20 * np.log10(np.abs(d["s11_fp64"][k]))
        e = np.abs(db_es - db_64)
        errs.append(float(np.max(e)))
        est.append(le)
        # What the early result gets right short of the 0.1 dB bound: the resonance (S11
        # minimum) and S11 wherever the antenna reflects at least -10 dB.
        res_err.append(abs(float(freqs[np.argmin(db_es)] - freqs[np.argmin(db_64)])))
        band_err.append(float(np.max(e[db_64 >= -10.0])) if np.any(db_64 >= -10.0) else 0.0)
    errs, est = np.array(errs), np.array(est)
    fall = est > th
    with_fb = np.where(fall, 0.0, errs)
    curve = []
    for q in np.linspace(0, 1, 11):
        t = np.quantile(est, q)
        keep = est <= t
        curve.append({"threshold": round(float(t), 4), "fallback_rate": round(float(np.mean(~keep)), 4),
                      "within_0p1db": round(float(np.mean((errs <= 0.1) | ~keep)), 4)})
    return {
        "variants": int(len(errs)),
        "threshold": th,
        "max_err_db_median": round(float(np.median(errs)), 4),
        "max_err_db_p90": round(float(np.percentile(errs, 90)), 4),
        "within_0p1db_no_fallback": round(float(np.mean(errs <= 0.1)), 4),
        "fallback_rate": round(float(np.mean(fall)), 4),
        "within_0p1db_with_fallback": round(float(np.mean(with_fb <= 0.1)), 4),
        "max_err_db_with_fallback_p90": round(float(np.percentile(with_fb, 90)), 4),
        "estimate_vs_error_corr": round(float(np.corrcoef(est, np.log10(errs + 1e-3))[0, 1]), 3),
        "resonance_err_mhz_median": round(float(np.median(res_err)) / 1e6, 1),
        "resonance_err_mhz_p90": round(float(np.percentile(res_err, 90)) / 1e6, 1),
        "max_err_db_where_s11_above_m10db_median": round(float(np.median(band_err)), 4),
        "curve": curve,
    }


def time_to_solution(weights_path, first, count, seed=4, workdir="/tmp/es-tts"):
    """Per variant, sequentially: the full-length FP64 run to S11 against the FP32 run to
    step 1500 + network + S11 (plus the FP64 run when the fallback triggers).  Builds,
    uploads and autotuning happen before each clock starts."""
    from .problem import read_problem
    from .sim import HipSim
    from .sweep import build_one, s11
    paths, _ = variant_scenarios(first, count, seed, workdir)
    freqs = frequency_grid({"maxFreq": 6e9, "step": 6e6})
    net = Predictor(weights_path)
    rows = []
    for path in paths:
        p64 = read_problem(build_one(path, workdir, fp32=False))
        p32 = read_problem(build_one(path, workdir, fp32=True))
        dt = p64.grid["dt"]
        _tail_phases(dt, freqs)                       # built once per dt, outside the clock
        sim = HipSim(p64, precision="fp64", boundary="matlab", port_dft_freqs=freqs, farfield=False)
        sim.autotune(3)
        sim.synchronize()
        t0 = time.perf_counter()
        sim.run_no_readback(0, STEPS)
        S64 = s11(sim, p64)
        t64 = time.perf_counter() - t0
        sim.close()
        sim = HipSim(p32, precision="fp32", boundary="matlab", port_dft_freqs=freqs, farfield=False)
        sim.autotune(3)
        sim.synchronize()
        t0 = time.perf_counter()
        v, i = sim.run(0, N_PREFIX)
        X = sim.port_dft()
        x, scale = features(v[0], i[0])
        tail, _, le = decode(net(x)[0], x)
        Vt, It = tail_dft(tail, scale, dt, freqs)
        S = s11_from(X[0] + Vt, X[sim.n_sv] + It)
        tes = time.perf_counter() - t0
        sim.close()
        fb = le > net.threshold
        err = float(np.max(np.abs(20 * np.log10(np.abs(S)) - 20 * np.log10(np.abs(S64)))))
        rows.append({"t_fp64_s": t64, "t_early_s": tes + (t64 if fb else 0.0), "fallback": bool(fb),
                     "max_err_db": 0.0 if fb else err})
        for f in (p64.path, p32.path):
            os.remove(f)
    net.close()
    t64 = sum(r["t_fp64_s"] for r in rows)
    tes = sum(r["t_early_s"] for r in rows)
    return {"variants": count, "fp64_s_mean": round(t64 / count, 4), "early_stop_s_mean": round(tes / count, 4),
            "speedup": round(t64 / tes, 2), "fallbacks": int(sum(r["fallback"] for r in rows)),
            "max_err_db": round(max(r["max_err_db"] for r in rows), 4)}


def _fp64_batch(p64, freqs, steps=STEPS):
    """Run problems together in one batched FP64 launch; returns (S11 list, seconds
    from launch to S11, tuning).  Uploads and the batch autotuner come before the clock."""
    from .sim import HipMulti, HipSim
    from .sweep import s11
    sims = [HipSim(p, precision="fp64", boundary="matlab", port_dft_freqs=freqs, farfield=False) for p in p64]
    multi = HipMulti(sims)
    tune = multi.autotune(2)
    multi.synchronize()
    t0 = time.perf_counter()
    multi.run_async(0, steps)
    multi.synchronize()
    S = [s11(sim, p) for sim, p in zip(sims, p64)]
    secs = time.perf_counter() - t0
    multi.close()
    for sim in sims:
        sim.close()
    return S, secs, tune


def time_to_solution_batched(weights_path, first, count, seed=4, workdir="/tmp/es-tts-batch", jobs=16):
    """The same comparison with every variant in one batched launch on each side: the
    FP64 reference runs all `count` variants for 7000 steps; the early-stopped path runs
    them in FP32 for 1500 steps, evaluates the network on the whole batch and forms
    S11; the variants whose S11-error estimate exceeds the calibrated threshold then run
    in full FP64 as a batch of their own, and that batch's measured time is added.
    Builds, uploads and the batch autotuners run outside the clocks; times are per
    variant."""
    from .problem import read_problem
    from .sim import HipMulti, HipSim
    paths, _ = variant_scenarios(first, count, seed, workdir)
    freqs = frequency_grid({"maxFreq": 6e9, "step": 6e6})
    p64 = [read_problem(f) for f in build_all(paths, os.path.join(workdir, "f64"), fp32=False, jobs=jobs)]
    p32 = [read_problem(f) for f in build_all(paths, os.path.join(workdir, "f32"), fp32=True, jobs=jobs)]
    for p in p64:
        _tail_phases(p.grid["dt"], freqs)

    S64, t64, tune64 = _fp64_batch(p64, freqs)

    net = Predictor(weights_path)
    sims = [HipSim(p, precision="fp32", boundary="matlab", port_dft_freqs=freqs, farfield=False) for p in p32]
    multi = HipMulti(sims)
    tune32 = multi.autotune(2)
    multi.synchronize()
    t0 = time.perf_counter()
    multi.run_async(0, N_PREFIX)
    feats = [features(*[a[0] for a in sim.traces(0, N_PREFIX)]) for sim in sims]
    Y = net(np.stack([f[0] for f in feats]))
    S, est = [], []
    for k, sim in enumerate(sims):
        X = sim.port_dft()
        tail, _, le = decode(Y[k], feats[k][0])
        Vt, It = tail_dft(tail, feats[k][1], p32[k].grid["dt"], freqs)
        S.append(s11_from(X[0] + Vt, X[sim.n_sv] + It))
        est.append(le)
    t32 = time.perf_counter() - t0
    multi.close()
    for sim in sims:
        sim.close()
    fb = np.array(est) > net.threshold
    net.close()
    t_fb = 0.0
    if fb.any():
        S_fb, t_fb, _ = _fp64_batch([p for p, f in zip(p64, fb) if f], freqs)
        for k, s in zip(np.nonzero(fb)[0], S_fb):
            S[k] = s
    for sub in ("f64", "f32"):
        d = os.path.join(workdir, sub)
        for f in os.listdir(d):
            if f.endswith(".fdtdp"):
                os.remove(os.path.join(d, f))

    err = np.array([np.max(np.abs(20 * np.log10(np.abs(a)) - 20 * np.log10(np.abs(b)))) for a, b in zip(S, S64)])
    t_early = t32 + t_fb
    return {"variants": count, "fp64_tuning": list(tune64), "fp32_tuning": list(tune32),
            "fp64_s_per_variant": round(t64 / count, 5), "early_stop_s_per_variant": round(t_early / count, 5),
            "prefix_and_network_s": round(t32, 4), "fallback_batch_s": round(t_fb, 4),
            "speedup": round(t64 / t_early, 2), "fallbacks": int(fb.sum()),
            "speedup_without_fallbacks": round(t64 / t32, 2),
            "max_err_db_max": round(float(err.max()), 4), "max_err_db_median": round(float(np.median(err)), 4),
            "within_0p1db": round(float(np.mean(err <= 0.1)), 4),
            "per_variant": {"fallback": [bool(f) for f in fb], "estimate": [round(float(e), 4) for e in est],
                            "max_err_db": [float(e) for e in err]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["gen-train", "gen-heldout", "evaluate", "tts", "tts-batched"])
    ap.add_argument("--weights", default=None)
    ap.add_argument("--heldout", default=None)
    ap.add_argument("--first", type=int, default=0)
    ap.add_argument("--count", type=int, default=100)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    if a.what == "evaluate":
        out = evaluate(a.weights, a.heldout)
    elif a.what == "tts":
        out = time_to_solution(a.weights, a.first, a.count, seed=a.seed)
    elif a.what == "tts-batched":
        out = time_to_solution_batched(a.weights, a.first, a.count, seed=a.seed)
    else:
        fn = gen_train if a.what == "gen-train" else gen_heldout
        out = fn(a.first, a.count, a.out, seed=a.seed, batch=a.batch)
    text = json.dumps(out)
    if a.out and a.what in ("evaluate", "tts", "tts-batched"):
        with open(a.out, "w") as f:
            f.write(text)
    print(text, flush=True)


if __name__ == "__main__":
    main()

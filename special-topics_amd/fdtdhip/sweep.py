"""Batched sweep of many problems on one GPU.

Each problem gets its own context (device state and HIP stream).  Two ways to run
them:
  * batched (``batch=N``): N contexts at a time are stepped together by one launch per
    phase (fdtd_hip_multi_*, tuned once by its autotuner), while the host uploads the
    next N;
  * in flight (``batch=0``): ``inflight`` contexts run at once, each enqueued on its
    own stream with fdtd_hip_run_batch_async.
Either way the host prepares the next problems (builder process, upload, coefficient
deduplication) while the GPU runs.  Per problem the sweep keeps the port traces, S11 from the on-GPU running port
DFT (sparameters.js:23-101 from dft.js:17-38 sums), the resonance (S11 minimum) and the
directivity cuts of js_fdtd/src/index.js:77-92 at the NTFF frequency nearest it
(farfield.js:257-607, evaluated on the GPU).

    python -m fdtdhip.sweep SCENARIO.json ... [--batch 64 | --batch 0 --inflight 4] [--precision fp32]
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

import numpy as np

from .problem import build_problem, read_problem
from . import native
from .sim import HipMulti, HipSim


def standard_cuts():
    ang = np.arange(360) * np.pi / 180
    th_xz = np.where(ang > np.pi, 2 * np.pi - ang, ang)
    ph_xz = np.where(ang > np.pi, np.pi, 0.0)
    return {"xy": (np.full(360, np.pi / 2), ang), "xz": (th_xz, ph_xz)}


def frequency_grid(dft: dict | None) -> np.ndarray:
    dft = dft or {}
    step = dft.get("step", 20e6)
    return (np.arange(int(np.floor(dft.get("maxFreq", 10e9) / step))) + 1) * step


def s11(sim: HipSim, problem) -> np.ndarray:
    X = sim.port_dft()
    port = next(p for p in problem.header["ports"] if p.get("is_source_port"))
    z = port["impedance"] if isinstance(port["impedance"], (int, float)) else port["impedance"]["re"]
    V = X[port["sampled_voltage_index"] - 1]
    I = X[sim.n_sv + port["sampled_current_index"] - 1]
    return (0.5 * (V - z * I)) / (0.5 * (V + z * I))


def build_one(path, workdir, fp32, steps=None):
    """Problem file of one scenario (built once, by the Node builder)."""
    os.makedirs(workdir, exist_ok=True)
    out = os.path.join(workdir, os.path.basename(path).replace(".json", "-f32.fdtdp" if fp32 else "-f64.fdtdp"))
    if not os.path.exists(out):
        build_problem(path, out + ".tmp", steps=steps, fp32=fp32)
        os.replace(out + ".tmp", out)
    return out


def build_all(scenarios, workdir, fp32, jobs=8, steps=None):
    """Problem files for every scenario, built by parallel builder processes."""
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        return list(ex.map(lambda p: build_one(p, workdir, fp32, steps), scenarios))


class LazyBuilds:
    """Scenario files built on demand by `jobs` builder processes, at most `ahead`
    ahead of the consumer; each problem file is deleted once its context holds it."""

    def __init__(self, scenarios, workdir, fp32, jobs=8, ahead=32):
        self.scenarios, self.workdir, self.fp32, self.ahead = list(scenarios), workdir, fp32, ahead
        self.ex = cf.ThreadPoolExecutor(max_workers=jobs)
        self.futures = {}

    def __len__(self):
        return len(self.scenarios)

    def get(self, k):
        for q in range(k, min(k + self.ahead, len(self.scenarios))):
            if q not in self.futures:
                self.futures[q] = self.ex.submit(build_one, self.scenarios[q], self.workdir, self.fp32)
        return self.futures.pop(k).result()

    def close(self):
        self.ex.shutdown(wait=True)


class Sweep:
    """Runs problems `batch` at a time in one batched launch (batch > 0), or with up to
    `inflight` single-problem contexts in flight (batch = 0); results in submission
    order."""

    def __init__(self, precision="fp32", boundary="matlab", inflight=4, tuning=None, keep_traces=True,
                 farfield=True, steps=None, s11=True, progress_every=0, batch=0, host_threads=8):
        self.precision, self.boundary, self.inflight = precision, boundary, inflight
        self.batch = batch
        # Batched mode: contexts of a group are created (uploads) and finished (traces,
        # S11, far field) by this many host threads; the library's calls release the GIL
        # and each context has its own stream.
        self.host_threads = max(1, host_threads)
        self.batch_tuning = None
        self.loop_seconds = 0.0       # batched runs only: launch to completion of the time loops
        self.autotune_seconds = 0.0
        self.tuning, self.keep_traces, self.farfield, self.steps = tuning, keep_traces, farfield, steps
        self.with_s11 = s11
        self.progress_every = progress_every
        self.cuts = standard_cuts()

    def _start(self, idx, problem, launch=True):
        freqs = frequency_grid(problem.header.get("dft"))
        sim = HipSim(problem, precision=self.precision, boundary=self.boundary,
                     port_dft_freqs=freqs if self.with_s11 else None, farfield=self.farfield)
        if self.tuning is not None and launch:
            sim.set_tuning(*self.tuning)
        steps = self.steps or problem.steps
        if launch:
            sim.run_async(0, steps)
        return {"idx": idx, "sim": sim, "problem": problem, "freqs": freqs, "steps": steps}

    def _finish(self, job):
        sim, p, freqs, steps = job["sim"], job["problem"], job["freqs"], job["steps"]
        v, i = sim.traces(0, steps)
        res = {"name": p.header.get("meta", {}).get("name"), "dims": list(p.dims), "steps": steps,
               "dt": p.grid["dt"]}
        if self.keep_traces:
            res["v"], res["i"] = v[0].copy(), i[0].copy()
        if not self.with_s11:
            sim.close()
            return res
        S = s11(sim, p)
        db = 20 * np.log10(np.abs(S))
        k = int(np.argmin(db))
        res.update({"s11": S, "s11_db": db, "resonance_hz": float(freqs[k]), "s11_min_db": float(db[k])})
        if self.farfield and sim.nfreq:
            ff = np.asarray(p.header["ff"]["frequencies"])
            f = int(np.argmin(np.abs(ff - freqs[k])))
            prad = sim.radiated_power()
            # Every cut at the one frequency used, in one call.
            names = list(self.cuts)
            th = np.concatenate([self.cuts[n][0] for n in names])
            ph = np.concatenate([self.cuts[n][1] for n in names])
            dt, dp = sim.directivity_at(f, th, ph, prad[f])
            cuts, a = {}, 0
            for n in names:
                m = len(self.cuts[n][0])
                cuts[n] = (dt[a:a + m], dp[a:a + m])
                a += m
            res.update({"farfield_hz": float(ff[f]), "prad_w": float(prad[f]), "cuts": cuts,
                        "peak_directivity_dbi": float(10 * np.log10(max(max(np.max(a), np.max(b))
                                                                        for a, b in cuts.values())))})
        sim.close()
        return res

    def _load(self, items, lazy, k):
        p = items.get(k) if lazy else items[k]
        path = p if isinstance(p, str) else None
        return read_problem(p) if path else p, path

    def _launch_group(self, jobs):
        """One batched launch over the group's contexts (tuned on the first group)."""
        steps = {j["steps"] for j in jobs}
        if len(steps) != 1:
            raise ValueError("a batched group needs one step count")
        if self.batch_tuning is None:
            if self.tuning is not None:
                self.batch_tuning = tuple(self.tuning) + (0,) * (4 - len(self.tuning))
            else:
                ta = time.perf_counter()
                m = HipMulti([j["sim"] for j in jobs])
                self.batch_tuning = m.autotune(2)
                m.close()
                self.autotune_seconds += time.perf_counter() - ta
        multi = HipMulti([j["sim"] for j in jobs], *self.batch_tuning)
        multi.t_launch = time.perf_counter()
        multi.run_async(0, steps.pop())
        return multi

    def _run_batched(self, items, lazy, t0):
        results = [None] * len(items)
        cell_steps = 0

        def group(first):
            ks = list(range(first, min(first + self.batch, len(items))))
            loaded = [self._load(items, lazy, k) for k in ks]
            jobs = list(pool.map(lambda kp: self._start(kp[0], kp[1][0], launch=False), zip(ks, loaded)))
            for p, path in loaded:
                if path and lazy:
                    os.remove(path)
            return jobs

        native.lib()                          # load once, before the worker threads call it
        pool = cf.ThreadPoolExecutor(max_workers=self.host_threads)
        nxt = 0
        jobs = group(nxt)
        while jobs:
            nxt += len(jobs)
            multi = self._launch_group(jobs)
            for j in jobs:
                cell_steps += int(np.prod(j["problem"].dims)) * j["steps"]
            following = group(nxt)            # uploads overlap the running group
            multi.synchronize()
            # Launch to completion; with several groups it also covers the next group's
            # uploads whenever those take longer than the loop.
            self.loop_seconds += time.perf_counter() - multi.t_launch
            multi.close()                     # before _finish closes the contexts it borrows
            for j, r in zip(jobs, pool.map(self._finish, jobs)):
                results[j["idx"]] = r
            if self.progress_every:
                print(f"[sweep] {nxt}/{len(items)} done, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
            jobs = following
        pool.shutdown()
        return results, time.perf_counter() - t0, cell_steps

    def run(self, problems):
        """problems: Problems, problem-file paths, or a LazyBuilds (files deleted after
        upload).  Returns (results, seconds, cell_steps)."""
        lazy = isinstance(problems, LazyBuilds)
        items = problems if lazy else list(problems)
        if self.batch > 0:
            return self._run_batched(items, lazy, time.perf_counter())
        results = [None] * len(items)
        running = []
        t0 = time.perf_counter()
        cell_steps = 0
        nxt = 0
        while nxt < len(items) or running:
            while nxt < len(items) and len(running) < self.inflight:
                p = items.get(nxt) if lazy else items[nxt]
                path = p if isinstance(p, str) else None
                p = read_problem(p) if path else p
                job = self._start(nxt, p)
                if lazy:
                    os.remove(path)            # the context holds its own copy now
                cell_steps += int(np.prod(p.dims)) * job["steps"]
                running.append(job)
                nxt += 1
            # Finish whichever job is done first; if none is, wait on the oldest.
            done = next((j for j in running if j["sim"].idle()), running[0])
            running.remove(done)
            results[done["idx"]] = self._finish(done)
            finished = sum(r is not None for r in results)
            if self.progress_every and finished % self.progress_every == 0:
                print(f"[sweep] {finished}/{len(items)} done, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        return results, time.perf_counter() - t0, cell_steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("scenarios", nargs="+")
    ap.add_argument("--workdir", default="/tmp/fdtd-sweep")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--boundary", default="matlab", choices=["matlab", "reference"])
    ap.add_argument("--inflight", type=int, default=4)
    ap.add_argument("--batch", type=int, default=64, help="variants per batched launch (0: streams in flight)")
    ap.add_argument("--tuning", default=None,
                    help="planes,cells,storage,cpml for every launch (default: the batch autotuner)")
    ap.add_argument("--steps", type=int, default=None, help="steps per variant (default: the scenario's)")
    ap.add_argument("--threads", type=int, default=8, help="host threads creating and finishing a batch's contexts")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    tuning = tuple(int(x) for x in a.tuning.split(",")) if a.tuning else None
    files = build_all(a.scenarios, a.workdir, fp32=a.precision == "fp32", jobs=16)
    sw = Sweep(a.precision, a.boundary, a.inflight, tuning=tuning, keep_traces=False, steps=a.steps, batch=a.batch,
               host_threads=a.threads)
    res, secs, cells = sw.run(files)
    summary = {"variants": len(res), "seconds": round(secs, 3), "mcell_updates_per_s": round(cells / secs / 1e6, 1),
               "loop_s": round(sw.loop_seconds, 3),
               "loop_mcell_updates_per_s": round(cells / sw.loop_seconds / 1e6, 1) if sw.loop_seconds else None,
               "results": [{k: v for k, v in r.items() if k not in ("s11", "s11_db", "cuts", "v", "i")} for r in res]}
    text = json.dumps(summary)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text)
    print(text)


if __name__ == "__main__":
    main()

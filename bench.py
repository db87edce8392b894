#!/usr/bin/env python3
"""Benchmark of the MI355X Yee FDTD time loop.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload roofline-256]

One "step" is one FDTD time step (H update, E update, port epilogue, NTFF when the
scenario has one) over the whole grid.  The default workload is the 256^3 FP32
stencil-roofline scenario (seeded random dielectric boxes, 8-cell CPML, centre
dipole; special-topics_amd/scenarios/roofline-256.json).  Prints one JSON line with
Mcell-updates/s (whole job), the HBM-roofline position of the dominant kernel
(timed with HIP events on the library's stream), and a CPU baseline: the oracle's
reference-structured loop (x-slab threads for the bulk, serial CPML/captures, FP64)
timed on this host for a bounded number of steps of the same scenario.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "special-topics_amd"))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FIELD_VALUES_PER_CELL_PHASE = 9  # 3 own-field reads + 3 writes + 3 neighbour-field reads
REF_VALUES_PER_CELL_PHASE = 18   # the same plus the reference's 9 per-cell coefficients
PSI_VALUES = 3                   # read Psi, write Psi, read CPsi


def cell_phase_bytes(esz: int, coef_mode: int) -> int:
    """Compulsory HBM bytes of one cell in one stencil phase for the coefficient storage
    in use: 9 field values, plus 3 per-cell coefficient indices (8- or 16-bit; the
    deduplicated coefficient tables themselves stay on chip) or 9 per-cell
    coefficient values."""
    if coef_mode == 1:
        return FIELD_VALUES_PER_CELL_PHASE * esz + 3
    if coef_mode == 2:
        return FIELD_VALUES_PER_CELL_PHASE * esz + 6
    return REF_VALUES_PER_CELL_PHASE * esz


def psi_counts(problem, axes="xyz"):
    """(H-phase Psi elements, E-phase Psi elements) of the problem's CPML faces along
    `axes`."""
    slot_comp = {"x": ["Hy", "Hz", "Ey", "Ez"], "y": ["Hx", "Hz", "Ez", "Ex"], "z": ["Hx", "Hy", "Ex", "Ey"]}
    nh = ne = 0
    for f in problem.header["cpml"]:
        if f["face"][0] not in axes:
            continue
        axis = "xyz".index(f["face"][0])
        for slot in range(4):
            d = list(problem.component_dims(slot_comp[f["face"][0]][slot]))
            d[axis] = f["nc"]
            n = d[0] * d[1] * d[2]
            if slot < 2:
                nh += n
            else:
                ne += n
    return nh, ne


def cpu_baseline(problem64, steps: int, threads: int, target_s: float = 15.0):
    """Oracle steps/s on the host; steps = 0 sizes the sample for about target_s."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle  # the checker, used here only as the CPU baseline
    sim = oracle.OracleSim(problem64, threads=threads)
    sim.run(0, 1)                      # first touch of all pages
    ts = 1
    if steps <= 0:
        t0 = time.perf_counter()
        sim.run(ts, 2)
        ts += 2
        per = (time.perf_counter() - t0) / 2
        steps = int(max(2, min(problem64.steps - ts, (target_s - 2 * per) / per)))
    t0 = time.perf_counter()
    sim.run(ts, steps)
    dt = time.perf_counter() - t0
    nx, ny, nz = problem64.dims
    del sim
    return nx * ny * nz * steps / dt / 1e6, dt, steps


def traffic_table(workload: str, precision: str, placement: int = 0) -> dict:
    """Per-kernel HBM bytes per launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (calibrated against a 1 GiB copy); {} if none.
    The CPML pass kernels keep their names across placements but not their tile lists,
    so each placement other than the separate pass has a file of its own
    (<workload>-<precision>-cpml<placement>-traffic.json)."""
    tag = "" if placement == 0 else f"-cpml{placement}"
    path = os.path.join(REPO, "profiles", f"{workload}-{precision}{tag}-traffic.json")
    try:
        with open(path) as f:
            return {k: int(v["bytes"]) for k, v in json.load(f)["kernels"].items()}
    except (OSError, ValueError, KeyError):
        return {}


def pmc_summary(fetch_dir: str, write_dir: str, copy_mib: int, out: str):
    """Per-kernel HBM bytes per dispatch from two rocprofv3 counter runs of the same
    command (`--pmc FETCH_SIZE`, then `--pmc WRITE_SIZE`, each with
    --copy-calibration-mib), each counter scaled so that the calibration copy
    (__amd_rocclr_copyBuffer of copy_mib MiB read and written) reads its true size."""
    import csv
    import glob

    def per_kernel(d, counter):
        rows = []
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                rows += [r for r in csv.DictReader(fh) if r["Counter_Name"] == counter]
        acc = {}
        for r in rows:
            acc.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
        # The calibration copy is the largest copy dispatch (uploads are copies too).
        return {k: (max(v) if k == "__amd_rocclr_copyBuffer" else sum(v) / len(v)) for k, v in acc.items()}

    fetch, write = per_kernel(fetch_dir, "FETCH_SIZE"), per_kernel(write_dir, "WRITE_SIZE")
    copy_kib = copy_mib * 1024
    fs = copy_kib / fetch["__amd_rocclr_copyBuffer"]
    ws = copy_kib / write["__amd_rocclr_copyBuffer"]
    kernels = {}
    for k in sorted(set(fetch) & set(write)):
        if k.startswith("fdtd::") or k.startswith("void fdtd::"):
            name = k[len("void "):] if k.startswith("void ") else k
            name = name.split("(")[0]
            kernels[name] = {"fetch_kib": fetch[k], "write_kib": write[k],
                             "bytes": (fetch[k] * fs + write[k] * ws) * 1024}
    doc = {"calibration": {"copy_bytes": copy_mib << 20, "fetch_scale": fs, "write_scale": ws}, "kernels": kernels}
    with open(out, "w") as f:
        json.dump(doc, f, indent=1)
    for k, v in kernels.items():
        print(f"{k}: {v['bytes'] / 1e6:.1f} MB/dispatch")


def standard_cuts():
    """The xy (theta = 90 deg) and xz cuts, 0..359 deg, of js_fdtd/src/index.js:77-92."""
    ang = np.arange(360) * np.pi / 180
    th_xz = np.where(ang > np.pi, 2 * np.pi - ang, ang)
    ph_xz = np.where(ang > np.pi, np.pi, 0.0)
    return (np.full(360, np.pi / 2), ang), (th_xz, ph_xz)


def s11_from_running_dft(sim, problem):
    """S11 of the source port from the on-GPU running port DFTs (sparameters.js:23-101)."""
    X = sim.port_dft()
    port = next(p for p in problem.header["ports"] if p.get("is_source_port"))
    z = port["impedance"] if isinstance(port["impedance"], (int, float)) else port["impedance"]["re"]
    V = X[port["sampled_voltage_index"] - 1]
    I = X[sim.n_sv + port["sampled_current_index"] - 1]
    a = 0.5 * (V + z * I) / np.sqrt(z)
    b = 0.5 * (V - z * I) / np.sqrt(z)
    return b / a


def ifa_run(problem, precision, boundary, freqs, cuts):
    """One IFA run to its outputs: S11 at `freqs` and directivity cuts at the NTFF
    frequencies.  Time-to-solution is the loop plus the on-GPU port DFT readback, S11
    and far-field evaluation; the autotuner runs before the clock starts."""
    from fdtdhip.sim import HipSim
    sim = HipSim(problem, precision=precision, boundary=boundary, port_dft_freqs=freqs)
    try:
        sim.autotune(3)
        sim.synchronize()
        t0 = time.perf_counter()
        sim.run_no_readback(0, problem.steps)
        s11 = s11_from_running_dft(sim, problem)
        prad = sim.radiated_power()
        d = [sim.directivity(th, ph, prad) for th, ph in cuts]
        secs = time.perf_counter() - t0
    finally:
        sim.close()
    return {"s11": s11, "directivity": d, "prad": prad, "secs": secs}


def ifa_study(workdir, scenario="ifa-config1-7000"):
    """FP32 vs FP64 on the IFA (S11 dB, resonance, directivity) and time-to-solution,
    with the reference's i = nx ghost plane ('reference') and without it ('matlab')."""
    from fdtdhip.problem import load_scenario_problem
    p = load_scenario_problem(scenario, workdir)
    dft = p.header["dft"] or {}
    step = dft.get("step", 20e6)
    freqs = (np.arange(int(np.floor(dft.get("maxFreq", 10e9) / step))) + 1) * step
    cuts = standard_cuts()
    nx, ny, nz = p.dims
    out = {"workload": f"{scenario} {nx}x{ny}x{nz}, {p.steps} steps, S11 at {freqs.size} f, "
                       f"NTFF at {len(p.header['ff'].get('frequencies', []))} f"}
    for boundary in ("matlab", "reference"):
        r64 = ifa_run(p, "fp64", boundary, freqs, cuts)
        r32 = ifa_run(p, "fp32", boundary, freqs, cuts)
        db64 = 20 * np.log10(np.abs(r64["s11"]))
        db32 = 20 * np.log10(np.abs(r32["s11"]))
        res64, res32 = freqs[np.argmin(db64)], freqs[np.argmin(db32)]
        ddir, dpeak = 0.0, 0.0
        for d64, d32 in zip(r64["directivity"], r32["directivity"]):
            for a64, a32 in zip(d64, d32):          # theta/phi components
                for f in range(a64.shape[0]):
                    m = a64[f] >= 1e-3 * a64[f].max()
                    if m.any():
                        ddir = max(ddir, float(np.max(np.abs(10 * np.log10(a32[f][m] / a64[f][m])))))
                    dpeak = max(dpeak, abs(10 * np.log10(a32[f].max() / a64[f].max())))
        out[f"boundary_{boundary}"] = {
            "fp64_s": round(r64["secs"], 4), "fp32_s": round(r32["secs"], 4),
            "fp32_speedup": round(r64["secs"] / r32["secs"], 2),
            "fp64_mcells": round(nx * ny * nz * p.steps / r64["secs"] / 1e6, 1),
            "fp32_mcells": round(nx * ny * nz * p.steps / r32["secs"] / 1e6, 1),
            "s11_max_abs_err_db": float(np.max(np.abs(db32 - db64))),
            "s11_min_db_fp64": round(float(db64.min()), 3),
            "resonance_ghz_fp64": float(res64 / 1e9), "resonance_err_mhz": float(abs(res32 - res64) / 1e6),
            "directivity_max_err_db_within_30db": ddir, "directivity_peak_err_db": dpeak,
        }
    return out


def sweep_study(workdir, count=64, batch=64, seed=4):
    """The sweep: `count` seeded IFA variants (uneven grids), FP32, `batch` of them stepped
    together by one launch per phase (tuned by the batch autotuner), 64-step chunks
    replayed from a captured graph."""
    sys.path.insert(0, os.path.join(REPO, "special-topics_amd", "scenarios"))
    import gen_ifa_variants
    from fdtdhip.sweep import Sweep, build_all
    wd = os.path.join(workdir, "ifa-variants")
    os.makedirs(wd, exist_ok=True)
    paths = []
    for i in range(count):
        sc = gen_ifa_variants.scenario(seed, i)
        path = os.path.join(wd, sc["name"] + ".json")
        with open(path, "w") as f:
            json.dump(sc, f)
        paths.append(path)
    t0 = time.perf_counter()
    files = build_all(paths, wd, fp32=True, jobs=max(1, min(host_cores(), 16)))
    build_s = time.perf_counter() - t0
    sweep = Sweep("fp32", "matlab", keep_traces=False, batch=batch)
    res, secs, cells = sweep.run(files)
    dims = np.array([r["dims"] for r in res])
    span = lambda a: f"{a.min()}-{a.max()}"
    res_ghz = np.array([r["resonance_hz"] for r in res]) / 1e9
    smin = np.array([r["s11_min_db"] for r in res])
    dpk = np.array([r["peak_directivity_dbi"] for r in res])
    return {
        "workload": f"{count} seeded IFA variants (gen_ifa_variants.py seed {seed}), grids "
                    f"{span(dims[:, 0])}x{span(dims[:, 1])}x{span(dims[:, 2])}, "
                    "FP32, 7000 steps, matlab boundary, S11 at 1000 f, NTFF at 16 f (matrix-core blocked)",
        "batch": batch, "batch_tuning": list(sweep.batch_tuning), "build_s": round(build_s, 2),
        "seconds": round(secs, 3),
        "mcell_updates_per_s": round(cells / secs / 1e6, 1), "ms_per_variant": round(secs / count * 1e3, 2),
        # The batched time loops alone (launch to completion), without uploads, the batch
        # autotuner, S11 or the far-field post.
        "loop_s": round(sweep.loop_seconds, 3), "autotune_s": round(sweep.autotune_seconds, 3),
        "loop_mcell_updates_per_s": round(cells / sweep.loop_seconds / 1e6, 1) if sweep.loop_seconds else None,
        "resonance_ghz_min_median_max": [round(float(x), 3) for x in (res_ghz.min(), np.median(res_ghz), res_ghz.max())],
        "s11_min_db_median": round(float(np.median(smin)), 2),
        "peak_directivity_dbi_median": round(float(np.median(dpk)), 2),
    }


EARLYSTOP_WEIGHTS = os.path.join(REPO, "special-topics_amd", "fdtdhip", "earlystop_weights.npz")


def early_stop_study(count=4, batch_count=64, first=4600):
    """Early-stop time-to-solution: the full-length FP64 run to S11 against the FP32 run
    stopped at step 1500 + the network (HIP kernels) + S11, with the FP64 fallback when
    the network's S11-error estimate exceeds its calibrated threshold; once with
    `batch_count` variants in one batched launch on each side (fallbacks re-run as their
    own FP64 batch, timed), once for `count` variants one at a time."""
    from fdtdhip import earlystop
    if not os.path.exists(EARLYSTOP_WEIGHTS):
        return {"skipped": f"no trained weights at {EARLYSTOP_WEIGHTS}"}
    out = {"workload": "IFA variants (gen_ifa_variants.py seed 4), 7000 steps, matlab boundary, S11 at 1000 f; "
                       "early stop at step 1500; variants from 4600 on are outside the training and held-out ranges"}
    progress(f"early stop: {batch_count} variants batched on each side")
    out["batched"] = earlystop.time_to_solution_batched(EARLYSTOP_WEIGHTS, first, batch_count)
    out["batched"].pop("per_variant", None)
    progress(f"early stop: {count} variants one at a time")
    out["single"] = earlystop.time_to_solution(EARLYSTOP_WEIGHTS, first, count)
    return out


def stencil_study(workload, precision, steps, warmup, timing_steps, workdir, ntff_mode="auto", fixed=None):
    """Whole-step throughput of `workload` and the HBM-roofline position of its
    dominant stencil kernel.

    Throughput: W untimed steps, then exactly K steps between device-wide syncs.
    Kernel durations: HIP events recorded on the library's stream around each launch
    kind, over `timing_steps` steps of a fresh run with the same tuning.  `achieved` is
    the kernel's algorithmic bytes for the coefficient storage in use (cell_phase_bytes
    per cell, plus the Psi bytes of CPML faces fused into it) over its average
    duration; `traffic` is what rocprofv3's calibrated FETCH_SIZE + WRITE_SIZE counted
    for the same kernel (profiles/<workload>-<precision>-traffic.json), and
    `traffic_frac` the fraction of peak those counted bytes give."""
    import torch
    from fdtdhip.problem import load_scenario_problem
    from fdtdhip.sim import HipSim

    fp32 = precision == "fp32"
    t_setup = time.perf_counter()
    progress(f"building {workload} ({precision})")
    problem = load_scenario_problem(workload, workdir, steps=max(warmup + steps, timing_steps + 1),
                                    fp32=fp32, no_phase=False)
    nx, ny, nz = problem.dims
    ncells = nx * ny * nz
    progress(f"uploading {nx}x{ny}x{nz}")
    sim = HipSim(problem, precision=precision)
    sim.set_ntff_mode(ntff_mode)
    progress("tuning")
    if fixed is None:
        tuning = sim.autotune(3)
    else:
        sim.set_tuning(*fixed)
        tuning = sim.tuning()
    setup_s = time.perf_counter() - t_setup

    progress(f"timing {steps} steps with {tuning}")
    sim.run_no_readback(0, warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.run_no_readback(warmup, steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    sim.close()

    sim = HipSim(problem, precision=precision)
    sim.set_ntff_mode(ntff_mode)
    sim.set_tuning(*tuning)
    sim.set_timing(True)
    sim.run_no_readback(0, timing_steps)
    ms, launches = sim.kernel_times()
    coef_mode = sim.coefficient_storage()[0]
    ntff_blocked = sim.ntff_blocked() if sim.nfreq else None
    sim.close()
    avg_ms = [m / max(n, 1) for m, n in zip(ms, launches)]

    esz = 4 if fp32 else 8
    nh, ne = psi_counts(problem)
    cb = cell_phase_bytes(esz, coef_mode)
    dom = 0 if avg_ms[0] >= avg_ms[1] else 1
    fused_axes = {0: "", 1: "z", 2: "xyz", 3: "z", 4: "xyz", 5: "xy"}[tuning[3]]
    fh, fe = psi_counts(problem, fused_axes) if fused_axes else (0, 0)
    psi_fused = PSI_VALUES * esz * (fh if dom == 0 else fe)
    dom_bytes = cb * ncells + psi_fused
    ref_bytes = REF_VALUES_PER_CELL_PHASE * esz * ncells + psi_fused
    dom_s = avg_ms[dom] * 1e-3
    achieved = dom_bytes / dom_s / 1e9

    ctype = "float" if fp32 else "double"
    # Six field arrays of at most 64 MiB: the library's stencils then use plain (cached)
    # loads and stores under placements 0 and 5 (the kernels' last template argument).
    sj = (nz + 1 + 3) // 4 * 4
    si = ((ny + 1) * sj + 1023) // 1024 * 1024
    stride = ((nx + 3) * si * esz + 255) // 256 * 256
    nt = "false" if 6 * stride <= 64 << 20 and tuning[3] in (0, 5) else "true"

    def stencil_name(phase):
        w = max(tuning[1], 1)
        if tuning[3] == 4:          # every slab cell in workgroups of their own
            return f"fdtd::k_update_{phase}_all<{ctype}, {coef_mode}, {w}>"
        if tuning[3] == 3:          # z slabs in workgroups of their own: the split kernels
            return f"fdtd::k_update_{phase}_split<{ctype}, {coef_mode}, {w}>"
        return f"fdtd::k_update_{phase}<{ctype}, {coef_mode}, {w}, {tuning[3]}, {nt}>"
    kname = stencil_name("h" if dom == 0 else "e")
    table = traffic_table(workload, precision, tuning[3])
    traffic = table.get(kname)
    step_s = elapsed / steps
    # Whole step, counted: the two stencils, the two CPML passes and the small E-tail
    # kernels as the timing steps launch them (separately).
    step_kernels = [stencil_name("h"), stencil_name("e")]
    if tuning[3] not in (2, 4):
        step_kernels += [f"fdtd::k_cpml_h<{ctype}>", f"fdtd::k_cpml_e<{ctype}>"]
    step_traffic = sum(table[k] for k in step_kernels) if all(k in table for k in step_kernels) else None
    step_bytes = 2 * cb * ncells + PSI_VALUES * esz * (nh + ne)

    return {
        "workload": f"{workload} {nx}x{ny}x{nz} {precision.upper()} fields+coefficients, FP64 port DFT/traces",
        "cells": ncells,
        "mcell_updates_per_s": round(ncells * steps / elapsed / 1e6, 1),
        "ms_per_step": round(step_s * 1e3, 4),
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_frac": round(traffic / dom_s / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
            "algorithmic_bytes_per_launch": dom_bytes,
            "bytes_per_cell": cb,
            "avg_launch_ms": round(avg_ms[dom], 5),
            # The same kernel priced as if it read the reference's 9 per-cell coefficient
            # arrays (18 values per cell): the rate the indexed storage is worth.
            "algorithmic_equivalent_frac": round(ref_bytes / dom_s / 1e9 / HBM_PEAK_GBS, 4),
        },
        "step_roofline": {
            "algorithmic_bytes": step_bytes,
            "frac": round(step_bytes / step_s / 1e9 / HBM_PEAK_GBS, 4),
            "traffic": step_traffic,
            "traffic_frac": round(step_traffic / step_s / 1e9 / HBM_PEAK_GBS, 4) if step_traffic else None,
        },
        "kernels_avg_ms": {"update_h": round(avg_ms[0], 5), "cpml_h": round(avg_ms[4], 5),
                           "update_e": round(avg_ms[1], 5), "cpml_e": round(avg_ms[5], 5),
                           "port_epilogue": round(avg_ms[2], 5), "ntff": round(avg_ms[3], 5)},
        "ntff_mode": None if ntff_blocked is None else ("blocked (matrix cores)" if ntff_blocked else "direct"),
        "setup_s": round(setup_s, 2),
        "tuning": {"planes_per_workgroup": tuning[0], "cells_per_thread": tuning[1],
                   "coefficient_storage": ["indexed", "arrays"][tuning[2]],
                   "cpml": ["separate pass", "z slabs fused into bulk", "all faces fused into bulk",
                            "z slabs in their own workgroups of the bulk launch",
                            "every slab cell in workgroups of its own in the bulk launch",
                            "x and y faces fused into bulk, z faces in the pass"][tuning[3]]},
    }


def progress(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=2000)
    # 64 warmup steps capture and replay the library's first 64-step graph, so the timed
    # steps are replays of an already captured step sequence.
    ap.add_argument("--warmup", type=int, default=64)
    ap.add_argument("--workload", default="roofline-256")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--timing-steps", type=int, default=100)
    ap.add_argument("--cpu-steps", type=int, default=0, help="0 = choose for ~20 s of CPU work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ifa", action="store_true", help="skip the IFA FP32-vs-FP64 time-to-solution study")
    ap.add_argument("--no-sweep", action="store_true", help="skip the 64-variant IFA sweep")
    ap.add_argument("--sweep-batch", type=int, default=64,
                    help="variants per batched launch in the sweep study (later groups upload while the previous one runs)")
    ap.add_argument("--no-early-stop", action="store_true", help="skip the early-stop time-to-solution study")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 roofline-256 sub-measurement")
    ap.add_argument("--no-patch512", action="store_true", help="skip the 512x512x384 sub-measurement")
    ap.add_argument("--ntff-mode", default="auto", choices=["auto", "direct", "blocked"],
                    help="NTFF running DFT: step-by-step (direct) or K-blocked on the matrix cores")
    ap.add_argument("--workdir", default=os.path.join(tempfile.gettempdir(), "fdtd-bench"))
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB (calibrates rocprof byte counters)")
    ap.add_argument("--planes", type=int, default=-1, help="x planes per workgroup (-1 = autotune)")
    ap.add_argument("--cells", type=int, default=0, help="cells per thread (with --planes)")
    ap.add_argument("--coef", type=int, default=0, help="coefficient storage (with --planes)")
    ap.add_argument("--fuse-cpml", type=int, default=0,
                    help="CPML placement with --planes: 0 separate pass, 1 z slabs fused, 2 all faces fused, "
                         "3 z slabs in workgroups of their own")
    ap.add_argument("--pmc-summary", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "OUT_JSON"),
                    help="write the calibrated per-kernel traffic summary of two rocprofv3 --pmc runs and exit")
    args = ap.parse_args()

    if args.pmc_summary:
        pmc_summary(args.pmc_summary[0], args.pmc_summary[1], args.copy_calibration_mib or 1024, args.pmc_summary[2])
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the time loop runs on one GPU; across GPUs it would split the grid into x-slabs that "
              "exchange one plane of halo per phase).", file=sys.stderr)
        sys.exit(2)

    import torch

    os.makedirs(args.workdir, exist_ok=True)
    if args.copy_calibration_mib:
        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    fixed = None if args.planes < 0 else (args.planes, args.cells, args.coef, args.fuse_cpml)
    st = stencil_study(args.workload, args.precision, args.steps, args.warmup, args.timing_steps, args.workdir,
                       args.ntff_mode, fixed)
    fp32 = args.precision == "fp32"
    line = {
        "metric": "Mcell-updates/s",
        "value": st["mcell_updates_per_s"],
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": st["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if fp32 else "f64",
        "data": "synthetic",
        "config": {"workload": st["workload"], "cells": st["cells"], "precision": args.precision,
                   "parallelism": "single GPU"},
    }
    line.update({k: v for k, v in st.items() if k not in ("mcell_updates_per_s", "ms_per_step", "workload", "cells")})

    if args.workload == "roofline-256" and not args.no_fp64:
        progress("FP64 roofline-256 (the reference's precision)")
        line["fp64_roofline_256"] = stencil_study("roofline-256", "fp64", 500, 64, 60, args.workdir, args.ntff_mode)

    if args.workload == "roofline-256" and not args.no_patch512:
        progress("patch-512 (512x512x384 FP32, NTFF at 16 f)")
        line["patch_512"] = stencil_study("patch-512", "fp32", 200, 64, 70, args.workdir, args.ntff_mode)

    if not args.no_ifa:
        progress("IFA FP32-vs-FP64 study")
        line["ifa"] = ifa_study(args.workdir)

    if not args.no_sweep:
        progress("64-variant IFA sweep")
        line["sweep"] = sweep_study(args.workdir, batch=args.sweep_batch)

    if not args.no_early_stop:
        progress("early-stop time-to-solution")
        line["early_stop"] = early_stop_study()

    if not args.no_cpu_baseline:
        from fdtdhip.problem import load_scenario_problem
        progress("CPU baseline")
        threads = max(1, min(host_cores() - 1, 15))   # the box gives one GPU a 16-core share
        total_steps = args.warmup + args.steps
        p64 = load_scenario_problem(args.workload, args.workdir, steps=max(total_steps, args.timing_steps + 1),
                                    fp32=False, no_phase=False)
        mc, secs, steps_cpu = cpu_baseline(p64, args.cpu_steps, threads)
        line["cpu_baseline"] = {
            "value": round(mc, 2), "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps_cpu} FP64 steps of {args.workload} ({secs:.1f} s): x-slab threads for the bulk "
                      "curls, CPML/captures on one thread, as the reference's wasm-cpu engine is structured",
        }
        # The reference's own published case: the IFA, 39x136x136, FP64 (published:
        # ~22 ms/step = 32.8 Mcell/s on 16 worker threads).
        pifa = load_scenario_problem("ifa-matlab-7000", args.workdir)
        mc, secs, steps_cpu = cpu_baseline(pifa, 0, threads, target_s=10.0)
        line["cpu_baseline"]["ifa_fp64"] = {
            "value": round(mc, 2), "unit": "Mcell-updates/s", "cores": threads,
            "ms_per_step": round(secs / steps_cpu * 1e3, 2),
            "sample": f"{steps_cpu} FP64 steps of ifa-matlab-7000 39x136x136 ({secs:.1f} s)",
        }
    print(json.dumps(line))


if __name__ == "__main__":
    main()

"""HIP time loop vs the reference golden vectors and the CPU oracle (needs the GPU).

FP64 results must be bit-identical: the kernels evaluate the reference's expression
trees without contraction.  FP32 results are compared against FP64 within stated
tolerances.
"""

import numpy as np
import pytest

import goldens as G
from oracle import OracleSim

pytestmark = pytest.mark.gpu


def _sim(problem, **kw):
    from fdtdhip.sim import HipSim
    return HipSim(problem, **kw)


def _check_run_golden(sim, fx, fields=True):
    steps = fx["steps"]
    v, i = sim.run(0, steps)
    for n, b in enumerate(fx["v"]):
        assert np.array_equal(v[n], G.f64(b)), f"V{n} differs from the reference trace"
    for n, b in enumerate(fx["i"]):
        assert np.array_equal(i[n], G.f64(b)), f"I{n} differs from the reference trace"
    if fields:
        for comp in ("Hx", "Hy", "Hz", "Ex", "Ey", "Ez"):
            assert G.sha(sim.field(comp)) == fx["fields"][comp], f"{comp} differs"
        for key, h in fx["fields"].items():
            if key.startswith("psi/"):
                _, face, nm = key.split("/")
                slot = G.PSI_NAMES[face[0]].index(nm)
                assert G.sha(sim.psi(face, slot)) == h, f"Psi {face}/{nm} differs"
    if fx.get("ff"):
        for n, name in enumerate(G.FF_ORDER):
            re, im = sim.farfield(n)
            assert G.sha(re) == fx["ff"][name]["re"] and G.sha(im) == fx["ff"][name]["im"], f"NTFF {name} differs"
    return v, i


@pytest.mark.parametrize("fuse", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("storage,expect", [(0, 1), (1, 0)])
@pytest.mark.parametrize("cells", [1, 2])
def test_golden_small_fp64_matches_reference_run(get_problem, storage, expect, cells, fuse):
    """Every coefficient storage (deduplicated indices, per-cell arrays), launch shape
    and CPML placement (separate pass, z slabs fused into the bulk kernels, every face
    fused, z slabs in workgroups of their own, every slab cell in workgroups of their
    own, x and y faces fused) reproduces the reference run bit for bit."""
    prob = get_problem("golden-small")
    fx = G.load("run-golden-small.json")
    sim = _sim(prob)
    sim.set_tuning(0, cells, storage, fuse)
    mode, entries = sim.coefficient_storage()
    assert mode == expect and (expect == 0 or 0 < max(entries) <= 256)
    _check_run_golden(sim, fx)
    sim.close()


@pytest.mark.parametrize("planes,cells", [(1, 1), (3, 2), (64, 1)])
def test_march_depth_does_not_change_results(get_problem, planes, cells):
    prob = get_problem("golden-small")
    fx = G.load("run-golden-small.json")
    sim = _sim(prob)
    sim.set_tuning(planes, cells, 0)
    _check_run_golden(sim, fx, fields=False)
    sim.close()


@pytest.mark.parametrize("fuse", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("cells", [1, 2, 4])
def test_fp32_launch_shapes_agree(get_problem, cells, fuse):
    """FP32 results do not depend on the launch shape or on where the z-slab CPML
    terms are applied (same arithmetic, same order per element)."""
    prob = get_problem("golden-small")
    ref = _sim(prob, precision="fp32")
    ref.set_tuning(1, 1, 1)
    v0, i0 = ref.run(0, 100)
    ref.close()
    sim = _sim(prob, precision="fp32")
    sim.set_tuning(5, cells, 0, fuse)
    v1, i1 = sim.run(0, 100)
    sim.close()
    assert np.array_equal(v0, v1) and np.array_equal(i0, i1)


def test_autotune_then_run_matches_reference(get_problem):
    """The autotuner runs candidate shapes on the live problem and then resets it."""
    prob = get_problem("golden-small")
    fx = G.load("run-golden-small.json")
    sim = _sim(prob)
    planes, cells, storage, fuse = sim.autotune(2)
    assert planes >= 1 and cells in (1, 2) and storage in (0, 1) and fuse in (0, 1, 2, 3, 4, 5)
    _check_run_golden(sim, fx)
    sim.close()


def test_random_state_one_step_fp64_matches_oracle(get_problem):
    """Seeded random fields and CPML state exercise every coefficient and Psi slot."""
    prob = get_problem("golden-small")
    fields, psi = G.random_state(prob, 7)
    sim = _sim(prob)
    ref = OracleSim(prob)
    for comp, a in fields.items():
        sim.set_field(comp, a)
        ref.field(comp)[:] = a
    for (face, slot), a in psi.items():
        sim.set_psi(face, slot, a)
        ref.psi(face, slot)[:] = a
    for ts in (0, 1, 2):
        v, i = sim.run(ts, 1)
        rv, ri = ref.run(ts, 1)
        assert np.array_equal(v[:, 0], rv[0]) and np.array_equal(i[:, 0], ri[0])
        for comp in fields:
            assert np.array_equal(sim.field(comp), ref.field(comp)), f"step {ts}: {comp}"
        for face, slot in psi:
            assert np.array_equal(sim.psi(face, slot), ref.psi(face, slot)), f"step {ts}: psi {face}/{slot}"
    sim.close()


@pytest.mark.parametrize("fuse", [0, 1, 2, 3, 4, 5])
def test_matlab_boundary_fp64_matches_oracle(get_problem, fuse):
    prob = get_problem("golden-small")
    sim = _sim(prob, boundary="matlab")
    sim.set_tuning(0, 0, 0, fuse)
    ref = OracleSim(prob, boundary="matlab")
    v, i = sim.run(0, 120)
    rv, ri = ref.run(0, 120)
    assert np.array_equal(v, rv.T) and np.array_equal(i, ri.T)
    for comp in ("Hx", "Hy", "Hz", "Ex", "Ey", "Ez"):
        assert np.array_equal(sim.field(comp), ref.field(comp)), comp
    sim.close()


def test_batches_compose(get_problem):
    """Running in uneven batches gives the same trace as one batch."""
    prob = get_problem("golden-small")
    a = _sim(prob)
    va, ia = a.run(0, 100)
    a.close()
    b = _sim(prob)
    parts = [b.run(s, n) for s, n in ((0, 1), (1, 37), (38, 50), (88, 12))]
    b.close()
    assert np.array_equal(va, np.concatenate([p[0] for p in parts], axis=1))
    assert np.array_equal(ia, np.concatenate([p[1] for p in parts], axis=1))


@pytest.mark.parametrize("scenario,fuse", [("ifa-matlab-7000", 0), ("ifa-matlab-7000", 2),
                                           ("ifa-matlab-7000", 3), ("ifa-config1-7000", 0),
                                           ("ifa-config1-7000", 1), ("ifa-config1-7000", 2),
                                           ("ifa-config1-7000", 3), ("ifa-config1-7000", 4),
                                           ("ifa-matlab-7000", 4), ("ifa-config1-7000", 5),
                                           ("ifa-matlab-7000", 5), ("ifa-fr4-substrate", 0),
                                           ("ifa-finefreq-sweep", 0)])
def test_ifa_fp64_full_run_matches_reference(get_problem, scenario, fuse):
    """The 39x136x136 IFA scenarios (7000 FP64 steps; 300 for the FR4-substrate and
    fine-frequency ones) against the reference's own per-step code."""
    prob = get_problem(scenario)
    fx = G.load(f"run-{scenario}.json")
    sim = _sim(prob)
    sim.set_tuning(0, 0, 0, fuse)
    _check_run_golden(sim, fx)
    sim.close()


@pytest.mark.parametrize("scenario", ["ifa-matlab-7000", "ifa-config1-7000"])
def test_ghost_plane_drives_the_late_time_growth(get_problem, scenario):
    """The reference's 7000-step IFA regrows from about step 5000; with the i = nx ghost
    plane skipped (boundary=matlab) the same run decays.  The growth is therefore the
    ghost plane's, not a CPML or source instability."""
    prob = get_problem(scenario)
    ref_v = G.f64(G.load(f"run-{scenario}.json")["v"][0])
    sim = _sim(prob, boundary="matlab")
    v, _ = sim.run(0, prob.steps)
    sim.close()
    win = lambda x, a: np.max(np.abs(x[a:a + 500]))
    assert np.all(np.isfinite(v))
    assert win(ref_v, 6500) > win(ref_v, 4500) * 100           # reference: growing
    assert win(v[0], 6500) < win(v[0], 4500) * 0.5               # matlab: still decaying
    assert win(v[0], 6500) < 1e-3 * win(v[0], 0)


def _ntff_runs(prob, precision, batches):
    out = {}
    for mode in ("direct", "blocked"):
        sim = _sim(prob, precision=precision)
        sim.set_ntff_mode(mode)
        assert sim.ntff_blocked() == (mode == "blocked")
        for s, n in batches:
            sim.run(s, n)
        out[mode] = [sim.farfield(q) for q in range(24)]
        sim.close()
    return out


def _assert_ntff_close(runs, rel):
    for (dr, di), (br, bi) in zip(runs["direct"], runs["blocked"]):
        scale = max(np.max(np.abs(dr)), np.max(np.abs(di)))
        assert scale > 0
        assert np.max(np.abs(br - dr)) <= rel * scale and np.max(np.abs(bi - di)) <= rel * scale


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_blocked_ntff_matches_direct(get_problem, precision):
    """The matrix-core K-blocked NTFF DFT equals the step-by-step accumulation to FP64
    rounding, across partial blocks at uneven batch boundaries."""
    prob = get_problem("golden-small")
    runs = _ntff_runs(prob, precision, ((0, 37), (37, 64), (101, 59)))
    _assert_ntff_close(runs, 1e-12)


def test_blocked_ntff_many_frequencies(get_problem, tmp_path):
    """20 NTFF frequencies: one launch for the first 16 (two 16-column phase tiles) and
    one for the remaining 4."""
    import json
    import os
    scen = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "scenarios", "golden-small.json")))
    scen["farfield"]["frequencies"] = [1e9 * (1 + 0.5 * q) for q in range(20)]
    path = tmp_path / "golden-small-ff20.json"
    path.write_text(json.dumps(scen))
    prob = get_problem(str(path))
    assert prob.header["ff"]["nFreq"] == 20
    _assert_ntff_close(_ntff_runs(prob, "fp64", ((0, 150),)), 1e-12)


def test_fp32_defaults_to_blocked_ntff(get_problem):
    prob = get_problem("golden-small")
    a, b = _sim(prob, precision="fp32"), _sim(prob)
    assert a.ntff_blocked() and not b.ntff_blocked()
    a.close()
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replays_match_direct_launches(get_problem, precision):
    """64-step chunks replayed from a captured HIP graph give the same traces, fields and
    NTFF accumulators, bit for bit, as launching every kernel directly -- over uneven
    batches (graph, direct remainder, a new graph after a partial NTFF block) with the
    host synchronising between them."""
    prob = get_problem("golden-small")
    batches = ((0, 70), (70, 64), (134, 26))
    out = {}
    for graphs in (False, True):
        sim = _sim(prob, precision=precision, port_dft_freqs=np.arange(1, 9) * 1e9)
        sim.set_graphs(graphs)
        parts = [sim.run(s, n) for s, n in batches]
        out[graphs] = (np.concatenate([p[0] for p in parts], axis=1), np.concatenate([p[1] for p in parts], axis=1),
                       [sim.field(c) for c in ("Hx", "Hy", "Hz", "Ex", "Ey", "Ez")],
                       [sim.farfield(q) for q in range(24)], sim.port_dft())
        sim.close()
    a, b = out[False], out[True]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for x, y in zip(a[2], b[2]):
        assert np.array_equal(x, y)
    for (xr, xi), (yr, yi) in zip(a[3], b[3]):
        assert np.array_equal(xr, yr) and np.array_equal(xi, yi)
    assert np.array_equal(a[4], b[4])


def test_port_running_dft_matches_trace_dft(get_problem):
    prob = get_problem("golden-small")
    freqs = np.arange(1, 41) * 0.5e9
    sim = _sim(prob, port_dft_freqs=freqs)
    v, i = sim.run(0, prob.steps)
    dft = sim.port_dft()
    sim.close()
    dt = prob.grid["dt"]
    n = np.arange(prob.steps)
    w = 2 * np.pi * freqs
    for o, (x, shift) in enumerate([(v[0], 0.0), (v[1], 0.0), (i[0], -dt / 2), (i[1], -dt / 2)]):
        t = (n + 1) * dt + shift
        ref = dt * (np.cos(np.outer(w, t)) @ x - 1j * (np.sin(np.outer(w, t)) @ x))
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(dft[o] - ref)) <= 1e-12 * scale


def test_port_dft_upload_order_does_not_matter(get_problem):
    """The running port DFT may be uploaded before the observers: its accumulators are
    sized at the first step from every observer uploaded by then, so both orders give
    the same transforms bit for bit, and reading them before any step gives zeros."""
    prob = get_problem("golden-small")
    freqs = np.arange(1, 41) * 0.5e9
    out = []
    for first in (False, True):
        sim = _sim(prob, port_dft_freqs=freqs, port_dft_first=first)
        assert not np.any(sim.port_dft())
        sim.run(0, prob.steps)
        out.append(sim.port_dft())
        sim.close()
    assert out[0].shape == (4, freqs.size) and np.any(out[0])
    assert np.array_equal(out[0], out[1])


def test_fp32_tracks_fp64(get_problem):
    prob = get_problem("golden-small")
    s64 = _sim(prob)
    v64, i64 = s64.run(0, prob.steps)
    s64.close()
    s32 = _sim(prob, precision="fp32")
    v32, i32 = s32.run(0, prob.steps)
    s32.close()
    for a, b in ((v32, v64), (i32, i64)):
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        assert err < 2e-4, err


def test_kernel_timing_counts_launches(get_problem):
    prob = get_problem("golden-small")
    sim = _sim(prob)
    sim.set_timing(True)
    sim.run(0, 5)
    ms, launches = sim.kernel_times()
    sim.close()
    assert launches[0] == 5 and launches[1] == 5 and ms[0] > 0 and ms[1] > 0

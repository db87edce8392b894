"""Point field probes (captureElectricFields / captureMagneticFields,
js_fdtd/src/sampling.js:144-292).

CPU: the oracle's probe capture is pinned by closed forms -- uniform fields give the
field value back exactly, the magnitude of a known vector, and a single non-zero
element reaches exactly the probe cells the reference's index offsets name, with
weight 1/2 (E) or 1/4 (H).  GPU: the HIP probes equal the oracle's bit for bit in
FP64 (both boundary modes), track them in FP32, agree between a batched and a single
run, come back through the Node backend, and out-of-range probes are refused.
"""

import copy
import itertools
import json
import math
import os

import numpy as np
import pytest

import goldens as G

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCEN = os.path.join(REPO, "tests", "golden", "scenarios", "golden-probes.json")
STEPS = 160


def _oracle(problem, probes=None):
    from oracle import OracleSim
    if probes is not None:
        problem = copy.copy(problem)
        problem.header = dict(problem.header, probes=probes)
    return OracleSim(problem)


def _field_ids():
    from oracle import FIELD_IDS
    return FIELD_IDS


def test_scenario_carries_every_probe_kind(get_problem):
    prob = get_problem(SCEN)
    kinds = {(p["magnetic"], p["component"]) for p in prob.header["probes"]}
    assert kinds == set(itertools.product((0, 1), range(4)))


def test_oracle_probes_of_uniform_fields(get_problem):
    """A uniform field averages to itself exactly; 'm' is the Euclidean norm."""
    prob = get_problem(SCEN)
    o = _oracle(prob)
    ids = _field_ids()
    vals = {"Hx": 2.5, "Hy": -1.0, "Hz": 0.5, "Ex": 1.5, "Ey": -2.0, "Ez": 4.0}
    for name, v in vals.items():
        o.array(ids[name])[:] = v
    got_h, got_e = o.capture_probes(True), o.capture_probes(False)
    for n, pr in enumerate(prob.header["probes"]):
        f = "H" if pr["magnetic"] else "E"
        got = (got_h if pr["magnetic"] else got_e)[n]
        if pr["component"] < 3:
            assert got == vals[f + "xyz"[pr["component"]]]
        else:
            want = math.sqrt(sum(vals[f + c] ** 2 for c in "xyz"))
            assert got == want


@pytest.mark.parametrize("magnetic", [0, 1])
@pytest.mark.parametrize("comp", [0, 1, 2])
def test_oracle_probe_offsets_follow_the_reference(get_problem, magnetic, comp):
    """One non-zero element at (i0, j0, k0) of F_comp.  The reference averages E_x
    over x-1..x, E_y over y-1..y, E_z over z-1..z, and H_x over (y-1..y, z-1..z),
    H_y over (x-1..x, z-1..z), H_z over (x-1..x, y-1..y): so exactly the probe cells
    one step up along those axes see it, with weight 1/2 or 1/4."""
    prob = get_problem(SCEN)
    i0, j0, k0 = 9, 8, 10
    cells = list(itertools.product((i0, i0 + 1), (j0, j0 + 1), (k0, k0 + 1)))
    probes = [{"magnetic": magnetic, "component": comp, "i": i, "j": j, "k": k} for i, j, k in cells]
    o = _oracle(prob, probes)
    name = ("H" if magnetic else "E") + "xyz"[comp]
    arr = o.field(name)
    arr[:] = 0.0
    nx, ny, nz = o.nx, o.ny, o.nz
    ext = {"Hx": (1, 0, 0), "Hy": (0, 1, 0), "Hz": (0, 0, 1), "Ex": (0, 1, 1), "Ey": (1, 0, 1), "Ez": (1, 1, 0)}[name]
    d1, d2 = ny + ext[1], nz + ext[2]
    arr[(i0 * d1 + j0) * d2 + k0] = 8.0
    got = o.capture_probes(bool(magnetic))
    axes = [comp] if not magnetic else [a for a in range(3) if a != comp]
    for n, (i, j, k) in enumerate(cells):
        up = (i - i0, j - j0, k - k0)
        sees = all(up[a] in (0, 1) for a in axes) and all(up[a] == 0 for a in range(3) if a not in axes)
        want = (8.0 / (2 if not magnetic else 4)) if sees else 0.0
        assert got[n] == want, (name, (i, j, k), got[n], want)


@pytest.mark.gpu
@pytest.mark.parametrize("boundary", ["reference", "matlab"])
def test_hip_probes_match_oracle_fp64(get_problem, boundary):
    from fdtdhip.sim import HipSim
    prob = get_problem(SCEN, steps=STEPS)
    sim = HipSim(prob, boundary=boundary)
    v, i = sim.run(0, STEPS)
    p = sim.probes(0, STEPS)
    sim.close()
    from oracle import OracleSim
    rv, ri, rp = OracleSim(prob, boundary=boundary).run_probes(0, STEPS)
    assert np.array_equal(v, rv.T) and np.array_equal(i, ri.T)
    assert np.array_equal(p, rp.T), "probe traces differ from the oracle"
    assert np.all(np.abs(p).max(axis=1) > 0), "every probe should see the pulse"


@pytest.mark.gpu
def test_hip_probes_fp32_track_fp64(get_problem):
    from fdtdhip.sim import HipSim
    p64 = HipSim(get_problem(SCEN, steps=STEPS))
    p64.run(0, STEPS)
    a = p64.probes(0, STEPS)
    p64.close()
    p32 = HipSim(get_problem(SCEN, steps=STEPS, fp32=True), precision="fp32")
    p32.run(0, STEPS)
    b = p32.probes(0, STEPS)
    p32.close()
    scale = np.abs(a).max(axis=1, keepdims=True)
    assert np.all(np.abs(a - b) <= 1e-4 * scale), np.max(np.abs(a - b) / scale)


@pytest.mark.gpu
def test_batched_probes_match_single(get_problem):
    from fdtdhip.sim import HipMulti, HipSim
    prob = get_problem(SCEN, steps=STEPS)
    single = HipSim(prob)
    single.run(0, STEPS)
    want = single.probes(0, STEPS)
    single.close()
    sims = [HipSim(prob), HipSim(get_problem("golden-small", steps=STEPS))]
    m = HipMulti(sims)
    m.run(0, STEPS)
    got = sims[0].probes(0, STEPS)
    m.close()
    for s in sims:
        s.close()
    assert np.array_equal(got, want)


@pytest.mark.gpu
def test_out_of_range_probe_is_refused(get_problem):
    from fdtdhip import native
    from fdtdhip.sim import HipSim
    sim = HipSim(get_problem("golden-small"))
    with pytest.raises(native.FdtdHipError, match="outside"):
        native.check(sim.L.fdtd_hip_upload_field_probe(sim.h, 0, 0, 0, 5, 5), "upload_field_probe")  # Ex(-1,..)
    with pytest.raises(native.FdtdHipError, match="component"):
        native.check(sim.L.fdtd_hip_upload_field_probe(sim.h, 1, 4, 5, 5, 5), "upload_field_probe")
    sim.close()


@pytest.mark.gpu
def test_node_backend_returns_probes(tmp_path, get_problem):
    """backend.run fills every probe's sampled_value with the library's traces."""
    import subprocess
    from fdtdhip.sim import HipSim
    job = {"scenario": SCEN, "steps": STEPS, "backend": "hip", "post": False, "farfield": False,
           "gpuFarfield": False, "options": {}, "out": str(tmp_path / "rec.json")}
    (tmp_path / "job.json").write_text(json.dumps(job))
    res = subprocess.run(["node", os.path.join(REPO, "tests", "js", "backend_driver.mjs"), str(tmp_path / "job.json")],
                         capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr[-3000:]
    rec = json.loads((tmp_path / "rec.json").read_text())
    sim = HipSim(get_problem(SCEN, steps=STEPS))
    sim.run(0, STEPS)
    want = sim.probes(0, STEPS)
    sim.close()
    assert len(rec["probes"]) == len(want)
    for n, b in enumerate(rec["probes"]):
        assert np.array_equal(G.f64(b), want[n]), f"probe {n} differs"

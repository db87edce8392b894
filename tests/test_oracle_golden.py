"""The CPU oracle against the reference's golden vectors.

steps-golden-small.json holds the outputs of the reference's own per-step functions
(updateMagneticCPML, captureSampledCurrents, updateElectricCPML, injectVoltageSources,
captureSampledVoltages, accumulateFarfieldDFT) applied to a seeded random state; the
oracle must reproduce each bit for bit.  run-*.json are full loops (reference
per-step code plus a restatement of the WASM bulk kernels, see gen_golden.mjs).
"""

import numpy as np

import goldens as G
from oracle import OracleSim


def _load_state(ref, prob, seed):
    fields, psi = G.random_state(prob, seed)
    for comp, a in fields.items():
        ref.field(comp)[:] = a
    for (face, slot), a in psi.items():
        ref.psi(face, slot)[:] = a
    return fields, psi


def test_per_step_functions_match_reference(get_problem):
    prob = get_problem("golden-small")
    fx = G.load("steps-golden-small.json")
    ref = OracleSim(prob)
    _load_state(ref, prob, fx["seed"])
    psi_key = lambda face, slot: f"{face}/{G.PSI_NAMES[face[0]][slot]}"
    faces = [f["face"] for f in prob.header["cpml"]]

    ref.update_h_cpml()
    for comp in ("Hx", "Hy", "Hz"):
        assert G.sha(ref.field(comp)) == fx["cpml_h"]["fields"][comp], comp
    for face in faces:
        for slot in (0, 1):
            assert G.sha(ref.psi(face, slot)) == fx["cpml_h"]["psi"][psi_key(face, slot)]

    assert [G.hexd(x) for x in ref.capture_currents()] == fx["capture_i"]

    ref.update_e_cpml()
    for comp in ("Ex", "Ey", "Ez"):
        assert G.sha(ref.field(comp)) == fx["cpml_e"]["fields"][comp], comp
    for face in faces:
        for slot in (2, 3):
            assert G.sha(ref.psi(face, slot)) == fx["cpml_e"]["psi"][psi_key(face, slot)]

    ref.inject(40)
    for comp in ("Ex", "Ey", "Ez"):
        assert G.sha(ref.field(comp)) == fx["inject"][comp], comp
    assert [G.hexd(x) for x in ref.capture_voltages()] == fx["capture_v"]

    ref.ff_accumulate(7)
    ref.ff_accumulate(8)
    for n, name in enumerate(G.FF_ORDER):
        assert G.sha(ref.ff(n, False)) == fx["ff"][name]["re"], name
        assert G.sha(ref.ff(n, True)) == fx["ff"][name]["im"], name


def test_full_run_golden_small(get_problem):
    prob = get_problem("golden-small")
    fx = G.load("run-golden-small.json")
    ref = OracleSim(prob)
    v, i = ref.run(0, fx["steps"])
    for n, b in enumerate(fx["v"]):
        assert np.array_equal(v[:, n], G.f64(b))
    for n, b in enumerate(fx["i"]):
        assert np.array_equal(i[:, n], G.f64(b))
    for comp in ("Hx", "Hy", "Hz", "Ex", "Ey", "Ez"):
        assert G.sha(ref.field(comp)) == fx["fields"][comp]
    for n, name in enumerate(G.FF_ORDER):
        assert G.sha(ref.ff(n, False)) == fx["ff"][name]["re"]


def test_threaded_oracle_is_bit_identical(get_problem):
    """x-slab threads (the reference's worker partition) do not change a bit."""
    prob = get_problem("golden-small")
    a = OracleSim(prob, threads=1)
    b = OracleSim(prob, threads=5)
    va, ia = a.run(0, 80)
    vb, ib = b.run(0, 80)
    assert np.array_equal(va, vb) and np.array_equal(ia, ib)
    for comp in ("Ey", "Ez"):
        assert np.array_equal(a.field(comp), b.field(comp))


def test_ifa_prefix_matches_reference(get_problem):
    """First 600 steps of the IFA loop (the full 7000 run against the same fixture is
    a GPU test; the oracle needs ~1 minute for it)."""
    prob = get_problem("ifa-matlab-7000")
    fx = G.load("run-ifa-matlab-7000.json")
    ref = OracleSim(prob, threads=4)
    v, i = ref.run(0, 600)
    assert np.array_equal(v[:, 0], G.f64(fx["v"][0])[:600])
    assert np.array_equal(i[:, 0], G.f64(fx["i"][0])[:600])

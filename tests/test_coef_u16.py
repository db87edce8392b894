"""The 16-bit coefficient-index path (more than 256 distinct (self, curl, curl)
coefficient triples in a component).

A seeded scenario of 180 small dielectric bricks with distinct eps_r and sigma (the
material grid is 8-bit, js_fdtd/src/materials.js:21-68, but the edge averaging of
materials.js:87-238 and the CPML grading multiply the distinct coefficient values)
gives several hundred triples per E component.  CPU: the count, from the builder's
arrays.  GPU: the library picks 16-bit indices for it, and the FP64 run equals the
oracle bit for bit.
"""

from __future__ import annotations

import json
import os

import numpy as np
import pytest

from oracle import OracleSim

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TRIPLES = {"Ex": ("Cexe", "Cexhz", "Cexhy"), "Ey": ("Ceye", "Ceyhx", "Ceyhz"), "Ez": ("Ceze", "Cezhy", "Cezhx")}


def scenario(path):
    with open(os.path.join(REPO, "tests", "golden", "scenarios", "golden-small.json")) as f:
        sc = json.load(f)
    rng = np.random.default_rng(16)
    mats = sc["materialTypes"][:3]
    bricks = []
    for q in range(180):
        mats.append({"eps_r": round(float(rng.uniform(1.5, 12.0)), 4), "mu_r": 1,
                     "sigma_e": round(float(rng.uniform(0.0, 0.05)), 5), "sigma_m": 0})
        x0, y0, z0 = (int(v) for v in rng.integers(0, 9, size=3))
        bricks.append({"min_x": x0 * 1e-3, "min_y": round(y0 * 0.8e-3, 7), "min_z": round(z0 * 1.2e-3, 7),
                       "max_x": (x0 + 2) * 1e-3, "max_y": round((y0 + 2) * 0.8e-3, 7),
                       "max_z": round((z0 + 2) * 1.2e-3, 7), "material_type": len(mats)})
    sc.update({"name": "coef-u16", "numberOfTimeSteps": 60, "materialTypes": mats,
               "bricks": bricks + sc["bricks"][1:], "spheres": []})
    with open(path, "w") as f:
        json.dump(sc, f)
    return path


@pytest.fixture(scope="module")
def problem(problem_dir):
    from fdtdhip.problem import load_scenario_problem
    return load_scenario_problem(scenario(os.path.join(problem_dir, "coef-u16.json")), problem_dir)


def distinct_triples(problem, comp):
    cols = [np.asarray(problem.coeff(n), dtype=np.float64) for n in TRIPLES[comp]]
    return len(np.unique(np.stack(cols, axis=1), axis=0))


def test_more_than_256_triples(problem):
    counts = {c: distinct_triples(problem, c) for c in TRIPLES}
    print("distinct E coefficient triples:", counts)
    assert max(counts.values()) > 256 and max(counts.values()) <= 65536


@pytest.mark.gpu
def test_u16_indices_fp64_match_oracle(problem):
    from fdtdhip.sim import HipSim
    sim = HipSim(problem, precision="fp64")
    try:
        mode, entries = sim.coefficient_storage()
        assert mode == 2 and max(entries) > 256, (mode, entries)
        v, i = sim.run(0, problem.steps)
        fields = {c: sim.field(c) for c in ("Hx", "Hy", "Hz", "Ex", "Ey", "Ez")}
    finally:
        sim.close()
    ref = OracleSim(problem)
    rv, ri = ref.run(0, problem.steps)
    assert np.array_equal(v, rv.T) and np.array_equal(i, ri.T)
    for c, a in fields.items():
        assert np.array_equal(a, ref.field(c)), c

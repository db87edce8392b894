"""CPsi slots whose elements are all equal are read at one element (Face::cmask).

A CPML face slot's CPsi array (cpml.js:163-210) is the curl coefficient of its slab
cells times the cell size, so it holds one value wherever the slab is one material,
as in every standard scenario here; the kernels then read element 0 only.  Where a
material reaches into the slab the array varies and is streamed per element.

golden-small with its air buffers removed puts its bricks' materials into the CPML:
each face then has uniform (H) and varying (E) slots.  CPU: that mix, from the
builder's arrays, and the standard scenario's all-uniform slots.  GPU: the FP64 run
equals the oracle bit for bit (traces, fields, every Psi) in every CPML placement.
"""

from __future__ import annotations

import copy
import json
import os

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STEPS = 60


def uniform_slots(problem):
    """{face: [True if every CPsi element of slot s has the same bits]}."""
    out = {}
    for f in problem.header["cpml"]:
        face = f["face"]
        row = []
        for s in range(4):
            a = np.ascontiguousarray(problem.arrays[f"cpml/{face}/CPsi{s}"], dtype=np.float64).view(np.uint64)
            row.append(bool(a.size) and bool(np.all(a == a[0])))
        out[face] = row
    return out


@pytest.fixture(scope="module")
def no_air(problem_dir):
    with open(os.path.join(REPO, "tests", "golden", "scenarios", "golden-small.json")) as f:
        sc = copy.deepcopy(json.load(f))
    for k in sc["boundary"]:
        if k.startswith("air_buffer"):
            sc["boundary"][k] = 0
    sc["name"] = "golden-small-no-air"
    path = os.path.join(problem_dir, sc["name"] + ".json")
    with open(path, "w") as f:
        json.dump(sc, f)
    return path


def test_standard_scenario_slots_are_uniform(get_problem):
    slots = uniform_slots(get_problem("golden-small", steps=STEPS))
    assert all(all(v) for v in slots.values()), slots


def test_materials_in_the_slab_vary_cpsi(get_problem, no_air):
    slots = uniform_slots(get_problem(no_air, steps=STEPS))
    flat = [u for v in slots.values() for u in v]
    assert any(flat) and not all(flat), slots


@pytest.mark.gpu
@pytest.mark.parametrize("fuse", [0, 1, 2, 3, 4, 5])
def test_varying_cpsi_fp64_matches_oracle(get_problem, no_air, fuse):
    from fdtdhip.sim import HipSim
    from oracle import OracleSim
    prob = get_problem(no_air, steps=STEPS)
    sim = HipSim(prob)
    sim.set_tuning(0, 0, 0, fuse)
    ref = OracleSim(prob)
    v, i = sim.run(0, STEPS)
    rv, ri = ref.run(0, STEPS)
    assert np.array_equal(v, rv.T) and np.array_equal(i, ri.T)
    for comp in ("Hx", "Hy", "Hz", "Ex", "Ey", "Ez"):
        assert np.array_equal(sim.field(comp), ref.field(comp)), comp
    for f in prob.header["cpml"]:
        for slot in range(4):
            assert np.array_equal(sim.psi(f["face"], slot), ref.psi(f["face"], slot)), f"Psi {f['face']}/{slot}"
    assert np.abs(sim.field("Ez")).max() > 0
    sim.close()

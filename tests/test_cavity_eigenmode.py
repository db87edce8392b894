"""The bulk curl updates against the mathematics of the Yee scheme: discrete eigenmodes
of a PEC cavity.

In an empty rectangular cavity with PEC walls, E_z(i, j) = sin(m pi i / nx) sin(n pi j / ny)
(uniform along z), with H = 0, is the E part of a discrete eigenmode of the Yee update
(js_fdtd/assembly/fdtd_kernels.ts:20-210 with air coefficients): the field keeps its
shape exactly, and its amplitude follows the leapfrog recurrence
    e(n+1) = (2 - lam) e(n) - e(n-1),   e(1) = (1 - lam) e(0),
    lam = dt^2/(eps0 mu0) * ((2/dx sin(m pi/2nx))^2 + (2/dy sin(n pi/2ny))^2),
whose solution is e(n) = cos((n + 1/2) theta) / cos(theta/2) with cos(theta) = 1 - lam/2
-- the Yee discrete dispersion relation sin(w dt/2) = (c dt/2) K.  The same holds for E_x
(modes in y, z) and E_y (in x, z); superposed, the three stay independent, so one run
exercises every E and H component through both of its curl terms.  A wrong neighbour
offset, sign or coefficient slot breaks either the shape or the frequency.

The outer E planes are PEC only without the reference's i = nx ghost update, so the
cavity runs with the matlab boundary (the ghost plane is pinned by the golden runs).
CPU: the oracle.  GPU: the HIP kernels, FP64 and FP32, direct launches and graphs.
"""

from __future__ import annotations

import math
import os

import numpy as np
import pytest

from oracle import OracleSim

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCENARIO = os.path.join(REPO, "tests", "golden", "scenarios", "pec-cavity.json")
EPS0, MU0 = 8.854187817e-12, 4 * math.pi * 1e-7
# Mode numbers of each E component along its two transverse axes (axis index, mode).
MODES = {"Ez": ((0, 2), (1, 3)), "Ex": ((1, 2), (2, 3)), "Ey": ((0, 3), (2, 2))}
CHECK_STEPS = (1, 2, 3, 4, 64, 128, 192, 256, 320, 384)


@pytest.fixture(scope="module")
def cavity(problem_dir):
    from fdtdhip.problem import load_scenario_problem
    return load_scenario_problem(SCENARIO, problem_dir)


def mode_shape(problem, comp):
    """Shape of `comp`'s mode in the reference layout (zero on the PEC walls)."""
    n = problem.dims
    dims = problem.component_dims(comp)
    grids = np.meshgrid(*[np.arange(d) for d in dims], indexing="ij")
    s = np.ones(dims)
    for axis, m in MODES[comp]:
        s = s * np.sin(m * math.pi * grids[axis] / n[axis])
    return s.ravel()


def theta(problem, comp):
    g = problem.grid
    d = (g["dx"], g["dy"], g["dz"])
    n = problem.dims
    K2 = sum((2 / d[a] * math.sin(m * math.pi / (2 * n[a]))) ** 2 for a, m in MODES[comp])
    lam = g["dt"] ** 2 / (EPS0 * MU0) * K2
    return math.acos(1 - lam / 2)


def expected(problem, comp, step):
    th = theta(problem, comp)
    return math.cos((step + 0.5) * th) / math.cos(th / 2)


def check(problem, read, step, amp_tol, shape_tol):
    for comp in MODES:
        s = mode_shape(problem, comp)
        e = read(comp)
        a = float(e @ s / (s @ s))
        assert abs(a - expected(problem, comp, step)) <= amp_tol, (comp, step, a, expected(problem, comp, step))
        assert np.max(np.abs(e - a * s)) <= shape_tol, (comp, step)
    for comp in ("Hx", "Hy", "Hz"):
        assert np.all(np.isfinite(read(comp)))


def test_dispersion_relation_is_the_yee_one(cavity):
    """theta(dt) matches w dt for the Yee relation sin(w dt/2) = (c dt / 2) K, and the
    three modes oscillate at different frequencies (so a swapped axis would show)."""
    ths = [theta(cavity, c) for c in MODES]
    assert len(set(round(t, 12) for t in ths)) == 3
    g = cavity.grid
    c = 1 / math.sqrt(EPS0 * MU0)
    for comp, th in zip(MODES, ths):
        K = math.sqrt(sum((2 / (g["dx"], g["dy"], g["dz"])[a] * math.sin(m * math.pi / (2 * cavity.dims[a]))) ** 2
                          for a, m in MODES[comp]))
        assert abs(math.sin(th / 2) - c * g["dt"] / 2 * K) <= 1e-15


def test_oracle_follows_discrete_eigenmodes(cavity):
    sim = OracleSim(cavity, boundary="matlab")
    for comp in MODES:
        sim.field(comp)[:] = mode_shape(cavity, comp)
    done = 0
    for step in CHECK_STEPS:
        sim.run(done, step - done)
        done = step
        check(cavity, lambda c: np.array(sim.field(c)), step, 1e-12, 1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,cells,storage,amp_tol,shape_tol", [
    ("fp64", 1, 0, 1e-12, 1e-12), ("fp64", 2, 1, 1e-12, 1e-12), ("fp32", 4, 0, 5e-5, 5e-5)])
def test_hip_follows_discrete_eigenmodes(cavity, precision, cells, storage, amp_tol, shape_tol):
    from fdtdhip.sim import HipSim
    sim = HipSim(cavity, precision=precision, boundary="matlab")
    sim.set_tuning(0, cells, storage, 0)
    try:
        for comp in MODES:
            sim.set_field(comp, mode_shape(cavity, comp))
        print(f"cavity eigenmodes: {precision}, {cells} cells per thread, storage {storage}", flush=True)
        done = 0
        for step in CHECK_STEPS:          # single steps first, then 64-step graph replays
            sim.run_no_readback(done, step - done)
            done = step
            check(cavity, sim.field, step, amp_tol, shape_tol)
    finally:
        sim.close()

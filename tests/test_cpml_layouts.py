"""The CPML pass over uneven slab layouts (needs the GPU): some faces absent (a plain
wall there instead of CPML), unequal depths, in every CPML placement, both boundary
modes and one- and two-cell vectors.  FP64 must match the oracle bit for bit (traces,
fields, every Psi); FP32 must give the same bits in every launch shape and placement
(a vector of 2 or 4 cells may start before a slab or end past it, at either end); the
batched launch must equal each single run.

The standard scenarios carry all six faces.  These layouts give tiles whose mask holds
one face of an axis, and boxes that end at a wall with no slab, which the
cpml_tiles_check.cpp enumeration covers on the host.
"""

import copy
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STEPS = 90
LAYOUTS = {
    "no-xp-yn-zp": {"type_xp": "pec", "type_yn": "pec", "type_zp": "pec", "cpml_cells_zn": 6},
    "x-only": {"type_yn": "pec", "type_yp": "pec", "type_zn": "pec", "type_zp": "pec",
               "cpml_cells_xn": 6, "cpml_cells_xp": 3},
}
FACES = ("xn", "xp", "yn", "yp", "zn", "zp")


@pytest.fixture(scope="module")
def layouts(problem_dir):
    with open(os.path.join(REPO, "tests", "golden", "scenarios", "golden-small.json")) as f:
        base = json.load(f)
    paths = {}
    for name, change in LAYOUTS.items():
        sc = copy.deepcopy(base)
        sc["name"] = f"golden-small-{name}"
        sc["boundary"].update(change)
        path = os.path.join(problem_dir, sc["name"] + ".json")
        with open(path, "w") as f:
            json.dump(sc, f)
        paths[name] = path
    return paths


def _present(problem):
    return [f["face"] for f in problem.header["cpml"]]


@pytest.mark.parametrize("cells", [1, 2])
@pytest.mark.parametrize("boundary", ["reference", "matlab"])
@pytest.mark.parametrize("fuse", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("layout", sorted(LAYOUTS))
def test_uneven_layout_fp64_matches_oracle(get_problem, layouts, layout, fuse, boundary, cells):
    from fdtdhip.sim import HipSim
    from oracle import OracleSim
    prob = get_problem(layouts[layout], steps=STEPS)
    faces = _present(prob)
    assert 0 < len(faces) < 6
    sim = HipSim(prob, boundary=boundary)
    sim.set_tuning(0, cells, 0, fuse)
    ref = OracleSim(prob, boundary=boundary)
    v, i = sim.run(0, STEPS)
    rv, ri = ref.run(0, STEPS)
    assert np.array_equal(v, rv.T) and np.array_equal(i, ri.T)
    for comp in ("Hx", "Hy", "Hz", "Ex", "Ey", "Ez"):
        assert np.array_equal(sim.field(comp), ref.field(comp)), comp
    for face in faces:
        for slot in range(4):
            assert np.array_equal(sim.psi(face, slot), ref.psi(face, slot)), f"Psi {face}/{slot}"
    assert np.abs(sim.field("Ez")).max() > 0
    sim.close()


@pytest.mark.parametrize("fuse", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("cells", [2, 4])
@pytest.mark.parametrize("layout", sorted(LAYOUTS))
def test_uneven_layout_fp32_shapes_agree(get_problem, layouts, layout, cells, fuse):
    """FP32 in every launch shape and placement gives the bits of one-cell vectors with
    the separate pass: the arithmetic and its order per element are the same."""
    from fdtdhip.sim import HipSim
    prob = get_problem(layouts[layout], steps=STEPS, fp32=True)
    faces = _present(prob)
    runs = []
    for tuning in ((1, 1, 0, 0), (3, cells, 0, fuse)):
        sim = HipSim(prob, precision="fp32")
        sim.set_tuning(*tuning)
        v, i = sim.run(0, STEPS)
        runs.append((v, i, [sim.field(c) for c in ("Hx", "Hy", "Hz", "Ex", "Ey", "Ez")],
                     [sim.psi(f, s) for f in faces for s in range(4)]))
        sim.close()
    (v0, i0, f0, p0), (v1, i1, f1, p1) = runs
    assert np.array_equal(v0, v1) and np.array_equal(i0, i1)
    assert all(np.array_equal(a, b) for a, b in zip(f0, f1))
    assert all(np.array_equal(a, b) for a, b in zip(p0, p1))
    assert np.abs(f0[5]).max() > 0


def test_uneven_layouts_batched_match_single(get_problem, layouts):
    """Both layouts and the full six-face scenario in one batched launch (FP32)."""
    from fdtdhip.sim import HipMulti, HipSim
    probs = [get_problem(layouts[n], steps=STEPS, fp32=True) for n in sorted(LAYOUTS)]
    probs.append(get_problem("golden-small", steps=STEPS, fp32=True))
    want = []
    for p in probs:
        s = HipSim(p, precision="fp32")
        s.set_tuning(0, 2, 0, 0)
        v, i = s.run(0, STEPS)
        want.append((v, i, s.field("Ey"), s.field("Hz")))
        s.close()
    sims = [HipSim(p, precision="fp32") for p in probs]
    m = HipMulti(sims, cells_per_thread=2, fuse_cpml=0)
    got = m.run(0, STEPS)
    for n, s in enumerate(sims):
        v, i = got[n]
        assert np.array_equal(v, want[n][0]) and np.array_equal(i, want[n][1])
        assert np.array_equal(s.field("Ey"), want[n][2]) and np.array_equal(s.field("Hz"), want[n][3])
    m.close()
    for s in sims:
        s.close()

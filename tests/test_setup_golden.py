"""The JavaScript problem builder against the reference's own set-up (golden hashes).

Bit-exact: material index grid, all 18 coefficient arrays, CPML profiles and CPsi
arrays, source edges / coefficients / waveforms, observer index lists and Csvf.
"""

import numpy as np
import pytest

import goldens as G

SCENARIOS = ["ifa-matlab-7000", "ifa-config1-7000", "ifa-fr4-substrate", "ifa-finefreq-sweep", "golden-small"]


@pytest.mark.parametrize("scenario", SCENARIOS)
def test_builder_matches_reference_setup(get_problem, scenario):
    P = get_problem(scenario if scenario != "golden-small" else "golden-small")
    fx = G.load(f"setup-{scenario}.json")
    g = P.grid
    assert (g["nx"], g["ny"], g["nz"]) == (fx["grid"]["nx"], fx["grid"]["ny"], fx["grid"]["nz"])
    assert G.hexd(g["dt"]) == fx["grid"]["dt"]
    assert G.hexd(g["min_x"]) == fx["grid"]["min_x"] and G.hexd(g["min_z"]) == fx["grid"]["min_z"]
    assert G.sha(P.arrays["mat"]) == fx["mat"], "material index grid"
    for name, h in fx["coeffs"].items():
        assert G.sha(P.coeff(name)) == h, name
    assert {f["face"] for f in P.header["cpml"]} == set(fx["cpml"])
    for f in P.header["cpml"]:
        rec = fx["cpml"][f["face"]]
        assert f["nc"] == rec["nc"]
        for p in ("b_e", "a_e", "b_m", "a_m"):
            assert [G.hexd(x) for x in P.arrays[f"cpml/{f['face']}/{p}"]] == rec[p], (f["face"], p)
        for s in range(4):
            assert G.sha(P.arrays[f"cpml/{f['face']}/CPsi{s}"]) == rec[f"CPsi{s}"], (f["face"], s)
    for n, vs in enumerate(fx["vsources"]):
        assert list(P.arrays[f"vsrc/{n}/idx"]) == vs["idx"]
        assert [G.hexd(x) for x in P.arrays[f"vsrc/{n}/coef"]] == vs["coef"]
        assert G.sha(P.arrays[f"vsrc/{n}/vpe"]) == vs["vpe"]
    for n, sv in enumerate(fx["svoltages"]):
        assert list(P.arrays[f"sv/{n}/idx"]) == sv["idx"]
        assert G.hexd(P.header["svoltages"][n]["Csvf"]) == sv["Csvf"]
    if fx["ff"]:
        ff = P.header["ff"]
        assert [ff["li"], ff["lj"], ff["lk"], ff["ui"], ff["uj"], ff["uk"]] == fx["ff"]["box"]
        assert [G.hexd(x) for x in P.arrays["ff/w"]] == fx["ff"]["w"]


def test_ifa_closed_forms(get_problem):
    """Facts the reference documents independently of its code: the IFA grid is
    39x136x136 with dt = 5.770246941973866e-13 s; the feed has 21 edges (3 Ex cells
    across 7 z-edges), the port voltage integrates 18 Ex edges with Csvf = -dx/6, free
    space has Cexe = 1 and Cexhz = dt/(eps0 dy)."""
    P = get_problem("ifa-matlab-7000")
    g = P.grid
    assert (g["nx"], g["ny"], g["nz"]) == (39, 136, 136)
    assert g["dt"] == 5.770246941973866e-13
    assert P.header["vsources"][0]["count"] == 21
    assert P.header["svoltages"][0]["count"] == 18
    assert P.header["svoltages"][0]["Csvf"] == -g["dx"] / 6
    eps0 = 8.854187817e-12
    cexe, cexhz = P.coeff("Cexe"), P.coeff("Cexhz")
    nx, ny, nz = 39, 136, 136
    # an interior free-space Ex edge away from every object and the CPML
    idx = (12 * (ny + 1) + 68) * (nz + 1) + 12   # air: past the 8 CPML cells, before the substrate
    assert cexe[idx] == 1.0
    assert np.isclose(cexhz[idx], g["dt"] / (eps0 * g["dy"]), rtol=1e-15)
    # PEC plates (sigma = 1e10) give Cexe = (2 eps0 - dt sigma)/(2 eps0 + dt sigma) ~ -1
    assert np.min(cexe) < -0.999999


def test_problem_file_fp32_variant(get_problem):
    P64 = get_problem("golden-small")
    P32 = get_problem("golden-small", fp32=True)
    for name in ("Cexe", "Chzey"):
        assert P32.coeff(name).dtype == np.float32
        assert np.array_equal(P32.coeff(name), P64.coeff(name).astype(np.float32))

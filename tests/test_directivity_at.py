"""fdtd_hip_directivity_at: the directivity at one NTFF frequency.

It runs fdtd_hip_directivity's kernels for that frequency alone, so its outputs must be
the bits of that frequency's row of the all-frequency result, with and without a
caller-supplied radiated power; a frequency outside the NTFF set fails cleanly.  The
batched sweep takes its pattern cuts at each variant's resonance through it.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_directivity_at_equals_its_row_of_every_frequency(get_problem, precision):
    from fdtdhip import native
    from fdtdhip.sim import HipSim
    from fdtdhip.sweep import standard_cuts
    prob = get_problem("golden-small")
    sim = HipSim(prob, precision=precision)
    try:
        sim.run(0, prob.steps)
        print(f"directivity_at: golden-small {precision}, {prob.steps} steps", flush=True)
        cuts = standard_cuts()
        th = np.concatenate([cuts["xy"][0], cuts["xz"][0][:97]])     # not a multiple of the angle group
        ph = np.concatenate([cuts["xy"][1], cuts["xz"][1][:97]])
        prad = sim.radiated_power()
        nf = len(prad)
        assert nf == 3
        for rp in (prad, None):
            want_t, want_p = sim.directivity(th, ph, rp)
            for f in range(nf):
                got_t, got_p = sim.directivity_at(f, th, ph, None if rp is None else rp[f])
                assert np.all(np.isfinite(got_t)) and np.all(np.isfinite(got_p))
                assert np.array_equal(got_t, want_t[f]) and np.array_equal(got_p, want_p[f]), f
        for f in (-1, nf):
            with pytest.raises(native.FdtdHipError, match="outside"):
                sim.directivity_at(f, th, ph)
    finally:
        sim.close()


def test_directivity_bits_do_not_depend_on_the_angle_count(get_problem):
    """Each angle's sums run over the surface cells in one order whatever the other angles of
    the call are (its block's group of four included), so a prefix of the angles, down to a
    partial group, gives the bits of the full call, and one frequency those of its row."""
    from fdtdhip.sim import HipSim
    prob = get_problem("golden-small")
    sim = HipSim(prob, precision="fp64")
    try:
        sim.run(0, prob.steps)
        rng = np.random.default_rng(11)
        th = rng.uniform(0.0, np.pi, 4100)
        ph = rng.uniform(0.0, 2 * np.pi, 4100)
        prad = sim.radiated_power()
        assert len(prad) == 3
        want_t, want_p = sim.directivity(th, ph, prad)
        for n in (1000, 601, 5):
            print(f"directivity: golden-small fp64, {n} of 4100 angles", flush=True)
            got_t, got_p = sim.directivity(th[:n], ph[:n], prad)
            assert np.array_equal(got_t, want_t[:, :n]) and np.array_equal(got_p, want_p[:, :n]), n
        for f in range(3):
            got_t, got_p = sim.directivity_at(f, th[:720], ph[:720], prad[f])
            assert np.array_equal(got_t, want_t[f, :720]) and np.array_equal(got_p, want_p[f, :720]), f
    finally:
        sim.close()

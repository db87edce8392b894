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

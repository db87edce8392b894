"""fdtd_hip_directivity_noalloc: the directivity entry point that allocates nothing.

It runs the same kernels as fdtd_hip_directivity on the library's fixed per-device
workspace, so its outputs must equal that entry point's bit for bit, also on a second
call with fewer angles (the workspace keeps the previous call's contents), with and
without a caller-supplied radiated power; requests beyond the workspace fail cleanly.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cuts(n):
    ang = np.arange(n) * (2 * np.pi / n)
    return np.full(n, np.pi / 2), ang, np.where(ang > np.pi, 2 * np.pi - ang, ang), np.where(ang > np.pi, np.pi, 0.0)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_directivity_equals_allocating_entry_point(get_problem, precision):
    from fdtdhip import native
    from fdtdhip.sim import HipSim
    prob = get_problem("golden-small")
    sim = HipSim(prob, precision=precision)
    try:
        sim.run(0, prob.steps)
        print(f"noalloc directivity: golden-small {precision}, {prob.steps} steps", flush=True)
        th_xy, ph_xy, th_xz, ph_xz = _cuts(360)
        prad = sim.radiated_power()
        for th, ph in ((th_xy, ph_xy), (th_xz[:97], ph_xz[:97])):       # second call: fewer angles
            for rp in (None, prad):
                want = sim.directivity(th, ph, rp)
                got = sim.directivity_noalloc(th, ph, rp)
                for w, g in zip(want, got):
                    assert w.shape == g.shape and np.all(np.isfinite(w))
                    assert np.array_equal(w, g), float(np.max(np.abs(w - g)))
        # The largest request the workspace holds goes through; one more angle does not.
        big = np.linspace(0, np.pi, 8192)
        sim.directivity_noalloc(big, big)
        with pytest.raises(native.FdtdHipError, match="fixed workspace"):
            sim.directivity_noalloc(np.zeros(8193), np.zeros(8193))
    finally:
        sim.close()

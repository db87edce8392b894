"""The early-stop representation (CPU): the pooled tail reproduces the full-rate port
DFT (js_fdtd/src/dft.js:17-38 conventions) for signals band-limited to 6 GHz, and the
input/target extraction keeps its documented shapes."""

import numpy as np

from fdtdhip import earlystop as ES


def _ringing(dt, n, f0, tau, phase=0.3):
    t = (np.arange(n) + 1) * dt
    return np.exp(-t / tau) * np.sin(2 * np.pi * f0 * t + phase)


def test_prefix_plus_pooled_tail_matches_full_dft():
    """Ringing that decays to 1e-4..1e-2 of its peak by step 7000, as the variants'
    port transients do; S11 from prefix + pooled tail within 0.02 dB of full rate."""
    dt = 7.785185503791985e-13           # the IFA variants' step
    n = ES.N_PREFIX + ES.TAIL
    freqs = np.arange(1, 1001) * 6e6
    for f0, tau in ((2.4e9, 0.8e-9), (5.5e9, 0.6e-9), (3.1e9, 1.0e-9), (2.0e9, 1.2e-9)):
        v = _ringing(dt, n, f0, tau)
        i = 0.02 * _ringing(dt, n, f0, tau, phase=1.1)
        w = 2 * np.pi * freqs
        k = np.arange(n)
        V = dt * (np.exp(-1j * np.outer(w, (k + 1) * dt)) @ v)
        I = dt * (np.exp(-1j * np.outer(w, (k + 1) * dt - dt / 2)) @ i)
        x, scale = ES.features(v, i)
        tail, logr = ES.targets(v, i, scale)
        assert x.shape == (2, ES.N_IN) and tail.shape == (2, ES.N_OUT)
        Vp, Ip = ES.prefix_dft(v, i, dt, freqs)
        Vt, It = ES.tail_dft(tail, scale, dt, freqs)
        s_full = ES.s11_from(V, I)
        s_split = ES.s11_from(Vp + Vt, Ip + It)
        err_db = np.abs(20 * np.log10(np.abs(s_split)) - 20 * np.log10(np.abs(s_full)))
        assert np.max(err_db) < 0.02, (f0, np.max(err_db))
        e = (v / scale[0]) ** 2 + (i / scale[1]) ** 2
        assert np.isclose(logr, np.log10(e[ES.N_PREFIX:].sum() / e.sum()), atol=1e-5)

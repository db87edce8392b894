"""The matrix-pencil tail extrapolation (fdtdhip/extrapolate.py), pinned on signals whose
continuation is known in closed form, and its DFT / S11 conventions against the early-stop
pipeline's (which follow js_fdtd/src/dft.js:17-38 and js_fdtd/src/sparameters.js:65-101).
CPU only."""

from __future__ import annotations

import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "special-topics_amd"))

from fdtdhip import earlystop as es          # noqa: E402
from fdtdhip import extrapolate as ex        # noqa: E402
from fdtdhip.sweep import frequency_grid     # noqa: E402

# Damped modes of a ringing port: (decay per step, radians per step); V and I share them.
MODES = [(2.0e-4, 0.0087), (6.0e-4, 0.0151), (1.5e-3, 0.0262)]


def ringdown(n, amps, phases):
    t = np.arange(n)
    return sum(a * np.exp(-d * t) * np.cos(w * t + p) for (d, w), a, p in zip(MODES, amps, phases))


def test_pencil_recovers_shared_poles():
    n = np.arange(400)
    z_true = np.array([np.exp(-8 * d + 1j * 8 * w) for d, w in MODES])
    z_true = np.concatenate([z_true, z_true.conj()])
    rng = np.random.default_rng(3)
    ys = [(rng.standard_normal(6) @ (z_true[:, None] ** n[None, :])).real for _ in range(2)]
    z = ex.matrix_pencil(ys, tol=1e-9)
    assert z.size == z_true.size
    d = np.abs(z[:, None] - z_true[None, :]).min(axis=1)
    assert d.max() < 1e-9, d


def test_extend_continues_a_ringdown_exactly():
    v = ringdown(7000, (1.0, 0.4, 0.2), (0.3, 1.1, -0.7))
    i = ringdown(7000, (0.02, -0.01, 0.005), (1.2, 0.1, 2.0))
    ve, ie = ex.extend(v, i, start=1100, n=2000, dec=8, tol=1e-9)
    assert np.array_equal(ve[:2000], v[:2000]) and np.array_equal(ie[:2000], i[:2000])
    assert np.abs(ve - v).max() < 1e-8 * np.abs(v).max()
    assert np.abs(ie - i).max() < 1e-8 * np.abs(i).max()
    # A continuation is needed: leaving the tail out is far off.
    assert np.abs(v[2000:]).max() > 0.1 * np.abs(v).max()


def test_growing_poles_are_held_on_the_unit_circle():
    t = np.arange(3000)
    v = np.exp(4e-4 * t) * np.cos(0.01 * t)
    ve, _ = ex.extend(v, 0.5 * v, start=500, n=1500, dec=4, tol=1e-9)
    assert np.abs(ve[1500:]).max() <= 1.01 * np.abs(v[:1500]).max()


def test_dft_and_s11_follow_the_pipeline_conventions():
    rng = np.random.default_rng(5)
    v, i = rng.standard_normal(es.N_PREFIX), 0.02 * rng.standard_normal(es.N_PREFIX)
    dt, freqs = 5.77e-13, frequency_grid(ex.FREQS)
    V, I = ex.port_dft(v, i, dt, freqs)
    Vp, Ip = es.prefix_dft(v, i, dt, freqs)
    assert np.allclose(V, Vp, rtol=1e-12, atol=0) and np.allclose(I, Ip, rtol=1e-12, atol=0)
    assert np.allclose(ex.s11_db(V, I), 20 * np.log10(np.abs(es.s11_from(V, I))), rtol=0, atol=1e-12)

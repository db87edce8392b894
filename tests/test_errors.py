"""The C ABI's error paths on a live device: every refused call returns FDTD_HIP_ERR
with a message naming the problem, and leaves the context usable (needs the GPU).

The reference addon throws "<call>: <fdtd_last_error()>" for the same classes of
misuse (js_fdtd/src/backends/cuda/native/fdtd_cuda.h:12-13,22); here the Python
binding raises FdtdHipError with the library's message.
"""

import ctypes as C

import numpy as np
import pytest

import goldens as G

pytestmark = pytest.mark.gpu


def _sim(problem, **kw):
    from fdtdhip.sim import HipSim
    return HipSim(problem, **kw)


def _raises(match, fn, *args):
    from fdtdhip import native
    with pytest.raises(native.FdtdHipError, match=match):
        fn(*args)


def test_init_rejects_bad_dimensions():
    from fdtdhip import native
    L = native.lib()
    h = C.c_void_p()
    for dims, match in [(dict(nx=1, ny=8, nz=8), "too small"), (dict(nx=8, ny=8, nz=8, precision=7), "precision"),
                        (dict(nx=8, ny=8, nz=8, boundary=9), "boundary"),
                        (dict(nx=8, ny=8, nz=8, numberOfTimeSteps=0), "positive")]:
        d = dict(dx=1e-3, dy=1e-3, dz=1e-3, dt=1e-12, numberOfTimeSteps=4, precision=0, boundary=0)
        d.update(dims)
        rc = L.fdtd_hip_init(C.byref(native.FdtdHipDims(**d)), C.byref(h))
        assert rc != 0 and match in L.fdtd_hip_last_error().decode()


def test_uploads_check_lengths_and_ranges(get_problem):
    from fdtdhip import native
    prob = get_problem("golden-small")
    sim = _sim(prob)
    L = sim.L
    short = np.zeros(5)
    p = short.ctypes.data_as(C.POINTER(C.c_double))
    _raises("expected", lambda: native.check(L.fdtd_hip_upload_coeff(sim.h, 0, p, 5), "upload_coeff"))
    _raises("out of range", lambda: native.check(L.fdtd_hip_upload_coeff(sim.h, 99, p, 5), "upload_coeff"))
    _raises("FP64", lambda: native.check(
        L.fdtd_hip_upload_coeff_f32(sim.h, 0, np.zeros(5, np.float32).ctypes.data_as(C.POINTER(C.c_float)), 5),
        "upload_coeff_f32"))
    idx = np.array([10 ** 9], dtype=np.int32)
    _raises("outside", lambda: native.check(L.fdtd_hip_upload_svoltage(
        sim.h, 0, idx.ctypes.data_as(C.POINTER(C.c_int)), 1, 1.0), "upload_svoltage"))
    _raises("bad arguments", lambda: native.check(L.fdtd_hip_upload_port_dft(sim.h, 0, None), "upload_port_dft"))
    one = np.ones(1)
    native.check(L.fdtd_hip_upload_port_dft(sim.h, 1, one.ctypes.data_as(C.POINTER(C.c_double))), "upload_port_dft")
    _raises("already uploaded", lambda: native.check(L.fdtd_hip_upload_port_dft(
        sim.h, 1, one.ctypes.data_as(C.POINTER(C.c_double))), "upload_port_dft"))
    # The context still runs and still matches the reference after the refused calls.
    fx = G.load("run-golden-small.json")
    v, i = sim.run(0, fx["steps"])
    assert np.array_equal(v[0], G.f64(fx["v"][0]))
    sim.close()


def test_run_ranges_and_empty_batches(get_problem):
    prob = get_problem("golden-small")
    sim = _sim(prob)
    steps = prob.steps
    _raises("outside", sim.run, steps - 1, 2)
    _raises("outside", sim.run, -1, 1)
    v, i = sim.run(0, 0)                        # an empty batch is a no-op
    assert v.shape[1] == 0 and i.shape[1] == 0
    _raises("outside", sim.traces, 0, steps + 1)
    fx = G.load("run-golden-small.json")
    v, _ = sim.run(0, fx["steps"])
    assert np.array_equal(v[0], G.f64(fx["v"][0]))
    sim.close()


def test_tuning_and_field_reads_are_checked(get_problem):
    from fdtdhip import native
    prob = get_problem("golden-small")
    sim = _sim(prob)
    _raises("cells per thread", sim.set_tuning, 0, 4, 0, 0)       # 4 cells is FP32 only
    _raises("cells per thread", sim.set_tuning, 0, 3, 0, 0)
    _raises("fuse_cpml", sim.set_tuning, 0, 1, 0, 6)
    _raises("coefficient storage", sim.set_tuning, 0, 1, 7, 0)
    out = np.zeros(3)
    _raises("elements", lambda: native.check(sim.L.fdtd_hip_read_field(
        sim.h, 0, out.ctypes.data_as(C.POINTER(C.c_double)), 3), "read_field"))
    _raises("out of range", lambda: native.check(sim.L.fdtd_hip_read_field(
        sim.h, 6, out.ctypes.data_as(C.POINTER(C.c_double)), 3), "read_field"))
    sim.close()


def test_batch_rejects_misuse(get_problem):
    from fdtdhip import native
    from fdtdhip.sim import HipMulti
    prob = get_problem("golden-small")
    a, b = _sim(prob), _sim(prob)
    _raises("repeats", HipMulti, [a, a])
    m = HipMulti([a, b])
    _raises("outside", m.run_async, 0, prob.steps + 1)
    # A problem that has stepped cannot be autotuned (that resets it).
    m.run(0, 3)
    _raises("before the first time step", m.autotune, 2)
    m.close()
    L = native.lib()
    h = C.c_void_p()
    arr = (C.c_void_p * 1)(a.h)
    assert L.fdtd_hip_multi_create(arr, 0, None, C.byref(h)) != 0
    assert "variants" in L.fdtd_hip_last_error().decode()
    a.close()
    b.close()


def test_batch_members_close_after_the_batch(get_problem):
    """A batch borrows its contexts: closing one while the batch lives is refused and
    leaves the batch usable; after the batch is closed the contexts close normally."""
    from fdtdhip.sim import HipMulti
    prob = get_problem("golden-small")
    a, b = _sim(prob), _sim(prob)
    m = HipMulti([a, b])
    with pytest.raises(RuntimeError, match="live HipMulti"):
        a.close()
    v, _ = m.run(0, 3)[0]
    assert v.shape[1] == 3
    m.close()
    a.close()
    b.close()
    assert a.h is None and b.h is None

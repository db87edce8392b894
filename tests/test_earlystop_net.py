"""The early-stop network's HIP inference kernels against an FP32 NumPy forward of the
same BF16-rounded weights (CPU pieces: packing and the reference itself)."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from fdtdhip import earlystop as ES
from fdtdhip import earlystop_net as NET


def _random_weights(seed=0):
    rng = np.random.default_rng(seed)
    n_w = (NET.CH * 2 * NET.KW + 3 * NET.CH * NET.CH * NET.KW + NET.HIDDEN * NET.CH * NET.LENS[-1]
           + NET.N_OUT * NET.HIDDEN)
    n_b = 4 * NET.CH + NET.HIDDEN + NET.N_OUT
    w = NET.bf16_bits((rng.standard_normal(n_w) * 0.05).astype(np.float32))
    b = (rng.standard_normal(n_b) * 0.01).astype(np.float32)
    return w, b


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-30], dtype=np.float32)
    r = NET.bf16_to_f32(NET.bf16_bits(x))
    assert r[0] == 1.0 and r[1] == 1.0            # halfway rounds to even
    assert r[2] == np.float32(1.0078125)          # above halfway rounds up
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)


def test_reference_forward_matches_torch():
    torch = pytest.importorskip("torch")
    model = NET.make_model()
    w, b = NET.pack(model)
    # Load the BF16-rounded weights back so both sides use identical values.
    wo = bo = 0
    f = NET.bf16_to_f32(w)
    with torch.no_grad():
        for m in model:
            if hasattr(m, "weight"):
                n = m.weight.numel()
                m.weight.copy_(torch.from_numpy(f[wo:wo + n].reshape(m.weight.shape)))
                wo += n
                m.bias.copy_(torch.from_numpy(b[bo:bo + m.bias.numel()]))
                bo += m.bias.numel()
        x = np.random.default_rng(1).standard_normal((3, 2, ES.N_IN)).astype(np.float32)
        ref = model(torch.from_numpy(x)).numpy()
    out = NET.reference_forward(w, b, x)
    assert np.allclose(out, ref, rtol=1e-4, atol=1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("batch", [1, 7, 64])
def test_hip_inference_matches_reference(batch):
    from fdtdhip import native
    L = native.lib()
    w, b = _random_weights()
    h = C.c_void_p()
    native.check(L.fdtd_hip_es_create(w.ctypes.data_as(C.POINTER(C.c_uint16)), w.size,
                                      b.ctypes.data_as(C.POINTER(C.c_float)), b.size, C.byref(h)), "es_create")
    try:
        x = np.random.default_rng(batch).standard_normal((batch, 2, ES.N_IN)).astype(np.float32)
        y = np.zeros((batch, NET.N_OUT), dtype=np.float32)
        native.check(L.fdtd_hip_es_predict(h, batch, x.ctypes.data_as(C.POINTER(C.c_float)),
                                           y.ctypes.data_as(C.POINTER(C.c_float))), "es_predict")
    finally:
        L.fdtd_hip_es_destroy(h)
    ref = NET.reference_forward(w, b, x)
    assert np.max(np.abs(y - ref)) <= 1e-4 * max(1.0, np.max(np.abs(ref)))

"""ctypes binding of the C ABI in include/fdtd_hip.h (libfdtd_hip.so).

The library is loaded from the package directory (built in-tree by ``build()``); a
missing library raises instead of falling back to anything else.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libfdtd_hip.so")
ADDON_PATH = os.path.join(PKG_DIR, "js", "fdtd_hip.node")
CSRC = os.path.join(PKG_DIR, "csrc")

FP64, FP32 = 0, 1
BOUNDARY_REFERENCE, BOUNDARY_MATLAB = 0, 1
NTFF_AUTO, NTFF_DIRECT, NTFF_BLOCKED = 0, 1, 2

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # The FP64 path must round every operation like the reference (no FMA).
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


class FdtdHipDims(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
                ("numberOfTimeSteps", C.c_int), ("precision", C.c_int), ("boundary", C.c_int)]


class FdtdHipTuning(C.Structure):
    _fields_ = [("planes_per_workgroup", C.c_int), ("cells_per_thread", C.c_int),
                ("coefficient_storage", C.c_int), ("fuse_cpml", C.c_int)]


class FdtdHipCounts(C.Structure):
    _fields_ = [("nSampledVoltages", C.c_int), ("nSampledCurrents", C.c_int), ("nProbes", C.c_int),
                ("nSources", C.c_int), ("nFarfieldFreq", C.c_int), ("nPortDftFreq", C.c_int)]


class FdtdHipKernelTimes(C.Structure):
    _fields_ = [("ms", C.c_double * 6), ("launches", C.c_longlong * 6)]


# Every symbol include/fdtd_hip.h declares.
EXPORTS = [
    "fdtd_hip_last_error", "fdtd_hip_device_count", "fdtd_hip_init", "fdtd_hip_upload_coeff",
    "fdtd_hip_upload_coeff_f32", "fdtd_hip_upload_cpml", "fdtd_hip_upload_vsource",
    "fdtd_hip_upload_svoltage", "fdtd_hip_upload_scurrent", "fdtd_hip_upload_field_probe",
    "fdtd_hip_read_probes", "fdtd_hip_upload_farfield",
    "fdtd_hip_upload_port_dft", "fdtd_hip_run_batch", "fdtd_hip_run_batch_async", "fdtd_hip_read_traces",
    "fdtd_hip_query", "fdtd_hip_get_step_info", "fdtd_hip_read_field", "fdtd_hip_write_field",
    "fdtd_hip_read_psi", "fdtd_hip_write_psi", "fdtd_hip_read_farfield", "fdtd_hip_read_port_dft", "fdtd_hip_radiated_power", "fdtd_hip_directivity", "fdtd_hip_directivity_at",
    "fdtd_hip_directivity_noalloc", "fdtd_hip_get_counts",
    "fdtd_hip_set_ntff_mode", "fdtd_hip_get_ntff_mode", "fdtd_hip_set_graphs",
    "fdtd_hip_set_tuning", "fdtd_hip_get_tuning", "fdtd_hip_autotune",
    "fdtd_hip_get_coefficient_storage", "fdtd_hip_set_timing", "fdtd_hip_get_kernel_times", "fdtd_hip_synchronize", "fdtd_hip_destroy",
    "fdtd_hip_multi_create", "fdtd_hip_multi_run_async", "fdtd_hip_multi_synchronize", "fdtd_hip_multi_set_graphs",
    "fdtd_hip_multi_get_tuning", "fdtd_hip_multi_autotune", "fdtd_hip_multi_destroy",
    "fdtd_hip_es_create", "fdtd_hip_es_predict", "fdtd_hip_es_destroy",
]


def build_library(verbose: bool = False) -> str:
    """Compile csrc/fdtd_hip.hip for gfx950 into the package directory."""
    src = os.path.join(CSRC, "fdtd_hip.hip")
    tmp = LIB_PATH + ".tmp"
    cmd = ["hipcc", *HIPCC_FLAGS, "-o", tmp, src]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed:\n{res.stderr[-4000:]}")
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f"built {LIB_PATH}")
    return LIB_PATH


_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        P, D, I, LL = C.c_void_p, C.POINTER(C.c_double), C.c_int, C.c_longlong
        IP, FP = C.POINTER(C.c_int), C.POINTER(C.c_float)
        L.fdtd_hip_last_error.restype = C.c_char_p
        L.fdtd_hip_last_error.argtypes = []
        L.fdtd_hip_device_count.argtypes = [IP]
        L.fdtd_hip_init.argtypes = [C.POINTER(FdtdHipDims), C.POINTER(P)]
        L.fdtd_hip_upload_coeff.argtypes = [P, I, D, LL]
        L.fdtd_hip_upload_coeff_f32.argtypes = [P, I, FP, LL]
        L.fdtd_hip_upload_cpml.argtypes = [P, I, I, D, D, D, D, D, LL, D, LL, D, LL, D, LL, I, I]
        L.fdtd_hip_upload_vsource.argtypes = [P, I, IP, D, I, D, I]
        L.fdtd_hip_upload_svoltage.argtypes = [P, I, IP, I, C.c_double]
        L.fdtd_hip_upload_scurrent.argtypes = [P, IP, IP, D, I, I]
        L.fdtd_hip_upload_field_probe.argtypes = [P, I, I, I, I, I]
        L.fdtd_hip_read_probes.argtypes = [P, I, I, D]
        L.fdtd_hip_upload_farfield.argtypes = [P, I, IP, D, I]
        L.fdtd_hip_upload_port_dft.argtypes = [P, I, D]
        L.fdtd_hip_run_batch.argtypes = [P, I, I, D, D]
        L.fdtd_hip_run_batch_async.argtypes = [P, I, I]
        L.fdtd_hip_read_traces.argtypes = [P, I, I, D, D]
        L.fdtd_hip_query.argtypes = [P, IP]
        L.fdtd_hip_get_step_info.argtypes = [P, IP]
        L.fdtd_hip_read_field.argtypes = [P, I, D, LL]
        L.fdtd_hip_write_field.argtypes = [P, I, D, LL]
        L.fdtd_hip_read_psi.argtypes = [P, I, I, D, LL]
        L.fdtd_hip_write_psi.argtypes = [P, I, I, D, LL]
        L.fdtd_hip_read_farfield.argtypes = [P, I, D, D, LL]
        L.fdtd_hip_read_port_dft.argtypes = [P, D, D, LL]
        L.fdtd_hip_radiated_power.argtypes = [P, D, I]
        L.fdtd_hip_directivity.argtypes = [P, D, D, I, D, D, D, D]
        L.fdtd_hip_directivity_noalloc.argtypes = [P, D, D, I, D, D, D, D]
        L.fdtd_hip_directivity_at.argtypes = [P, I, C.c_double, D, I, D, D, D, D]
        L.fdtd_hip_get_counts.argtypes = [P, C.POINTER(FdtdHipCounts)]
        L.fdtd_hip_set_ntff_mode.argtypes = [P, I]
        L.fdtd_hip_get_ntff_mode.argtypes = [P, IP]
        L.fdtd_hip_set_graphs.argtypes = [P, I]
        L.fdtd_hip_set_tuning.argtypes = [P, C.POINTER(FdtdHipTuning)]
        L.fdtd_hip_get_tuning.argtypes = [P, C.POINTER(FdtdHipTuning)]
        L.fdtd_hip_autotune.argtypes = [P, I, C.POINTER(FdtdHipTuning)]
        L.fdtd_hip_get_coefficient_storage.argtypes = [P, IP, IP]
        L.fdtd_hip_set_timing.argtypes = [P, I]
        L.fdtd_hip_get_kernel_times.argtypes = [P, C.POINTER(FdtdHipKernelTimes)]
        L.fdtd_hip_synchronize.argtypes = [P]
        L.fdtd_hip_destroy.argtypes = [P]
        L.fdtd_hip_destroy.restype = None
        L.fdtd_hip_multi_create.argtypes = [C.POINTER(P), I, C.POINTER(FdtdHipTuning), C.POINTER(P)]
        L.fdtd_hip_multi_run_async.argtypes = [P, I, I]
        L.fdtd_hip_multi_synchronize.argtypes = [P]
        L.fdtd_hip_multi_set_graphs.argtypes = [P, I]
        L.fdtd_hip_multi_get_tuning.argtypes = [P, C.POINTER(FdtdHipTuning)]
        L.fdtd_hip_multi_autotune.argtypes = [P, I, C.POINTER(FdtdHipTuning)]
        L.fdtd_hip_multi_destroy.argtypes = [P]
        L.fdtd_hip_multi_destroy.restype = None
        L.fdtd_hip_es_create.argtypes = [C.POINTER(C.c_uint16), LL, FP, LL, C.POINTER(P)]
        L.fdtd_hip_es_predict.argtypes = [P, I, FP, FP]
        L.fdtd_hip_es_destroy.argtypes = [P]
        L.fdtd_hip_es_destroy.restype = None
        _lib = L
    return _lib


class FdtdHipError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().fdtd_hip_last_error().decode("utf-8", "replace")
        raise FdtdHipError(f"{what}: {msg}")


def device_count() -> int:
    n = C.c_int(0)
    check(lib().fdtd_hip_device_count(C.byref(n)), "device_count")
    return n.value

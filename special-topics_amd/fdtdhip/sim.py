"""Python host of the HIP time loop: uploads a problem file through the C ABI and runs it.

``HipSim`` mirrors what the Node backend (js/backends/hipBackend.mjs) does with a
js_fdtd problem object; the test-suite and bench.py use it to drive the library from
the same problem files the JavaScript builder writes.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import native
from .problem import COEFF_ORDER, FACE_NAMES, FIELD_ORDER, Problem


def _d(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _i(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


SLOT_COMP = {"x": ["Hy", "Hz", "Ey", "Ez"], "y": ["Hx", "Hz", "Ez", "Ex"], "z": ["Hx", "Hy", "Ex", "Ey"]}


class HipSim:
    """One problem resident on the GPU."""

    def __init__(self, problem: Problem, precision: str = "fp64", boundary: str = "reference",
                 port_dft_freqs: np.ndarray | None = None, farfield: bool = True, port_dft_first: bool = False):
        self.L = native.lib()
        self.problem = problem
        g = problem.grid
        self.fp32 = precision == "fp32"
        dims = native.FdtdHipDims(
            nx=g["nx"], ny=g["ny"], nz=g["nz"], dx=g["dx"], dy=g["dy"], dz=g["dz"], dt=g["dt"],
            numberOfTimeSteps=problem.steps, precision=native.FP32 if self.fp32 else native.FP64,
            boundary=native.BOUNDARY_REFERENCE if boundary == "reference" else native.BOUNDARY_MATLAB)
        h = C.c_void_p()
        native.check(self.L.fdtd_hip_init(C.byref(dims), C.byref(h)), "init")
        self.h = h
        try:
            self._upload(problem, port_dft_freqs, farfield, port_dft_first)
        except Exception:
            self.close()
            raise

    def _upload(self, P: Problem, port_dft_freqs, farfield: bool, port_dft_first: bool = False):
        L, h = self.L, self.h
        self.ndft = 0
        if port_dft_first and port_dft_freqs is not None and len(port_dft_freqs) > 0:
            fr = np.ascontiguousarray(port_dft_freqs, dtype=np.float64)
            native.check(L.fdtd_hip_upload_port_dft(h, fr.size, _d(fr)), "upload_port_dft")
            self.ndft = fr.size
        for w, name in enumerate(COEFF_ORDER):
            a = P.arrays["coeff/" + name]
            if a.dtype == np.float32:
                if not self.fp32:
                    raise ValueError("FP32 problem file cannot drive an FP64 run")
                a = np.ascontiguousarray(a)
                native.check(L.fdtd_hip_upload_coeff_f32(h, w, _f(a), a.size), f"upload_coeff {name}")
            else:
                a = np.ascontiguousarray(a, dtype=np.float64)
                native.check(L.fdtd_hip_upload_coeff(h, w, _d(a), a.size), f"upload_coeff {name}")
        for f in P.header["cpml"]:
            face = f["face"]
            prof = [np.ascontiguousarray(P.arrays[f"cpml/{face}/{p}"], dtype=np.float64)
                    for p in ("b_e", "a_e", "b_m", "a_m")]
            cps = [np.ascontiguousarray(P.arrays[f"cpml/{face}/CPsi{s}"], dtype=np.float64)
                   for s in range(4)]
            args = []
            for a in cps:
                args += [_d(a), a.size]
            native.check(L.fdtd_hip_upload_cpml(h, FACE_NAMES.index(face), f["nc"], *[_d(a) for a in prof],
                                                *args, f["m_start"], f["e_start"]), f"upload_cpml {face}")
        for n, vs in enumerate(P.header["vsources"]):
            idx = np.ascontiguousarray(P.arrays[f"vsrc/{n}/idx"], dtype=np.int32)
            coef = np.ascontiguousarray(P.arrays[f"vsrc/{n}/coef"], dtype=np.float64)
            vpe = np.ascontiguousarray(P.arrays[f"vsrc/{n}/vpe"], dtype=np.float64)
            native.check(L.fdtd_hip_upload_vsource(h, vs["dir"], _i(idx), _d(coef), idx.size, _d(vpe), vpe.size),
                         "upload_vsource")
        for n, sv in enumerate(P.header["svoltages"]):
            idx = np.ascontiguousarray(P.arrays[f"sv/{n}/idx"], dtype=np.int32)
            native.check(L.fdtd_hip_upload_svoltage(h, sv["dir"], _i(idx), idx.size, float(sv["Csvf"])),
                         "upload_svoltage")
        for n, si in enumerate(P.header["scurrents"]):
            comp = np.ascontiguousarray(P.arrays[f"si/{n}/comp"], dtype=np.int32)
            index = np.ascontiguousarray(P.arrays[f"si/{n}/index"], dtype=np.int32)
            w = np.ascontiguousarray(P.arrays[f"si/{n}/weight"], dtype=np.float64)
            native.check(L.fdtd_hip_upload_scurrent(h, _i(comp), _i(index), _d(w), comp.size, int(si["negate"])),
                         "upload_scurrent")
        self.probe_list = P.header.get("probes", [])
        for pr in self.probe_list:
            native.check(L.fdtd_hip_upload_field_probe(h, pr["magnetic"], pr["component"], pr["i"], pr["j"], pr["k"]),
                         "upload_field_probe")
        ff = P.header["ff"]
        self.nfreq = 0
        if farfield and ff.get("nFreq", 0) > 0 and "ff/phase" in P.arrays:
            box = np.array([ff["li"], ff["lj"], ff["lk"], ff["ui"], ff["uj"], ff["uk"]], dtype=np.int32)
            ph = np.ascontiguousarray(P.arrays["ff/phase"], dtype=np.float64)
            native.check(L.fdtd_hip_upload_farfield(h, ff["nFreq"], _i(box), _d(ph), P.steps), "upload_farfield")
            self.nfreq = ff["nFreq"]
            self.ff_sizes = ([ff["nj_ff"] * ff["nk_ff"]] * 8 + [ff["ni_ff"] * ff["nk_ff"]] * 8 +
                             [ff["ni_ff"] * ff["nj_ff"]] * 8)
        self.n_sv = len(P.header["svoltages"])
        self.n_si = len(P.header["scurrents"])
        if not port_dft_first and port_dft_freqs is not None and len(port_dft_freqs) > 0:
            fr = np.ascontiguousarray(port_dft_freqs, dtype=np.float64)
            native.check(L.fdtd_hip_upload_port_dft(h, fr.size, _d(fr)), "upload_port_dft")
            self.ndft = fr.size

    # ------------------------------------------------------------------------------
    def close(self):
        """Frees the context.  A batch borrows its contexts, so a context in a live
        HipMulti cannot be closed: close the batch first."""
        if getattr(self, "_batches", 0):
            raise RuntimeError("HipSim.close: the context is part of a live HipMulti; close the batch first")
        if getattr(self, "h", None):
            self.L.fdtd_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        # A live batch holds a reference to this object, so this runs in a batch only at
        # interpreter exit; the context is then left to the process teardown.
        if not getattr(self, "_batches", 0):
            self.close()

    def run(self, ts0: int, count: int):
        """Steps [ts0, ts0+count); returns (V[n_sv, count], I[n_si, count])."""
        v = np.zeros((max(self.n_sv, 1), count))
        i = np.zeros((max(self.n_si, 1), count))
        native.check(self.L.fdtd_hip_run_batch(self.h, ts0, count, _d(v), _d(i)), "run_batch")
        return v[: self.n_sv], i[: self.n_si]

    def run_async(self, ts0: int, count: int):
        """Enqueue steps [ts0, ts0+count) and return at once (see traces / idle)."""
        native.check(self.L.fdtd_hip_run_batch_async(self.h, ts0, count), "run_batch_async")

    def traces(self, ts0: int, count: int):
        """Wait for the context's work, then (V[n_sv, count], I[n_si, count]) of [ts0, ts0+count)."""
        v = np.zeros((max(self.n_sv, 1), count))
        i = np.zeros((max(self.n_si, 1), count))
        native.check(self.L.fdtd_hip_read_traces(self.h, ts0, count, _d(v), _d(i)), "read_traces")
        return v[: self.n_sv], i[: self.n_si]

    def probes(self, ts0: int, count: int) -> np.ndarray:
        """Point field probe traces of steps [ts0, ts0+count): [n_probes, count]."""
        n = len(self.probe_list)
        out = np.zeros((max(n, 1), count))
        native.check(self.L.fdtd_hip_read_probes(self.h, ts0, count, _d(out)), "read_probes")
        return out[:n]

    def idle(self) -> bool:
        b = C.c_int(0)
        native.check(self.L.fdtd_hip_query(self.h, C.byref(b)), "query")
        return bool(b.value)

    def run_no_readback(self, ts0: int, count: int):
        native.check(self.L.fdtd_hip_run_batch(self.h, ts0, count, None, None), "run_batch")

    def synchronize(self):
        native.check(self.L.fdtd_hip_synchronize(self.h), "synchronize")

    def field(self, comp: str) -> np.ndarray:
        out = np.zeros(self.problem.component_len(comp))
        native.check(self.L.fdtd_hip_read_field(self.h, FIELD_ORDER.index(comp), _d(out), out.size), "read_field")
        return out

    def set_field(self, comp: str, values: np.ndarray):
        a = np.ascontiguousarray(values, dtype=np.float64)
        native.check(self.L.fdtd_hip_write_field(self.h, FIELD_ORDER.index(comp), _d(a), a.size), "write_field")

    def psi_len(self, face: str, slot: int) -> int:
        f = next(f for f in self.problem.header["cpml"] if f["face"] == face)
        d = list(self.problem.component_dims(SLOT_COMP[face[0]][slot]))
        d["xyz".index(face[0])] = f["nc"]
        return d[0] * d[1] * d[2]

    def psi(self, face: str, slot: int) -> np.ndarray:
        out = np.zeros(self.psi_len(face, slot))
        native.check(self.L.fdtd_hip_read_psi(self.h, FACE_NAMES.index(face), slot, _d(out), out.size), "read_psi")
        return out

    def set_psi(self, face: str, slot: int, values: np.ndarray):
        a = np.ascontiguousarray(values, dtype=np.float64)
        native.check(self.L.fdtd_hip_write_psi(self.h, FACE_NAMES.index(face), slot, _d(a), a.size), "write_psi")

    def farfield(self, which: int):
        n = self.ff_sizes[which] * self.nfreq
        re, im = np.zeros(n), np.zeros(n)
        native.check(self.L.fdtd_hip_read_farfield(self.h, which, _d(re), _d(im), n), "read_farfield")
        return re, im

    def radiated_power(self) -> np.ndarray:
        out = np.zeros(self.nfreq)
        native.check(self.L.fdtd_hip_radiated_power(self.h, _d(out), self.nfreq), "radiated_power")
        return out

    def directivity(self, theta, phi, rad_power=None):
        """(dataTheta, dataPhi), each [nFreq][nAngles], at the NTFF frequencies."""
        th = np.ascontiguousarray(theta, dtype=np.float64)
        ph = np.ascontiguousarray(phi, dtype=np.float64)
        fr = np.ascontiguousarray(self.problem.header["ff"]["frequencies"], dtype=np.float64)
        dt = np.zeros((self.nfreq, th.size))
        dp = np.zeros((self.nfreq, th.size))
        rp = None if rad_power is None else _d(np.ascontiguousarray(rad_power, dtype=np.float64))
        native.check(self.L.fdtd_hip_directivity(self.h, _d(fr), rp, th.size, _d(th), _d(ph), _d(dt), _d(dp)),
                     "directivity")
        return dt, dp

    def directivity_at(self, f, theta, phi, rad_power=None):
        """(dataTheta, dataPhi), each [nAngles], at NTFF frequency f only (row f of
        directivity()).  rad_power: that frequency's radiated power, or None."""
        th = np.ascontiguousarray(theta, dtype=np.float64)
        ph = np.ascontiguousarray(phi, dtype=np.float64)
        freqs = self.problem.header["ff"]["frequencies"]
        fr = float(freqs[f]) if 0 <= f < len(freqs) else 0.0      # the library refuses f outside
        dt = np.zeros(th.size)
        dp = np.zeros(th.size)
        rp = None if rad_power is None else _d(np.array([rad_power], dtype=np.float64))
        native.check(self.L.fdtd_hip_directivity_at(self.h, int(f), fr, rp, th.size, _d(th), _d(ph), _d(dt), _d(dp)),
                     "directivity_at")
        return dt, dp

    def directivity_noalloc(self, theta, phi, rad_power=None):
        """As directivity(), through the entry point that allocates no device memory."""
        th = np.ascontiguousarray(theta, dtype=np.float64)
        ph = np.ascontiguousarray(phi, dtype=np.float64)
        fr = np.ascontiguousarray(self.problem.header["ff"]["frequencies"], dtype=np.float64)
        dt = np.zeros((self.nfreq, th.size))
        dp = np.zeros((self.nfreq, th.size))
        rp = None if rad_power is None else _d(np.ascontiguousarray(rad_power, dtype=np.float64))
        native.check(self.L.fdtd_hip_directivity_noalloc(self.h, _d(fr), rp, th.size, _d(th), _d(ph), _d(dt),
                                                         _d(dp)), "directivity_noalloc")
        return dt, dp

    def port_dft(self):
        n = (self.n_sv + self.n_si) * self.ndft
        re, im = np.zeros(n), np.zeros(n)
        native.check(self.L.fdtd_hip_read_port_dft(self.h, _d(re), _d(im), n), "read_port_dft")
        return (re + 1j * im).reshape(self.n_sv + self.n_si, self.ndft)

    def set_tuning(self, planes_per_workgroup: int = 0, cells_per_thread: int = 0, coefficient_storage: int = 0,
                   fuse_cpml: int = 0):
        t = native.FdtdHipTuning(int(planes_per_workgroup), int(cells_per_thread), int(coefficient_storage),
                                 int(fuse_cpml))
        native.check(self.L.fdtd_hip_set_tuning(self.h, C.byref(t)), "set_tuning")

    @staticmethod
    def _tuple(t):
        return t.planes_per_workgroup, t.cells_per_thread, t.coefficient_storage, t.fuse_cpml

    def tuning(self):
        t = native.FdtdHipTuning()
        native.check(self.L.fdtd_hip_get_tuning(self.h, C.byref(t)), "get_tuning")
        return self._tuple(t)

    def autotune(self, steps_per_candidate: int = 3):
        t = native.FdtdHipTuning()
        native.check(self.L.fdtd_hip_autotune(self.h, int(steps_per_candidate), C.byref(t)), "autotune")
        return self._tuple(t)

    def coefficient_storage(self):
        mode = C.c_int(0)
        n = (C.c_int * 6)()
        native.check(self.L.fdtd_hip_get_coefficient_storage(self.h, C.byref(mode), n), "get_coefficient_storage")
        return mode.value, list(n)

    def set_ntff_mode(self, mode: str):
        """'auto' | 'direct' | 'blocked' (matrix-core K-blocked NTFF DFT)."""
        m = {"auto": native.NTFF_AUTO, "direct": native.NTFF_DIRECT, "blocked": native.NTFF_BLOCKED}[mode]
        native.check(self.L.fdtd_hip_set_ntff_mode(self.h, m), "set_ntff_mode")

    def ntff_blocked(self) -> bool:
        b = C.c_int(0)
        native.check(self.L.fdtd_hip_get_ntff_mode(self.h, C.byref(b)), "get_ntff_mode")
        return bool(b.value)

    def set_graphs(self, enable: bool):
        """Replay 64-step chunks from a captured HIP graph (default) or launch directly."""
        native.check(self.L.fdtd_hip_set_graphs(self.h, 1 if enable else 0), "set_graphs")

    def set_timing(self, enable: bool):
        native.check(self.L.fdtd_hip_set_timing(self.h, 1 if enable else 0), "set_timing")

    def kernel_times(self):
        t = native.FdtdHipKernelTimes()
        native.check(self.L.fdtd_hip_get_kernel_times(self.h, C.byref(t)), "get_kernel_times")
        return list(t.ms), list(t.launches)

    def e_tail_merged(self) -> bool:
        """Whether the last step ran E CPML, port epilogue and NTFF as one launch."""
        v = C.c_int(0)
        native.check(self.L.fdtd_hip_get_step_info(self.h, C.byref(v)), "get_step_info")
        return bool(v.value)


class HipMulti:
    """Several HipSim contexts stepped together, one launch per phase over all of them
    (fdtd_hip_multi_*).  The contexts stay usable for reads (traces, port_dft,
    farfield, field, ...) after run(); they must not be run on their own meanwhile.
    Each context's results equal its own run with the same tuning, bit for bit."""

    def __init__(self, sims, planes_per_workgroup: int = 0, cells_per_thread: int = 0,
                 coefficient_storage: int = 0, fuse_cpml: int = 0):
        self.sims = list(sims)
        self.L = native.lib()
        self.h = C.c_void_p()
        arr = (C.c_void_p * len(self.sims))(*[s.h for s in self.sims])
        t = native.FdtdHipTuning(int(planes_per_workgroup), int(cells_per_thread), int(coefficient_storage),
                                 int(fuse_cpml))
        native.check(self.L.fdtd_hip_multi_create(arr, len(self.sims), C.byref(t), C.byref(self.h)), "multi_create")
        for s in self.sims:
            s._batches = getattr(s, "_batches", 0) + 1

    def close(self):
        if getattr(self, "h", None):
            self.L.fdtd_hip_multi_destroy(self.h)
            self.h = None
            for s in self.sims:
                s._batches -= 1

    def __del__(self):
        self.close()

    def run_async(self, ts0: int, count: int):
        native.check(self.L.fdtd_hip_multi_run_async(self.h, ts0, count), "multi_run_async")

    def synchronize(self):
        native.check(self.L.fdtd_hip_multi_synchronize(self.h), "multi_synchronize")

    def run(self, ts0: int, count: int):
        """Steps [ts0, ts0+count) of every context; returns [(V, I)] per context."""
        self.run_async(ts0, count)
        return [s.traces(ts0, count) for s in self.sims]

    def set_graphs(self, enable: bool):
        native.check(self.L.fdtd_hip_multi_set_graphs(self.h, 1 if enable else 0), "multi_set_graphs")

    def tuning(self):
        t = native.FdtdHipTuning()
        native.check(self.L.fdtd_hip_multi_get_tuning(self.h, C.byref(t)), "multi_get_tuning")
        return HipSim._tuple(t)

    def autotune(self, steps_per_candidate: int = 3):
        t = native.FdtdHipTuning()
        native.check(self.L.fdtd_hip_multi_autotune(self.h, int(steps_per_candidate), C.byref(t)), "multi_autotune")
        return HipSim._tuple(t)


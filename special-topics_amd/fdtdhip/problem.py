"""Reader for the binary problem files written by ``js/problem/dump.mjs``.

A problem file is the FP64 (or FP32-coefficient) set-up of one scenario exactly as
the JavaScript builder produced it: a JSON header followed by 64-byte aligned
little-endian arrays.  ``build_problem`` runs the Node builder to make one.
"""

from __future__ import annotations

import json
import os
import shutil
import subprocess
from dataclasses import dataclass, field

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JS_DIR = os.path.join(PKG_DIR, "js")
SCENARIO_DIR = os.path.join(PKG_DIR, "scenarios")

_DTYPES = {"f64": np.float64, "f32": np.float32, "i32": np.int32, "u8": np.uint8}

COEFF_ORDER = [
    "Cexe", "Cexhz", "Cexhy", "Ceye", "Ceyhx", "Ceyhz", "Ceze", "Cezhy", "Cezhx",
    "Chxh", "Chxez", "Chxey", "Chyh", "Chyex", "Chyez", "Chzh", "Chzey", "Chzex",
]
FIELD_ORDER = ["Hx", "Hy", "Hz", "Ex", "Ey", "Ez"]
FACE_NAMES = ["xn", "xp", "yn", "yp", "zn", "zp"]


@dataclass
class Problem:
    header: dict
    arrays: dict = field(default_factory=dict)
    path: str = ""

    @property
    def grid(self) -> dict:
        return self.header["grid"]

    @property
    def dims(self):
        g = self.header["grid"]
        return g["nx"], g["ny"], g["nz"]

    @property
    def steps(self) -> int:
        return int(self.header["grid"]["numberOfTimeSteps"])

    def coeff(self, name: str) -> np.ndarray:
        return self.arrays["coeff/" + name]

    def component_dims(self, comp: str):
        nx, ny, nz = self.dims
        ext = {"Hx": (1, 0, 0), "Hy": (0, 1, 0), "Hz": (0, 0, 1),
               "Ex": (0, 1, 1), "Ey": (1, 0, 1), "Ez": (1, 1, 0)}[comp]
        return nx + ext[0], ny + ext[1], nz + ext[2]

    def component_len(self, comp: str) -> int:
        d = self.component_dims(comp)
        return d[0] * d[1] * d[2]


def read_problem(path: str, mmap: bool = True) -> Problem:
    with open(path, "rb") as f:
        pre = f.read(16)
        if pre[:8] != b"FDTDPRB1":
            raise ValueError(f"{path}: not a problem file")
        hlen = int.from_bytes(pre[8:12], "little")
        header = json.loads(f.read(hlen).decode("utf-8"))
    raw = np.memmap(path, dtype=np.uint8, mode="r") if mmap else np.fromfile(path, dtype=np.uint8)
    arrays = {}
    for name, (dtype, off, count) in header["arrays"].items():
        dt = np.dtype(_DTYPES[dtype])
        arrays[name] = raw[off: off + count * dt.itemsize].view(dt)
    return Problem(header=header, arrays=arrays, path=path)


def node_binary() -> str:
    exe = shutil.which("node")
    if exe is None:
        raise RuntimeError("node is required to build problems (the set-up is the JavaScript host's)")
    return exe


def build_problem(scenario: str, out: str, steps: int | None = None, fp32: bool = False,
                  no_phase: bool = False, timeout: float = 900.0) -> dict:
    """Run the Node problem builder; returns its one-line JSON report."""
    cmd = [node_binary(), os.path.join(JS_DIR, "build_problem.mjs"),
           f"--scenario={scenario}", f"--out={out}"]
    if steps:
        cmd.append(f"--steps={int(steps)}")
    if fp32:
        cmd.append("--fp32")
    if no_phase:
        cmd.append("--no-phase")
    env = dict(os.environ)
    env.pop("FDTD_STEPS", None)
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env)
    if res.returncode != 0:
        raise RuntimeError(f"problem build failed ({res.returncode}): {res.stderr.strip()[-2000:]}")
    lines = [ln for ln in res.stdout.splitlines() if ln.startswith("{")]
    return json.loads(lines[-1])


def load_scenario_problem(scenario: str, workdir: str, steps: int | None = None,
                          fp32: bool = False, no_phase: bool = False) -> Problem:
    os.makedirs(workdir, exist_ok=True)
    tag = os.path.basename(scenario).replace(".json", "")
    out = os.path.join(workdir, f"{tag}-{steps or 'n'}-{'f32' if fp32 else 'f64'}.fdtdp")
    if not os.path.exists(out):
        build_problem(scenario, out + ".tmp", steps=steps, fp32=fp32, no_phase=no_phase)
        os.replace(out + ".tmp", out)
    return read_problem(out)

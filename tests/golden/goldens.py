"""Helpers to read the golden fixtures written by tests/golden/gen_golden.mjs."""

from __future__ import annotations

import base64
import hashlib
import json
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = os.path.join(HERE, "fixtures")
SCENARIOS = os.path.join(HERE, "scenarios")

FF_ORDER = ["cjyxp", "cjzxp", "cmyxp", "cmzxp", "cjyxn", "cjzxn", "cmyxn", "cmzxn",
            "cjxyp", "cjzyp", "cmxyp", "cmzyp", "cjxyn", "cjzyn", "cmxyn", "cmzyn",
            "cjxzp", "cjyzp", "cmxzp", "cmyzp", "cjxzn", "cjyzn", "cmxzn", "cmyzn"]
PSI_NAMES = {"x": ["hyx", "hzx", "eyx", "ezx"], "y": ["hxy", "hzy", "ezy", "exy"],
             "z": ["hxz", "hyz", "exz", "eyz"]}


def load(name: str) -> dict:
    with open(os.path.join(FIXTURES, name)) as f:
        return json.load(f)


def f64(b64: str) -> np.ndarray:
    return np.frombuffer(base64.b64decode(b64), dtype=np.float64)


def b64(a: np.ndarray) -> str:
    return base64.b64encode(np.ascontiguousarray(a, dtype="<f8").tobytes()).decode("ascii")


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def hexd(x: float) -> str:
    return struct.pack("<d", float(x)).hex()


def splitmix_uniform(seed: int):
    """Same generator as gen_golden.mjs: splitmix64, uniform doubles in [-1, 1)."""
    M = (1 << 64) - 1
    s = seed

    def nxt() -> float:
        nonlocal s
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        return (z >> 11) / 9007199254740992 * 2 - 1
    return nxt


def random_state(problem, seed: int):
    """Fields (Hx..Ez) and Psi arrays (faces xn..zp, slots 0..3) in the fill order of
    gen_golden.mjs stepsRecord()."""
    rnd = splitmix_uniform(seed)
    fields = {}
    for comp in ("Hx", "Hy", "Hz", "Ex", "Ey", "Ez"):
        fields[comp] = np.array([rnd() for _ in range(problem.component_len(comp))])
    psi = {}
    slot_comp = {"x": ["Hy", "Hz", "Ey", "Ez"], "y": ["Hx", "Hz", "Ez", "Ex"], "z": ["Hx", "Hy", "Ex", "Ey"]}
    for f in problem.header["cpml"]:
        face = f["face"]
        axis = "xyz".index(face[0])
        for slot in range(4):
            d = list(problem.component_dims(slot_comp[face[0]][slot]))
            d[axis] = f["nc"]
            psi[(face, slot)] = np.array([rnd() for _ in range(d[0] * d[1] * d[2])])
    return fields, psi

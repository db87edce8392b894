"""The C ABI: every function include/fdtd_hip.h declares is exported by the built
library, the ctypes binding knows exactly that set, and the N-API addon exposes the
methods its header comment documents.  No GPU work is launched."""

from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest

from fdtdhip import native, addon

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "fdtd_hip.h")
ADDON_SRC = os.path.join(REPO, "special-topics_amd", "csrc", "napi_addon.c")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(fdtd_hip_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(native.LIB_PATH):
        native.build_library()
    if not os.path.exists(native.ADDON_PATH):
        addon.build_addon()
    return native.LIB_PATH


def test_header_matches_binding():
    assert declared() == sorted(native.EXPORTS)


def test_library_exports_every_declared_symbol(built):
    nm = subprocess.run(["nm", "-D", "--defined-only", built], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (fdtd_hip_[a-z0-9_]+)$", nm, re.M))
    assert set(declared()) <= exported
    lib = ctypes.CDLL(built)
    for name in declared():
        assert hasattr(lib, name), name


def test_library_reports_errors_without_a_device(built):
    """Calls that need a context fail with a message instead of crashing."""
    L = native.lib()
    n = ctypes.c_int(-1)
    assert L.fdtd_hip_device_count(ctypes.byref(n)) in (0, 1)
    assert n.value >= 0
    assert L.fdtd_hip_run_batch(None, 0, 1, None, None) != 0
    assert L.fdtd_hip_last_error()


def test_addon_exposes_documented_methods(built):
    doc = open(ADDON_SRC).read().split("*/", 1)[0]
    documented = set(re.findall(r"\b([a-z][A-Za-z]+)\(", doc)) - {"fdtd_hip_last_error"}
    res = subprocess.run(["node", "-e", f"console.log(Object.keys(require({native.ADDON_PATH!r})).join(' '))"],
                         capture_output=True, text=True, check=True)
    keys = set(res.stdout.split())
    assert documented <= keys, documented - keys

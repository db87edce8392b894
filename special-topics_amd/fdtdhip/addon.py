"""Build of the Node-API addon (csrc/napi_addon.c -> js/fdtd_hip.node).

The addon is a plain C shared object compiled against node_api.h; it links the HIP
library with an $ORIGIN-relative rpath so the tree can move as a whole.
"""

from __future__ import annotations

import os
import subprocess

from .native import ADDON_PATH, CSRC, LIB_PATH, PKG_DIR

NODE_INCLUDE = "/usr/include/node"


def build_addon(verbose: bool = False) -> str:
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("build the HIP library first")
    src = os.path.join(CSRC, "napi_addon.c")
    tmp = ADDON_PATH + ".tmp"
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-Wall", "-Wextra", "-Wno-unused-parameter",
           "-DNODE_GYP_MODULE_NAME=fdtd_hip", f"-I{NODE_INCLUDE}", "-o", tmp, src,
           f"-L{PKG_DIR}", "-l:libfdtd_hip.so", "-Wl,-rpath,$ORIGIN/..", "-Wl,--allow-shlib-undefined"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"addon build failed:\n{res.stderr[-4000:]}")
    os.replace(tmp, ADDON_PATH)
    if verbose:
        print(f"built {ADDON_PATH}")
    return ADDON_PATH

"""Loader of the native extension (_pyrenderer over libdvr_hip.so).

Fails loudly when the HIP library has not been built: there is no fallback
implementation of the renderer.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "libdvr_hip.so")


def _load():
    if not os.path.exists(LIBRARY_PATH):
        raise ImportError(
            f"{LIBRARY_PATH} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()' or make -C diffdvr_amd/csrc)")
    # load the C-ABI library explicitly so the extension's $ORIGIN rpath is not the only route
    ctypes.CDLL(LIBRARY_PATH, mode=ctypes.RTLD_GLOBAL)
    from . import _pyrenderer  # noqa: E402  (built next to this file)
    return _pyrenderer


native = _load()

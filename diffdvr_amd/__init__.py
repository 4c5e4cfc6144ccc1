"""MI355X-native differentiable direct volume rendering (DiffDVR's ray-marching
forward pass and adjoint on gfx950).

    from diffdvr_amd import pyrenderer          # the pyrenderer API
    from diffdvr_amd.diffdvr import Renderer    # torch.nn.Module / autograd wrapper

The renderer runs on HIP device tensors only.
"""
from __future__ import annotations

import os

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(PACKAGE_DIR, "libdvr_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(PACKAGE_DIR), "include", "dvr.h")

__all__ = ["PACKAGE_DIR", "LIBRARY_PATH", "HEADER_PATH"]

"""umbreon_amd -- MI355X (gfx950) per-ray backend of libumbreon.

The product is the C-ABI library ``lib/libumbreon_gpu.so`` (hand-written HIP
kernels + host scene builder, declared in ``include/umbreon_gpu.h``) and the
C++ facade ``lib/libumbreon.so`` (``umbreon::render``). This Python package is
a thin host-side mirror of the same interface used by the tests and the
benchmark; it loads the native library and never computes an image itself.
"""
from .capi import UmbreonError  # noqa: F401
from .scene import (Camera, DistantLight, Fog, Material, RenderOptions,  # noqa: F401
                    Scene)
from .render import BuiltScene, Device, FrameResult, render  # noqa: F401

__all__ = ["UmbreonError", "Camera", "DistantLight", "Fog", "Material", "RenderOptions", "Scene",
           "BuiltScene", "Device", "FrameResult", "render"]

"""raymond_amd: MI355X (gfx950) wavefront path tracer behind a C ABI.

The native library (libraymond_amd.so) is required; importing the binding
raises if it has not been built (python raymond_amd/build.py).
"""
from . import api  # noqa: F401  (loads the native library)
from .api import Device, Renderer, Scene, RmError  # noqa: F401

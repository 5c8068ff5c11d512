"""MI355X-native render path of gorender: HIP kernels behind a C ABI.

The compute lives in libgorender_hip.so (gorender_amd/csrc); this package holds
its ctypes binding (_lib), a device context wrapper (device), the scene data
model with seeded generators (scenes), asset and scene-manifest loading
(assets) and the serving-loop wrapper (serve).
"""
from . import _lib  # noqa: F401

__all__ = ["_lib"]

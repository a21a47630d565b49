"""astrosimgpu-amd: the simulation step of astrosimgpu's tripartite
neuron-astrocyte network, native to AMD Instinct MI355X (gfx950).

The step runs in hand-written HIP kernels behind the C ABI of include/asg.h;
this package is the Python binding of that ABI (ctypes) plus the host-side
helpers the tests, the benchmark and the serving loop use.
"""

from ._capi import AsgError, LIB_PATH, lib  # noqa: F401
from .network import Batch, DeviceNetwork, Model, device_math  # noqa: F401

__all__ = ["AsgError", "Batch", "DeviceNetwork", "Model", "device_math", "lib", "LIB_PATH"]

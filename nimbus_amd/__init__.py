"""MI355X (gfx950) HIP compute backend for the Nimbus Klemp-Wilhelmson large step.

The compute path is the C-ABI library libnimbus_hip.so (include/nimbus_hip.h);
`nimbus_amd.lib` is its ctypes binding and `nimbus_amd.serving` the ensemble
request loop built on it.
"""
from . import lib  # noqa: F401

__all__ = ["lib"]

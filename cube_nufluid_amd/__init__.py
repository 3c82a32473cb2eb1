"""MI355X-native implementation of CUBE's per-step gravity and update path.

The compute path is libcubehip.so (HIP, gfx950) behind the C ABI declared in
include/cube_hip.h.  This package holds the host-side mirror of CUBE's
driver (driver.Cube), its parameter and checkpoint formats, and the seeded
initial-condition writer.
"""
from .params import BoxConfig, CONFIGS, read_kernel_tables  # noqa: F401
from .checkpoint import BoxState, read_checkpoint, write_checkpoint  # noqa: F401


def Cube(*args, **kwargs):
    """driver.Cube, imported lazily so that importing the package never needs the GPU."""
    from .driver import Cube as _Cube
    return _Cube(*args, **kwargs)

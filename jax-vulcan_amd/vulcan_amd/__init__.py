"""MI355X-native Ros2 column step for VULCAN (gfx950, HIP, FP64).

Host modules mirror the reference's step API: ``chem`` (chem.py),
``solver`` (solver.py), ``jax_step`` (jax_step.py).  Every compute call goes
through the C ABI of ``libvulcan_hip.so`` (include/vulcan_ros2.h).
"""
from ._types import AtmStatic, NetworkArrays  # noqa: F401

__version__ = "0.1.0"

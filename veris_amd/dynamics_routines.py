"""GPU drop-ins for ``veris/dynamics_routines.py`` (implemented in :mod:`veris_amd.kernels`)."""

from .kernels import SeaIceStrength  # noqa: F401

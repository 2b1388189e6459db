"""GPU drop-in for ``veris/dynsolver.py`` (implemented in :mod:`veris_amd.kernels`)."""

from .kernels import WindForcingXY, IceVelocities  # noqa: F401

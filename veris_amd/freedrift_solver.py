"""GPU drop-in for ``veris/freedrift_solver.py`` (implemented in :mod:`veris_amd.kernels`)."""

from .kernels import freedrift_solver  # noqa: F401

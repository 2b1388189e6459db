"""GPU drop-in for ``veris/evp_solver.py`` (implemented in :mod:`veris_amd.kernels`)."""

from .kernels import evp_solver  # noqa: F401

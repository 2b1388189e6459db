"""GPU drop-in for ``veris/advection.py`` (implemented in :mod:`veris_amd.kernels`)."""

from .kernels import Advection  # noqa: F401

"""GPU drop-in for ``veris/growth.py`` (implemented in :mod:`veris_amd.kernels`)."""

from .kernels import Growth  # noqa: F401

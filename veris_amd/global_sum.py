"""GPU drop-in for ``veris/global_sum.py`` (implemented in :mod:`veris_amd.kernels`)."""

from .kernels import global_sum  # noqa: F401

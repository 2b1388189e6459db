"""GPU drop-in for ``veris/clean_up.py`` (implemented in :mod:`veris_amd.kernels`)."""

from .kernels import clean_up_advection, ridging  # noqa: F401

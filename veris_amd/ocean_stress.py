"""GPU drop-in for ``veris/ocean_stress.py`` (implemented in :mod:`veris_amd.kernels`)."""

from .kernels import OceanStressUV  # noqa: F401

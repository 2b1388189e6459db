"""GPU drop-in for ``veris/area_mass.py`` (implemented in :mod:`veris_amd.kernels`)."""

from .kernels import SeaIceMass, AreaWS  # noqa: F401

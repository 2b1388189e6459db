"""GPU drop-in for ``veris/fill_overlap.py`` (implemented in :mod:`veris_amd.kernels`)."""

from .kernels import fill_overlap, fill_overlap_uv  # noqa: F401

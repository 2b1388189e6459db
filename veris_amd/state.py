"""Veros-free model state: ``state.settings`` and ``state.variables``.

The reference kernels are pure functions of a Veros ``state`` object whose
``variables`` and ``settings`` attributes expose arrays and scalars by name
(e.g. ``veris/evp_solver.py:263-264``).  :class:`VerisState` provides exactly
that surface without Veros, plus the three host-model settings the plugin reads
from Veros (``nx``, ``ny``, ``enable_cyclic_x``).
"""

from __future__ import annotations

import numpy as np

from .settings import SettingsNamespace, make_settings
from .variables import VariablesNamespace


class VerisState:
    def __init__(self, settings: SettingsNamespace, variables: VariablesNamespace | dict):
        self.settings = settings
        self.variables = (
            variables if isinstance(variables, VariablesNamespace) else VariablesNamespace(variables)
        )

    @property
    def shape(self):
        s = self.settings
        return (s.nx + 2 * s.olx, s.ny + 2 * s.oly)

    def to(self, device="cuda", dtype=None):
        """State with every variable as a torch tensor on ``device``.

        ``dtype`` (a torch dtype) converts floating-point variables; default
        keeps float64.
        """
        import torch

        dt = torch.float64 if dtype is None else dtype

        def conv(a):
            t = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a))
            return t.to(device=device, dtype=dt).contiguous()

        return self.copy(conv)

    def numpy(self):
        """State with every variable as a float64 NumPy array on the host."""

        def conv(a):
            if hasattr(a, "detach"):
                return a.detach().to("cpu").double().numpy()
            return np.asarray(a, dtype=np.float64)

        return self.copy(conv)

    def copy(self, convert=None):
        """Shallow state copy; ``convert(array)`` maps every variable if given."""
        conv = convert or (lambda a: a)
        vals = {k: conv(v) for k, v in self.variables.items()}
        return VerisState(SettingsNamespace(self.settings.as_dict()), VariablesNamespace(vals))


def make_state_settings(nx: int, ny: int, enable_cyclic_x: bool = True, **overrides):
    """Plugin settings plus the host-model keys ``nx``, ``ny``, ``enable_cyclic_x``."""
    base = make_settings(**overrides)
    vals = base.as_dict()
    vals.update(nx=int(nx), ny=int(ny), enable_cyclic_x=bool(enable_cyclic_x))
    return SettingsNamespace(vals)

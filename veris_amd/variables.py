"""Two-dimensional variables of the sea-ice plugin.

Names, units and the ``("xt", "yt")`` grid are kept identical to
``veris/variables.py:5-106`` (83 entries): every array is shaped
``(nx + 2*olx, ny + 2*oly)`` in C order, axis 0 = x (zonal, index ``i``),
axis 1 = y (meridional, index ``j``, contiguous).
"""

from __future__ import annotations

from dataclasses import dataclass

GRID = ("xt", "yt")


@dataclass(frozen=True)
class Variable:
    description: str
    dims: tuple
    units: str


# (name, units, help) in the reference's order, grouped as there.
_SEA_ICE = [
    ("hIceMean", "m", "mean ice thickness"),
    ("hSnowMean", "m", "mean snow thickness"),
    ("Area", " ", "ice cover fraction"),
    ("TSurf", "K", "ice/snow surface temperature"),
    ("SeaIceMassC", "kg", "ice+snow mass at c-points"),
    ("SeaIceMassU", "kg", "ice+snow mass at u-points"),
    ("SeaIceMassV", "kg", "ice+snow mass at v-points"),
    ("SeaIceStrength", "N/m", "ice strength P"),
    ("os_hIceMean", "m", "advective ice-thickness overshoot"),
    ("os_hSnowMean", "m/s", "advective snow-thickness overshoot"),
    ("AreaW", " ", "ice cover fraction at u-points"),
    ("AreaS", " ", "ice cover fraction at v-points"),
    ("uIce", "m/s", "zonal ice velocity"),
    ("vIce", "m/s", "meridional ice velocity"),
    ("sigma1", "N/m2", "stress eigenvalue sigma1"),
    ("sigma2", "N/m2", "stress eigenvalue sigma2"),
    ("sigma12", "N/m2", "stress component sigma12"),
    ("WindForcingX", "N", "zonal wind and tilt forcing"),
    ("WindForcingY", "N", "meridional wind and tilt forcing"),
    ("recip_hIceMean", "1/m", "regularised 1/hIceMean"),
    ("SeaIceLoad", "kg/m2", "ice+snow load on the ocean surface"),
    ("IcePenetSW", "W/m2", "short-wave flux through the ice"),
]
_OCEAN = [
    ("uOcean", "m/s", "zonal ocean surface velocity"),
    ("vOcean", "m/s", "meridional ocean surface velocity"),
    ("theta", "K", "ocean surface temperature"),
    ("ocSalt", "g/kg", "ocean surface salinity"),
    ("Qnet", "W/m2", "net heat flux out of the ocean"),
    ("OceanStressU", "N/m2", "zonal stress on the ocean surface"),
    ("OceanStressV", "N/m2", "meridional stress on the ocean surface"),
    ("saltflux", "m/s", "salt flux into the ocean"),
    ("R_low", "m", "sea floor depth"),
    ("ssh_an", "m", "sea surface height anomaly"),
]
_ATMOSPHERE = [
    ("Qsw", "W/m2", "surface short-wave flux (+ up)"),
    ("uWind", "m/s", "zonal wind"),
    ("vWind", "m/s", "meridional wind"),
    ("wSpeed", "m/s", "wind speed"),
    ("surfPress", "P", "surface pressure"),
    ("SWdown", "W/m2", "downward short-wave radiation"),
    ("LWdown", "W/m2", "downward long-wave radiation"),
    ("ATemp", "K", "air temperature"),
    ("aqh", "kg/kg", "specific humidity"),
    ("precip", "m/s", "precipitation rate"),
    ("snowfall", "m/s", "snowfall rate"),
    ("evap", "m/s", "evaporation rate over open ocean"),
    ("runoff", "m/s", "runoff into the ocean"),
    ("EmPmR", "kg/m2 s", "evaporation minus precipitation minus runoff"),
]
_MASKS = [
    ("maskInC", " ", "open-boundary mask at c-points"),
    ("maskInU", " ", "open-boundary mask at u-points"),
    ("maskInV", " ", "open-boundary mask at v-points"),
    ("iceMask", " ", "mask at c-points"),
    ("iceMaskU", " ", "mask at u-points"),
    ("iceMaskV", " ", "mask at v-points"),
    ("k1AtC", " ", "metric term k1 at c-points"),
    ("k2AtC", " ", "metric term k2 at c-points"),
    ("k1AtZ", " ", "metric term k1 at z-points"),
    ("k2AtZ", " ", "metric term k2 at z-points"),
    ("Fu", " ", "coastline form factor, u"),
    ("Fv", " ", "coastline form factor, v"),
]
_GRIDVARS = [
    ("fCori", "1/s", "Coriolis parameter"),
    ("dxC", "m", "zonal centre spacing across the west face"),
    ("dyC", "m", "meridional centre spacing across the south face"),
    ("dxG", "m", "zonal face spacing along the south face"),
    ("dyG", "m", "meridional face spacing along the west face"),
    ("dxU", "m", "zonal u-point spacing through the centre"),
    ("dyU", "m", "meridional u-point spacing through the SW corner"),
    ("dxV", "m", "zonal v-point spacing through the SW corner"),
    ("dyV", "m", "meridional v-point spacing through the centre"),
    ("recip_dxC", "1/m", "1/dxC"),
    ("recip_dyC", "1/m", "1/dyC"),
    ("recip_dxG", "1/m", "1/dxG"),
    ("recip_dyG", "1/m", "1/dyG"),
    ("recip_dxU", "1/m", "1/dxU"),
    ("recip_dyU", "1/m", "1/dyU"),
    ("recip_dxV", "1/m", "1/dxV"),
    ("recip_dyV", "1/m", "1/dyV"),
    ("rA", "m2", "cell area at c-points"),
    ("rAu", "m2", "cell area at u-points"),
    ("rAv", "m2", "cell area at v-points"),
    ("rAz", "m2", "cell area at z-points"),
    ("recip_rA", "1/m2", "1/rA"),
    ("recip_rAu", "1/m2", "1/rAu"),
    ("recip_rAv", "1/m2", "1/rAv"),
    ("recip_rAz", "1/m2", "1/rAz"),
]

VARIABLES = {
    name: Variable(desc, GRID, units)
    for group in (_SEA_ICE, _OCEAN, _ATMOSPHERE, _MASKS, _GRIDVARS)
    for name, units, desc in group
}

VARIABLE_NAMES = tuple(VARIABLES)

# variables that never change within a model step (grid, masks)
GRID_NAMES = tuple(n for n, _, _ in _MASKS + _GRIDVARS)


class VariablesNamespace:
    """Attribute namespace of arrays (``state.variables.<name>``).

    ``_version`` counts assignments, so a caller can tell whether any
    variable has been replaced since it last looked (in-place changes of an
    array keep its identity and do not count)."""

    def __init__(self, arrays: dict | None = None):
        object.__setattr__(self, "_arrays", dict(arrays or {}))
        object.__setattr__(self, "_version", 0)

    def __getattr__(self, name):
        try:
            return self._arrays[name]
        except KeyError:
            raise AttributeError(name) from None

    def __setattr__(self, name, value):
        self._arrays[name] = value
        object.__setattr__(self, "_version", self._version + 1)

    def __getitem__(self, name):
        return self._arrays[name]

    def __setitem__(self, name, value):
        self._arrays[name] = value
        object.__setattr__(self, "_version", self._version + 1)

    def __contains__(self, name):
        return name in self._arrays

    def keys(self):
        return self._arrays.keys()

    def items(self):
        return self._arrays.items()

    def get(self, name, default=None):
        return self._arrays.get(name, default)

    def as_dict(self):
        return dict(self._arrays)

"""Scalar settings of the sea-ice plugin.

The names, defaults and types are the plugin's public configuration surface and
are kept identical to ``veris/settings.py:4-176`` (134 entries) so that a Veros
setup that sets ``state.settings.<name>`` keeps working.  The help strings are
short summaries.  ``recip_*`` entries are independent settings exactly as in the
reference (they are *not* derived from their partner); :func:`make_settings`
keeps them consistent when the caller overrides the partner.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Any


@dataclass(frozen=True)
class Setting:
    """One scalar setting: default value, python type and a one-line help."""

    default: Any
    type: type
    description: str


def _S(default, typ, desc):
    return Setting(default, typ, desc)


# (name, default, type, help) in the reference's order.
_TABLE = [
    ("deltatTherm", 86400, float, "thermodynamic time step [s]"),
    ("recip_deltatTherm", 1 / 86400, float, "1/deltatTherm [1/s]"),
    ("deltatDyn", 86400, float, "dynamic time step [s]"),
    ("recip_deltatDyn", 1 / 86400, float, "1/deltatDyn [1/s]"),
    ("olx", 2, int, "zonal overlap width [cells]"),
    ("oly", 2, int, "meridional overlap width [cells]"),
    ("gridcellWidth", 444709.893408, float, "grid cell width [m]"),
    ("nITC", 5, int, "number of ice thickness categories"),
    ("recip_nITC", 1 / 5, float, "1/nITC"),
    ("noSlip", True, bool, "no-slip lateral boundary condition"),
    ("useRelativeWind", True, bool, "wind stress from wind minus ice velocity"),
    ("secondOrderBC", False, bool, "second-order no-slip boundary terms"),
    ("extensiveFld", True, bool, "advected fields are extensive"),
    ("useRealFreshWaterFlux", False, bool, "sea-ice load enters the surface tilt"),
    ("useFreedrift", False, bool, "use the free-drift solver"),
    ("useEVP", True, bool, "use the EVP solver"),
    ("evpAlpha", 500, float, "EVP stress relaxation parameter"),
    ("evpBeta", 500, float, "EVP momentum relaxation parameter"),
    ("useAdaptiveEVP", False, bool, "adaptive relaxation parameters"),
    ("aEVPalphaMin", 5, float, "lower limit of adaptive alpha/beta"),
    ("aEvpCoeff", 0.5, float, "largest stabilised frequency (adaptive EVP)"),
    ("explicitDrag", True, bool, "explicit momentum drag"),
    ("nEVPsteps", 400, int, "EVP sub-cycling iterations per step"),
    ("computeEvpResidual", False, bool, "compute EVP residuals"),
    ("veros_fill", True, bool, "halo fill through Veros enforce_boundaries"),
    ("use_coastline", False, bool, "coastline form factor for lateral drag"),
    ("rhoIce", 900, float, "ice density [kg/m3]"),
    ("rhoFresh", 1000, float, "fresh water density [kg/m3]"),
    ("rhoSea", 1026, float, "sea water density [kg/m3]"),
    ("rhoAir", 1.3, float, "air density [kg/m3]"),
    ("rhoSnow", 330, float, "snow density [kg/m3]"),
    ("recip_rhoFresh", 1 / 1000, float, "1/rhoFresh"),
    ("recip_rhoSea", 1 / 1026, float, "1/rhoSea"),
    ("rhoIce2rhoSnow", 900 / 330, float, "rhoIce/rhoSnow"),
    ("rhoIce2rhoFresh", 900 / 1000, float, "rhoIce/rhoFresh"),
    ("rhoFresh2rhoSnow", 1000 / 330, float, "rhoFresh/rhoSnow"),
    ("dryIceAlb", 0.75, float, "dry ice albedo"),
    ("dryIceAlb_south", 0.75, float, "dry ice albedo, southern hemisphere"),
    ("wetIceAlb", 0.66, float, "wet ice albedo"),
    ("wetIceAlb_south", 0.66, float, "wet ice albedo, southern hemisphere"),
    ("drySnowAlb", 0.84, float, "dry snow albedo"),
    ("drySnowAlb_south", 0.84, float, "dry snow albedo, southern hemisphere"),
    ("wetSnowAlb", 0.7, float, "wet snow albedo"),
    ("wetSnowAlb_south", 0.7, float, "wet snow albedo, southern hemisphere"),
    ("wetAlbTemp", 0, float, "temperature above which wet albedos apply [degC]"),
    ("lhFusion", 3.34e5, float, "latent heat of fusion [J/kg]"),
    ("lhEvap", 2.5e6, float, "latent heat of evaporation [J/kg]"),
    ("lhSublim", 3.34e5 + 2.5e6, float, "latent heat of sublimation [J/kg]"),
    ("cpAir", 1005, float, "heat capacity of air [J/kg/K]"),
    ("cpWater", 3986, float, "heat capacity of water [J/kg/K]"),
    ("stefBoltz", 5.67e-8, float, "Stefan-Boltzmann constant [W/m2/K4]"),
    ("iceEmiss", 0.95, float, "ice long-wave emissivity"),
    ("snowEmiss", 0.95, float, "snow long-wave emissivity"),
    ("iceConduct", 2.1656, float, "ice conductivity"),
    ("snowConduct", 0.31, float, "snow conductivity"),
    ("hCut", 0.15, float, "snow thickness of full snow albedo [m]"),
    ("shortwave", 0.3, float, "short-wave penetration factor"),
    ("tempFrz", -1.96, float, "freezing temperature [degC]"),
    ("dtempFrz_dS", 0, float, "d(freezing temperature)/d(salinity)"),
    ("saltIce_ref", 0, float, "sea-ice reference salinity [g/kg]"),
    ("saltOcn_ref", 34.7, float, "ocean reference salinity [g/kg]"),
    ("minLWdown", 60, float, "minimum downward long-wave radiation"),
    ("maxTIce", 30, float, "maximum ice temperature"),
    ("minTIce", -50, float, "minimum ice temperature"),
    ("minTAir", -50, float, "minimum air temperature"),
    ("dalton", 0.00175, float, "Dalton number"),
    ("Area_reg", 0.15**2, float, "regularisation of the ice concentration"),
    ("hIce_reg", 0.10**2, float, "regularisation of the ice thickness"),
    ("celsius2K", 273.15, float, "degC to K offset"),
    ("stantonNr", 0.0056, float, "Stanton number"),
    ("uStarBase", 0.0125, float, "friction velocity under ice [m/s]"),
    ("McPheeTaperFac", 12.5, float, "ice-bottom tapering factor"),
    ("h0", 0.5, float, "lead closing parameter"),
    ("recip_h0", 1 / 0.5, float, "1/h0"),
    ("h0_south", 0.5, float, "lead closing parameter, southern hemisphere"),
    ("recip_h0_south", 1 / 0.5, float, "1/h0_south"),
    ("airTurnAngle", 0, float, "air-ice stress turning angle [deg]"),
    ("waterTurnAngle", 0, float, "water-ice stress turning angle [deg]"),
    ("sinWat", 0, float, "sin of the water turning angle (EVP)"),
    ("cosWat", 1, float, "cos of the water turning angle (EVP)"),
    ("wSpeedMin", 1e-10, float, "minimum wind speed [m/s]"),
    ("hIce_min", 1e-5, float, "minimum ice thickness [m]"),
    ("Area_min", 1e-5, float, "minimum ice cover fraction"),
    ("airIceDrag", 0.0012, float, "air-ice drag coefficient"),
    ("airIceDrag_south", 0.0012, float, "air-ice drag coefficient, south"),
    ("waterIceDrag", 0.0055, float, "water-ice drag coefficient"),
    ("waterIceDrag_south", 0.0055, float, "water-ice drag coefficient, south"),
    ("cDragMin", 0.25, float, "minimum linear ice-ocean drag coefficient"),
    ("seaIceLoadFac", 1, float, "sea-ice load scale factor"),
    ("gravity", 9.81, float, "gravitational acceleration [m/s2]"),
    ("PlasDefCoeff", 2, float, "yield-curve axis ratio"),
    ("deltaMin", 2e-9, float, "minimum Delta"),
    ("pressReplFac", 1, int, "replacement-pressure flag"),
    ("pStar", 27.5e3, float, "ice strength parameter P*"),
    ("cStar", 20, float, "ice strength parameter C*"),
    ("basalDragU0", 5e-5, float, "basal drag velocity scale"),
    ("basalDragK1", 8, float, "basal drag keel factor"),
    ("basalDragK2", 0, float, "basal drag magnitude"),
    ("cBasalStar", 20, float, "basal drag concentration factor"),
    ("tensileStrFac", 0, float, "tensile strength factor"),
    ("CrMax", 1e6, float, "flux limiter slope-ratio cap"),
    ("sideDragCoeff", 0.001, float, "side drag coefficient"),
    ("sideDragU0", 0.01, float, "side drag velocity scale [m/s]"),
    ("umin_o", 0.5, float, "minimum wind speed over ocean [m/s]"),
    ("umin_i", 1.0, float, "minimum wind speed over ice [m/s]"),
    ("zref", 10.0, float, "wind reference height [m]"),
    ("ztref", 2.0, float, "air temperature reference height [m]"),
    ("bolzc", 1.38065e-23, float, "Boltzmann constant [J/K/molecule]"),
    ("avogad", 6.02214e26, float, "Avogadro number [molecules/kmole]"),
    ("rgas", 8314.47, float, "ideal gas constant [J/K/kmole]"),
    ("mwdair", 28.966, float, "molecular weight of dry air"),
    ("mwwv", 18.016, float, "molecular weight of water vapour"),
    ("rdair", 287.042, float, "dry air gas constant [J/K/kg]"),
    ("rwv", 461.505, float, "water vapour gas constant [J/K/kg]"),
    ("zvir", 0.608, float, "rwv/rdair - 1"),
    ("cpdair", 1.00464e3, float, "specific heat of dry air [J/K/kg]"),
    ("cpwv", 1.810e3, float, "specific heat of water vapour [J/K/kg]"),
    ("cpvir", 0.802, float, "cpwv/cpdair - 1"),
    ("karman", 0.4, float, "von Karman constant"),
    ("latvap", 2.501e6, float, "latent heat of evaporation [J/kg]"),
    ("p0", 1e5, float, "reference pressure [Pa]"),
    ("cappa", 0.286, float, "R/cp"),
    ("zzsice", 0.0005, float, "ice surface roughness [m]"),
    ("ch", 1e-3, float, "sensible heat transfer coefficient"),
    ("ce", 1.15e-3, float, "latent heat transfer coefficient"),
    ("eps2", 1e-20, float, "threshold value"),
    ("emissivity", 1, float, "surface emissivity"),
    ("ocean_emissivity", 0.985, float, "ocean surface emissivity"),
    ("snow_emissivity", 0.98, float, "snow surface emissivity"),
    ("ice_emissivity", 0.98, float, "ice surface emissivity"),
    ("tf0kel", 273.15, float, "fresh water freezing temperature [K]"),
    ("gamma_blk", 0.010, float, "adiabatic lapse rate [K/m]"),
    ("ocean_albedo", 0.1, float, "ocean albedo"),
    ("ice_albedo", 0.7, float, "ice albedo"),
]

SETTINGS = {name: _S(d, t, h) for name, d, t, h in _TABLE}

# settings that are reciprocals of another setting
_RECIPROCALS = {
    "deltatTherm": "recip_deltatTherm",
    "deltatDyn": "recip_deltatDyn",
    "nITC": "recip_nITC",
    "rhoFresh": "recip_rhoFresh",
    "rhoSea": "recip_rhoSea",
    "h0": "recip_h0",
    "h0_south": "recip_h0_south",
}


class SettingsNamespace:
    """Attribute namespace of setting values (``state.settings.<name>``);
    ``_version`` counts assignments."""

    def __init__(self, values: dict):
        object.__setattr__(self, "_values", dict(values))
        object.__setattr__(self, "_version", 0)

    def __getattr__(self, name):
        try:
            return self._values[name]
        except KeyError:
            raise AttributeError(name) from None

    def __setattr__(self, name, value):
        if name not in SETTINGS and name not in self._values:
            raise AttributeError(f"unknown setting {name!r}")
        self._values[name] = value
        object.__setattr__(self, "_version", self._version + 1)

    def __getitem__(self, name):
        return self._values[name]

    def __contains__(self, name):
        return name in self._values

    def keys(self):
        return self._values.keys()

    def as_dict(self):
        return dict(self._values)


def make_settings(**overrides) -> SettingsNamespace:
    """Defaults of every setting, with ``overrides`` applied.

    Overriding a setting that has a reciprocal partner (``deltatDyn`` etc.) and
    not the partner itself sets the partner to ``1/value`` so that the two stay
    consistent, since the reference reads both independently.
    """
    vals = {k: s.default for k, s in SETTINGS.items()}
    for k, v in overrides.items():
        if k not in SETTINGS:
            raise KeyError(f"unknown setting {k!r}")
        vals[k] = v
    for base, recip in _RECIPROCALS.items():
        if base in overrides and recip not in overrides:
            vals[recip] = 1 / overrides[base]
    return SettingsNamespace(vals)


def check_consistent(sett) -> None:
    """Raise if a reciprocal setting disagrees with its partner."""
    for base, recip in _RECIPROCALS.items():
        a, b = float(sett[base]), float(sett[recip])
        if not math.isclose(a * b, 1.0, rel_tol=1e-12):
            raise ValueError(f"inconsistent settings: {recip}={b} but {base}={a}")
